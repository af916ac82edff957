// Hand-written adjoint (vector-Jacobian product) of the CDD flux and of one
// TVD Runge-Kutta step, with the time-step schedule frozen from the forward
// run (the CFL rule of FieldMover.py:94-96 is a non-differentiable global
// max; the adjoint treats the recorded dt as a constant, and finite-difference
// checks replay the same schedule).
//
// L(beta) = flux(beta, sigma(beta)) per RK stage (CentralUpwindHJ3D.py:63-271
// with the BetaP velocities/Hamiltonians).  Its transpose applied to a
// cotangent rbar is assembled from three pieces:
//   1. k_flux_vjp_cell: per cell, recompute the forward (one-sided derivatives
//      u, stress, corner velocities, speeds, weights) and propagate rbar back
//      to ubar (the 54 derivative cotangents), to the stress cotangent
//      (symmetrised, 9 planes) and to Lambda.  The max/min speed selections
//      and the minmod branches are replayed from the forward values.
//   2. k_flux_vjp_gather: the transpose of the limited-derivative stencils,
//      written as a gather over limiter sites (each cell collects from the 4
//      sites and 5 cells whose stencils reach it) so no atomics are needed;
//      the minmod selections are the codes stored by the cell kernel.
//   3. the stress adjoint: the spectral passes with the transposed Green's
//      contraction (green_sigma_T) applied to the symmetrised stress
//      cotangent, plus the transpose of the local part
//      -lam tr(beta) d - mu (beta + beta^T).
#include "cdd_kernels.h"

namespace cdd {

template <typename T>
__device__ __forceinline__ T mm3(T a, T b, T c) {
  const T mn = fmin(a, fmin(b, c));
  const T mx = fmax(a, fmax(b, c));
  return fmax(mn, T(0)) + fmin(mx, T(0));
}

// Which argument minmod3(a, b, c) returns: 0, 1, 2, or -1 for the zero case
// (NumericalMethods.py:367-378); ties resolve to the first argument.
template <typename T>
__device__ __forceinline__ int mm3_sel(T a, T b, T c) {
  const T mn = fmin(a, fmin(b, c));
  const T mx = fmax(a, fmax(b, c));
  if (mn > T(0)) return a == mn ? 0 : (b == mn ? 1 : 2);
  if (mx <= T(0)) return a == mx ? 0 : (b == mx ? 1 : 2);
  return -1;
}

template <typename T>
__device__ __forceinline__ void lderivs(const T f[5], T N, T& up, T& um) {
  const T Dm2 = f[1] - f[0], Dm1 = f[2] - f[1], D0 = f[3] - f[2], D1 = f[4] - f[3];
  const T m_i = mm3(D1 - D0, (D1 - Dm1) * T(0.5), D0 - Dm1);
  const T m_im1 = mm3(D0 - Dm1, (D0 - Dm2) * T(0.5), Dm1 - Dm2);
  up = N * fma(T(-0.5), m_i, D0);
  um = N * fma(T(0.5), m_im1, Dm1);
}

// ------------------------------------------------------------------ R1
template <typename T, int D, int DYN>
__host__ __device__ void cell_flux_vjp(const T (&u)[D][2][9], const T (&sg)[9], T Lambda, const T (&rb)[9],
                              T (&ub)[D][2][9], T (&sgb)[9], T& lamb) {
  constexpr int NCOR = 1 << D;
  const T ME = T(kME);
  const bool climb_sup = (DYN == DYN_GLIDE_CLIMB) && (Lambda != T(0));
  const T third = T(1.0 / 3.0);
  const T st = climb_sup ? Lambda * (sg[0] + sg[4] + sg[8]) * third : T(0);
  const T jfac = -Lambda * third;
  // ---------------- forward recompute
  T w[D][2][3], zz[D][2], tr[D][2], P[D][2];
  for (int a = 0; a < D; ++a)
    for (int s = 0; s < 2; ++s) {
      const T* ua = u[a][s];
      for (int l = 0; l < 3; ++l)
        w[a][s][l] = ua[3 * l] * sg[3 * a] + ua[3 * l + 1] * sg[3 * a + 1] + ua[3 * l + 2] * sg[3 * a + 2];
      T z = T(0);
      for (int c = 0; c < 9; ++c) z += ua[c] * sg[c];
      zz[a][s] = z;
      tr[a][s] = ua[0] + ua[4] + ua[8];
      T p = T(0);
      for (int m = 0; m < 3; ++m)
        if (m != a)
          for (int jj = 0; jj < 3; ++jj) p += ua[3 * m + jj] * ua[3 * m + jj];
      P[a][s] = p;
    }
  T R[NCOR][3], q[NCOR][3], v0p[NCOR][3], v0[NCOR][3], v[NCOR][3], rinv[NCOR], J[NCOR];
  T rr[NCOR], rv[NCOR], fpr[NCOR];
  for (int c = 0; c < NCOR; ++c) {
    int sg_[3] = {0, 0, 0};
    for (int a = 0; a < D; ++a) sg_[a] = (c >> (D - 1 - a)) & 1;
    for (int l = 0; l < 3; ++l) {
      T r = T(0), vv = T(0);
      for (int a = 0; a < D; ++a) {
        r += u[a][sg_[a]][3 * l + a];
        vv += w[a][sg_[a]][l];
      }
      R[c][l] = r;
      v0p[c][l] = vv;
    }
    for (int a = 0; a < D; ++a) v0p[c][a] -= zz[a][sg_[a]];
    for (int a = 0; a < 3; ++a) q[c][a] = (a < D ? tr[a % D][sg_[a]] : T(0)) - R[c][a];
    for (int l = 0; l < 3; ++l) v0[c][l] = v0p[c][l];
    rr[c] = rv[c] = fpr[c] = T(0);
    if (DYN == DYN_GLIDE_CLIMB) {
      if (climb_sup) {
        for (int a = 0; a < D; ++a) v0[c][a] += st * tr[a][sg_[a]];
        for (int l = 0; l < 3; ++l) v0[c][l] -= st * R[c][l];
      }
    } else {
      T rrr = ME;
      for (int a = 0; a < D; ++a) rrr += tr[a][sg_[a]] * tr[a][sg_[a]] - T(2) * tr[a][sg_[a]] * R[c][a];
      for (int l = 0; l < 3; ++l) rrr += R[c][l] * R[c][l];
      const T rvv = ME + v0p[c][0] * q[c][0] + v0p[c][1] * q[c][1] + v0p[c][2] * q[c][2];
      rr[c] = rrr;
      rv[c] = rvv;
      fpr[c] = rvv / rrr;
      for (int a = 0; a < 3; ++a) v0[c][a] += q[c][a] * fpr[c];
    }
    T r2 = ME;
    for (int a = 0; a < D; ++a) r2 += P[a][sg_[a]];
    T cross = T(0);
    for (int jj = 0; jj < 3; ++jj) {
      cross += u[0][sg_[0]][3 + jj] * u[1][sg_[1]][jj];
      if (D == 3) {
        cross += u[1][sg_[1]][6 + jj] * u[D - 1][sg_[D - 1]][3 + jj];
        cross += u[D - 1][sg_[D - 1]][jj] * u[0][sg_[0]][6 + jj];
      }
    }
    r2 -= T(2) * cross;
    rinv[c] = rsqrt_nr(r2);
    for (int l = 0; l < 3; ++l) v[c][l] = v0[c][l] * rinv[c];
    J[c] = climb_sup ? jfac * (v[c][0] * q[c][0] + v[c][1] * q[c][1] + v[c][2] * q[c][2]) : T(0);
  }
  T ap[3] = {0, 0, 0}, am[3] = {0, 0, 0}, at[3] = {1, 1, 1};
  int cmax[3] = {0, 0, 0}, cmin[3] = {0, 0, 0};
  for (int a = 0; a < D; ++a) {
    T mxv = v[0][a], mnv = v[0][a];
    for (int c = 1; c < NCOR; ++c) {
      if (v[c][a] > mxv) {
        mxv = v[c][a];
        cmax[a] = c;
      }
      if (v[c][a] < mnv) {
        mnv = v[c][a];
        cmin[a] = c;
      }
    }
    ap[a] = fmax(mxv, T(0));
    am[a] = -fmin(mnv, T(0));
    at[a] = ap[a] + am[a] + ME;
  }
  T Y[D][2][3], Jsum = T(0), wcs[NCOR];
  for (int a = 0; a < D; ++a)
    for (int s = 0; s < 2; ++s)
      for (int l = 0; l < 3; ++l) Y[a][s][l] = T(0);
  for (int c = 0; c < NCOR; ++c) {
    T wc = T(1);
    for (int a = 0; a < D; ++a) wc *= ((c >> (D - 1 - a)) & 1) == 0 ? am[a] : ap[a];
    wcs[c] = wc;
    for (int a = 0; a < D; ++a)
      for (int l = 0; l < 3; ++l) Y[a][(c >> (D - 1 - a)) & 1][l] += wc * v[c][l];
    Jsum += wc * J[c];
  }
  T inv = T(1);
  for (int a = 0; a < D; ++a) inv *= at[a];
  inv = T(1) / inv;
  T twv[3] = {0, 0, 0};
  T iat[3] = {1, 1, 1};
  for (int a = 0; a < D; ++a) {
    T o = inv;
    for (int b = 0; b < D; ++b)
      if (b != a) o *= at[b];
    iat[a] = o;
  }
  for (int a = 0; a < D; ++a) twv[a] = am[a] * ap[a] * iat[a];
  T Hs[9];
  for (int i = 0; i < 3; ++i)
    for (int jj = 0; jj < 3; ++jj) {
      const int cc = 3 * i + jj;
      T h = (i == jj) ? Jsum : T(0);
      for (int a = 0; a < D; ++a)
        for (int s = 0; s < 2; ++s) h += u[a][s][cc] * Y[a][s][a];
      if (i < D)
        for (int s = 0; s < 2; ++s)
          for (int l = 0; l < 3; ++l) h -= u[i % D][s][3 * l + jj] * Y[i % D][s][l];
      Hs[cc] = h;
    }
  // ---------------- reverse
  for (int a = 0; a < D; ++a)
    for (int s = 0; s < 2; ++s)
      for (int c = 0; c < 9; ++c) ub[a][s][c] = T(0);
  for (int c = 0; c < 9; ++c) sgb[c] = T(0);
  lamb = T(0);
  T Hb[9], invb = T(0), twb[3] = {0, 0, 0};
  for (int cc = 0; cc < 9; ++cc) {
    Hb[cc] = -rb[cc] * inv;
    invb -= rb[cc] * Hs[cc];
    for (int a = 0; a < D; ++a) {
      twb[a] += rb[cc] * (u[a][0][cc] - u[a][1][cc]);
      ub[a][0][cc] += rb[cc] * twv[a];
      ub[a][1][cc] -= rb[cc] * twv[a];
    }
  }
  const T Jsumb = Hb[0] + Hb[4] + Hb[8];
  T Yb[D][2][3];
  for (int a = 0; a < D; ++a)
    for (int s = 0; s < 2; ++s)
      for (int l = 0; l < 3; ++l) Yb[a][s][l] = T(0);
  for (int i = 0; i < 3; ++i)
    for (int jj = 0; jj < 3; ++jj) {
      const int cc = 3 * i + jj;
      for (int a = 0; a < D; ++a)
        for (int s = 0; s < 2; ++s) {
          ub[a][s][cc] += Hb[cc] * Y[a][s][a];
          Yb[a][s][a] += Hb[cc] * u[a][s][cc];
        }
      if (i < D)
        for (int s = 0; s < 2; ++s)
          for (int l = 0; l < 3; ++l) {
            ub[i % D][s][3 * l + jj] -= Hb[cc] * Y[i % D][s][l];
            Yb[i % D][s][l] -= Hb[cc] * u[i % D][s][3 * l + jj];
          }
    }
  T apb[3] = {0, 0, 0}, amb[3] = {0, 0, 0};
  for (int a = 0; a < D; ++a) {
    amb[a] += twb[a] * ap[a] * iat[a];
    apb[a] += twb[a] * am[a] * iat[a];
    T atb = -twb[a] * am[a] * ap[a] * iat[a] * iat[a] - invb * inv * iat[a];
    apb[a] += atb;
    amb[a] += atb;
  }
  T vb[NCOR][3], Jb[NCOR];
  for (int c = 0; c < NCOR; ++c) {
    T wcb = Jsumb * J[c];
    Jb[c] = Jsumb * wcs[c];
    for (int l = 0; l < 3; ++l) {
      T acc = T(0);
      for (int a = 0; a < D; ++a) {
        const T yb = Yb[a][(c >> (D - 1 - a)) & 1][l];
        acc += yb;
        wcb += yb * v[c][l];
      }
      vb[c][l] = acc * wcs[c];
    }
    for (int a = 0; a < D; ++a) {
      T others = T(1);
      for (int b2 = 0; b2 < D; ++b2)
        if (b2 != a) others *= ((c >> (D - 1 - b2)) & 1) == 0 ? am[b2] : ap[b2];
      if (((c >> (D - 1 - a)) & 1) == 0) amb[a] += wcb * others;
      else apb[a] += wcb * others;
    }
  }
  for (int a = 0; a < D; ++a) {
    if (ap[a] > T(0)) vb[cmax[a]][a] += apb[a];
    if (am[a] > T(0)) vb[cmin[a]][a] -= amb[a];
  }
  T Pb[D][2], wb[D][2][3], zzb[D][2], trb[D][2];
  for (int a = 0; a < D; ++a)
    for (int s = 0; s < 2; ++s) {
      Pb[a][s] = zzb[a][s] = trb[a][s] = T(0);
      for (int l = 0; l < 3; ++l) wb[a][s][l] = T(0);
    }
  T stb = T(0);
  for (int c = 0; c < NCOR; ++c) {
    int sg_[3] = {0, 0, 0};
    for (int a = 0; a < D; ++a) sg_[a] = (c >> (D - 1 - a)) & 1;
    T qb[3] = {0, 0, 0}, Rb[3] = {0, 0, 0};
    if (climb_sup) {
      const T vq = v[c][0] * q[c][0] + v[c][1] * q[c][1] + v[c][2] * q[c][2];
      for (int l = 0; l < 3; ++l) {
        vb[c][l] += Jb[c] * jfac * q[c][l];
        qb[l] += Jb[c] * jfac * v[c][l];
      }
      lamb -= Jb[c] * vq * third;
    }
    T v0b[3], rinvb = T(0);
    for (int l = 0; l < 3; ++l) {
      v0b[l] = vb[c][l] * rinv[c];
      rinvb += vb[c][l] * v0[c][l];
    }
    const T r2b = rinvb * T(-0.5) * rinv[c] * rinv[c] * rinv[c];
    for (int a = 0; a < D; ++a) Pb[a][sg_[a]] += r2b;
    const T crossb = T(-2) * r2b;
    for (int jj = 0; jj < 3; ++jj) {
      ub[0][sg_[0]][3 + jj] += crossb * u[1][sg_[1]][jj];
      ub[1][sg_[1]][jj] += crossb * u[0][sg_[0]][3 + jj];
      if (D == 3) {
        ub[1][sg_[1]][6 + jj] += crossb * u[D - 1][sg_[D - 1]][3 + jj];
        ub[D - 1][sg_[D - 1]][3 + jj] += crossb * u[1][sg_[1]][6 + jj];
        ub[D - 1][sg_[D - 1]][jj] += crossb * u[0][sg_[0]][6 + jj];
        ub[0][sg_[0]][6 + jj] += crossb * u[D - 1][sg_[D - 1]][jj];
      }
    }
    T v0pb[3] = {v0b[0], v0b[1], v0b[2]};
    if (DYN == DYN_GLIDE_CLIMB) {
      if (climb_sup) {
        for (int a = 0; a < D; ++a) {
          stb += v0b[a] * tr[a][sg_[a]];
          trb[a][sg_[a]] += v0b[a] * st;
        }
        for (int l = 0; l < 3; ++l) {
          stb -= v0b[l] * R[c][l];
          Rb[l] -= v0b[l] * st;
        }
      }
    } else {
      T fb = T(0);
      for (int a = 0; a < 3; ++a) {
        fb += v0b[a] * q[c][a];
        qb[a] += v0b[a] * fpr[c];
      }
      const T rvb = fb / rr[c];
      const T rrb = -fb * rv[c] / (rr[c] * rr[c]);
      for (int a = 0; a < 3; ++a) {
        v0pb[a] += rvb * q[c][a];
        qb[a] += rvb * v0p[c][a];
      }
      for (int a = 0; a < D; ++a) {
        trb[a][sg_[a]] += rrb * (T(2) * tr[a][sg_[a]] - T(2) * R[c][a]);
        Rb[a] += rrb * (T(-2) * tr[a][sg_[a]]);
      }
      for (int l = 0; l < 3; ++l) Rb[l] += rrb * T(2) * R[c][l];
    }
    for (int a = 0; a < 3; ++a) {
      if (a < D) trb[a % D][sg_[a]] += qb[a];
      Rb[a] -= qb[a];
    }
    for (int l = 0; l < 3; ++l)
      for (int a = 0; a < D; ++a) wb[a][sg_[a]][l] += v0pb[l];
    for (int a = 0; a < D; ++a) zzb[a][sg_[a]] -= v0pb[a];
    for (int l = 0; l < 3; ++l)
      for (int a = 0; a < D; ++a) ub[a][sg_[a]][3 * l + a] += Rb[l];
  }
  if (climb_sup) {
    sgb[0] += stb * Lambda * third;
    sgb[4] += stb * Lambda * third;
    sgb[8] += stb * Lambda * third;
    lamb += stb * (sg[0] + sg[4] + sg[8]) * third;
  }
  for (int a = 0; a < D; ++a)
    for (int s = 0; s < 2; ++s) {
      for (int m = 0; m < 3; ++m)
        if (m != a)
          for (int jj = 0; jj < 3; ++jj) ub[a][s][3 * m + jj] += T(2) * Pb[a][s] * u[a][s][3 * m + jj];
      for (int l = 0; l < 3; ++l)
        for (int nn = 0; nn < 3; ++nn) {
          ub[a][s][3 * l + nn] += wb[a][s][l] * sg[3 * a + nn];
          sgb[3 * a + nn] += wb[a][s][l] * u[a][s][3 * l + nn];
        }
      for (int c = 0; c < 9; ++c) {
        ub[a][s][c] += zzb[a][s] * sg[c];
        sgb[c] += zzb[a][s] * u[a][s][c];
      }
      ub[a][s][0] += trb[a][s];
      ub[a][s][4] += trb[a][s];
      ub[a][s][8] += trb[a][s];
    }
}

// The reverse of cell_flux with a small working set, for the 3D kernel
// (cell_flux_vjp holds every corner's quantities and all 54 derivative
// cotangents at once, which spills in 3D).  Every corner quantity depends on
// the derivatives only through per-(axis, sign) sums F: the diagonal entries
// u_a[l, a], the stress contractions w_a[l] = u_a[l, n] sigma[a, n] and
// z_a = u_a : sigma, the traces, the |rho|^2 pieces P_a and the |rho| cross
// products of two axes.  The forward pass forms F, parks each corner's
// velocity and climb rate in vbuf (4 values per corner, stride vstride, as
// the forward flux kernel does) and Y, Hb in ybuf; the reverse corner sweeps
// read only F and accumulate the cotangents of F; then reload(u) refills the
// derivatives (dead during the sweeps) and each derivative cotangent
// ub[a][s][c] is formed once and handed to store(a, s, c, value).  Same
// arithmetic as cell_flux_vjp up to the order of the sums.
template <typename T, int D, int DYN, typename Store, typename Reload>
__host__ __device__ __forceinline__ void cell_flux_vjp_lean(T (&u)[D][2][9], const T (&sg)[9], T Lambda,
                                                            const T (&rb)[9], T* __restrict__ vbuf,
                                                            T* __restrict__ ybuf, int vstride, T (&sgb)[9], T& lamb,
                                                            Store store, Reload reload) {
  constexpr int NCOR = 1 << D;
  const T ME = T(kME);
  const bool climb_sup = (DYN == DYN_GLIDE_CLIMB) && (Lambda != T(0));
  const T third = T(1.0 / 3.0);
  const T st = climb_sup ? Lambda * (sg[0] + sg[4] + sg[8]) * third : T(0);
  const T jfac = -Lambda * third;
  auto sgn = [](int c, int a) { return (c >> (D - 1 - a)) & 1; };
  // ---------------- F
  T Ru[D][2][3], w[D][2][3], zz[D][2], tr[D][2], P[D][2], C01[2][2], C12[2][2], C20[2][2];
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const T* ua = u[a][s];
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        Ru[a][s][l] = ua[3 * l + a];
        w[a][s][l] = ua[3 * l] * sg[3 * a] + ua[3 * l + 1] * sg[3 * a + 1] + ua[3 * l + 2] * sg[3 * a + 2];
      }
      T z = T(0);
#pragma unroll
      for (int cc = 0; cc < 9; ++cc) z += ua[cc] * sg[cc];
      zz[a][s] = z;
      tr[a][s] = ua[0] + ua[4] + ua[8];
      T p = T(0);
#pragma unroll
      for (int m = 0; m < 3; ++m)
        if (m != a)
#pragma unroll
          for (int jj = 0; jj < 3; ++jj) p += ua[3 * m + jj] * ua[3 * m + jj];
      P[a][s] = p;
    }
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      T c01 = T(0), c12 = T(0), c20 = T(0);
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) {
        c01 += u[0][s][3 + jj] * u[1][t][jj];
        if (D == 3) {
          c12 += u[1][s][6 + jj] * u[D - 1][t][3 + jj];
          c20 += u[D - 1][s][jj] * u[0][t][6 + jj];
        }
      }
      C01[s][t] = c01;
      C12[s][t] = c12;
      C20[s][t] = c20;
    }
  // forward quantities of corner c (cell_flux_vjp's R, q, v0p, v0, rr, rv, fpr, rinv) from F
  auto corner = [&](int c, T (&R)[3], T (&q)[3], T (&v0p)[3], T (&v0)[3], T& rr, T& rv, T& fpr, T& rinv) {
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      T r = T(0), vv = T(0);
#pragma unroll
      for (int a = 0; a < D; ++a) {
        r += Ru[a][sgn(c, a)][l];
        vv += w[a][sgn(c, a)][l];
      }
      R[l] = r;
      v0p[l] = vv;
    }
#pragma unroll
    for (int a = 0; a < D; ++a) v0p[a] -= zz[a][sgn(c, a)];
#pragma unroll
    for (int a = 0; a < 3; ++a) q[a] = (a < D ? tr[a % D][sgn(c, a % D)] : T(0)) - R[a];
#pragma unroll
    for (int l = 0; l < 3; ++l) v0[l] = v0p[l];
    rr = rv = fpr = T(0);
    if (DYN == DYN_GLIDE_CLIMB) {
      if (climb_sup) {
#pragma unroll
        for (int a = 0; a < D; ++a) v0[a] += st * tr[a][sgn(c, a)];
#pragma unroll
        for (int l = 0; l < 3; ++l) v0[l] -= st * R[l];
      }
    } else {
      T rrr = ME;
#pragma unroll
      for (int a = 0; a < D; ++a) {
        const T ta = tr[a][sgn(c, a)];
        rrr += ta * ta - T(2) * ta * R[a];
      }
#pragma unroll
      for (int l = 0; l < 3; ++l) rrr += R[l] * R[l];
      rr = rrr;
      rv = ME + v0p[0] * q[0] + v0p[1] * q[1] + v0p[2] * q[2];
      fpr = rv / rr;
#pragma unroll
      for (int a = 0; a < 3; ++a) v0[a] += q[a] * fpr;
    }
    T r2 = ME;
#pragma unroll
    for (int a = 0; a < D; ++a) r2 += P[a][sgn(c, a)];
    T cross = C01[sgn(c, 0)][sgn(c, 1)];
    if (D == 3) cross += C12[sgn(c, 1)][sgn(c, D - 1)] + C20[sgn(c, D - 1)][sgn(c, 0)];
    r2 -= T(2) * cross;
    rinv = rsqrt_nr(r2);
  };
  // the climb term's corner rate J = -Lambda/3 v . q, q from F
  auto climb_rate = [&](int c) {
    T j = T(0);
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      T r = T(0);
#pragma unroll
      for (int a = 0; a < D; ++a) r += Ru[a][sgn(c, a)][l];
      const T ql = (l < D ? tr[l % D][sgn(c, l % D)] : T(0)) - r;
      j += vbuf[(3 * c + l) * vstride] * ql;
    }
    return jfac * j;
  };
  auto weight = [&](int c, const T (&ap)[3], const T (&am)[3]) {
    T wc = T(1);
#pragma unroll
    for (int a = 0; a < D; ++a) wc *= sgn(c, a) == 0 ? am[a] : ap[a];
    return wc;
  };
  // ---------------- forward: corner velocities and climb rates, speeds
  T mxv[3], mnv[3];
  int cmax[3] = {0, 0, 0}, cmin[3] = {0, 0, 0};
#pragma unroll
  for (int c = 0; c < NCOR; ++c) {
    T R[3], q[3], v0p[3], v0[3], rr, rv, fpr, rinv;
    corner(c, R, q, v0p, v0, rr, rv, fpr, rinv);
    T v[3];
#pragma unroll
    for (int l = 0; l < 3; ++l) v[l] = v0[l] * rinv;
#pragma unroll
    for (int l = 0; l < 3; ++l) vbuf[(3 * c + l) * vstride] = v[l];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      if (c == 0) {
        mxv[a] = mnv[a] = v[a];
      } else {
        if (v[a] > mxv[a]) {
          mxv[a] = v[a];
          cmax[a] = c;
        }
        if (v[a] < mnv[a]) {
          mnv[a] = v[a];
          cmin[a] = c;
        }
      }
    }
  }
  T ap[3] = {0, 0, 0}, am[3] = {0, 0, 0}, at[3] = {1, 1, 1};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    ap[a] = fmax(mxv[a], T(0));
    am[a] = -fmin(mnv[a], T(0));
    at[a] = ap[a] + am[a] + ME;
  }
  // Y and Hb wait in ybuf (stride vstride) for the final pass: Y at
  // 6 a + 3 s + l, Hb at 6 D + c
  auto yat = [&](int a, int s, int l) -> T& { return ybuf[(6 * a + 3 * s + l) * vstride]; };
  auto hbat = [&](int c) -> T& { return ybuf[(6 * D + c) * vstride]; };
  auto sgat = [&](int c) -> T& { return ybuf[(6 * D + 9 + c) * vstride]; };
#pragma unroll
  for (int c = 0; c < 9; ++c) sgat(c) = sg[c];
  T invb = T(0);
  {
    T Y[D][2][3];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int l = 0; l < 3; ++l) Y[a][s][l] = T(0);
    T Jsum = T(0);
#pragma unroll
    for (int c = 0; c < NCOR; ++c) {
      const T wc = weight(c, ap, am);
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        const T vl = vbuf[(3 * c + l) * vstride];
#pragma unroll
        for (int a = 0; a < D; ++a) Y[a][sgn(c, a)][l] += wc * vl;
      }
      if (climb_sup) Jsum += wc * climb_rate(c);
    }
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int l = 0; l < 3; ++l) yat(a, s, l) = Y[a][s][l];
    T inv0 = T(1);
#pragma unroll
    for (int a = 0; a < D; ++a) inv0 *= at[a];
    inv0 = T(1) / inv0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) {
        const int cc = 3 * i + jj;
        T h = (i == jj) ? Jsum : T(0);
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
          for (int s = 0; s < 2; ++s) h += u[a][s][cc] * Y[a][s][a];
        if (i < D)
#pragma unroll
          for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int l = 0; l < 3; ++l) h -= u[i % D][s][3 * l + jj] * Y[i % D][s][l];
        // H is needed only for the cotangent of 1 / prod(at) and, through
        // Hb = -rbar / prod(at), for everything else
        invb -= rb[cc] * h;
        hbat(cc) = -rb[cc] * inv0;
      }
  }
  T inv = T(1);
#pragma unroll
  for (int a = 0; a < D; ++a) inv *= at[a];
  inv = T(1) / inv;
  // tw2[a] = am ap / at_a times prod(at): the diffusion weight over the
  // factor of Hb = -rbar / prod(at), so that the final pass takes rbar from Hb
  T iat[3] = {1, 1, 1}, tw2[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    T o = inv, p = T(1);
#pragma unroll
    for (int b = 0; b < D; ++b)
      if (b != a) {
        o *= at[b];
        p *= at[b];
      }
    iat[a] = o;
    tw2[a] = am[a] * ap[a] * p;
  }
  // ---------------- reverse
  T twb[3] = {0, 0, 0};
#pragma unroll
  for (int cc = 0; cc < 9; ++cc)
#pragma unroll
    for (int a = 0; a < D; ++a) twb[a] += rb[cc] * (u[a][0][cc] - u[a][1][cc]);
  const T Jsumb = hbat(0) + hbat(4) + hbat(8);
  // Yb[a][s][l] = [l == a] sum_c Hb[c] u[a][s][c] - sum_j Hb[3a + j] u[a][s][3l + j]
  T Yb[D][2][3];
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        T y = T(0);
        if (l == a)
#pragma unroll
          for (int cc = 0; cc < 9; ++cc) y += hbat(cc) * u[a][s][cc];
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) y -= hbat(3 * a + jj) * u[a][s][3 * l + jj];
        Yb[a][s][l] = y;
      }
  T apb[3] = {0, 0, 0}, amb[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    amb[a] = twb[a] * ap[a] * iat[a];
    apb[a] = twb[a] * am[a] * iat[a];
    const T atb = -twb[a] * am[a] * ap[a] * iat[a] * iat[a] - invb * inv * iat[a];
    apb[a] += atb;
    amb[a] += atb;
  }
  // corner sweep 1: the weights' cotangents
#pragma unroll
  for (int c = 0; c < NCOR; ++c) {
    T wcb = climb_sup ? Jsumb * climb_rate(c) : T(0);
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const T vl = vbuf[(3 * c + l) * vstride];
#pragma unroll
      for (int a = 0; a < D; ++a) wcb += Yb[a][sgn(c, a)][l] * vl;
    }
#pragma unroll
    for (int a = 0; a < D; ++a) {
      T others = T(1);
#pragma unroll
      for (int b2 = 0; b2 < D; ++b2)
        if (b2 != a) others *= sgn(c, b2) == 0 ? am[b2] : ap[b2];
      if (sgn(c, a) == 0) amb[a] += wcb * others;
      else apb[a] += wcb * others;
    }
  }
  // corner sweep 2: velocity cotangents back to the cotangents of F
  T Pb[D][2], wb[D][2][3], zzb[D][2], trb[D][2], RbA[D][2][3], X01[2][2], X12[2][2], X20[2][2];
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      Pb[a][s] = zzb[a][s] = trb[a][s] = T(0);
#pragma unroll
      for (int l = 0; l < 3; ++l) wb[a][s][l] = RbA[a][s][l] = T(0);
    }
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int t = 0; t < 2; ++t) X01[s][t] = X12[s][t] = X20[s][t] = T(0);
  T stb = T(0);
  lamb = T(0);
#pragma unroll
  for (int c = 0; c < NCOR; ++c) {
    T R[3], q[3], v0p[3], v0[3], rr, rv, fpr, rinv;
    corner(c, R, q, v0p, v0, rr, rv, fpr, rinv);
    const T wc = weight(c, ap, am);
    T vb[3];
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      T acc = T(0);
#pragma unroll
      for (int a = 0; a < D; ++a) acc += Yb[a][sgn(c, a)][l];
      vb[l] = acc * wc;
    }
#pragma unroll
    for (int a = 0; a < D; ++a) {
      if (c == cmax[a] && ap[a] > T(0)) vb[a] += apb[a];
      if (c == cmin[a] && am[a] > T(0)) vb[a] -= amb[a];
    }
    T qb[3] = {0, 0, 0}, Rb[3] = {0, 0, 0};
    if (climb_sup) {
      const T Jb = Jsumb * wc;
      T v[3];
#pragma unroll
      for (int l = 0; l < 3; ++l) v[l] = vbuf[(3 * c + l) * vstride];
      const T vq = v[0] * q[0] + v[1] * q[1] + v[2] * q[2];
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        vb[l] += Jb * jfac * q[l];
        qb[l] += Jb * jfac * v[l];
      }
      lamb -= Jb * vq * third;
    }
    T v0b[3], rinvb = T(0);
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      v0b[l] = vb[l] * rinv;
      rinvb += vb[l] * v0[l];
    }
    const T r2b = rinvb * T(-0.5) * rinv * rinv * rinv;
#pragma unroll
    for (int a = 0; a < D; ++a) Pb[a][sgn(c, a)] += r2b;
    const T crossb = T(-2) * r2b;
    X01[sgn(c, 0)][sgn(c, 1)] += crossb;
    if (D == 3) {
      X12[sgn(c, 1)][sgn(c, D - 1)] += crossb;
      X20[sgn(c, D - 1)][sgn(c, 0)] += crossb;
    }
    T v0pb[3] = {v0b[0], v0b[1], v0b[2]};
    if (DYN == DYN_GLIDE_CLIMB) {
      if (climb_sup) {
#pragma unroll
        for (int a = 0; a < D; ++a) {
          stb += v0b[a] * tr[a][sgn(c, a)];
          trb[a][sgn(c, a)] += v0b[a] * st;
        }
#pragma unroll
        for (int l = 0; l < 3; ++l) {
          stb -= v0b[l] * R[l];
          Rb[l] -= v0b[l] * st;
        }
      }
    } else {
      T fb = T(0);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        fb += v0b[a] * q[a];
        qb[a] += v0b[a] * fpr;
      }
      const T rvb = fb / rr;
      const T rrb = -fb * rv / (rr * rr);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        v0pb[a] += rvb * q[a];
        qb[a] += rvb * v0p[a];
      }
#pragma unroll
      for (int a = 0; a < D; ++a) {
        const T ta = tr[a][sgn(c, a)];
        trb[a][sgn(c, a)] += rrb * (T(2) * ta - T(2) * R[a]);
        Rb[a] += rrb * (T(-2) * ta);
      }
#pragma unroll
      for (int l = 0; l < 3; ++l) Rb[l] += rrb * T(2) * R[l];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (a < D) trb[a % D][sgn(c, a % D)] += qb[a];
      Rb[a] -= qb[a];
    }
#pragma unroll
    for (int a = 0; a < D; ++a) {
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        wb[a][sgn(c, a)][l] += v0pb[l];
        RbA[a][sgn(c, a)][l] += Rb[l];
      }
      zzb[a][sgn(c, a)] -= v0pb[a];
    }
  }
#pragma unroll
  for (int c = 0; c < 9; ++c) sgb[c] = T(0);
  if (climb_sup) {
    sgb[0] += stb * Lambda * third;
    sgb[4] += stb * Lambda * third;
    sgb[8] += stb * Lambda * third;
    lamb += stb * st / Lambda;
  }
  // ---------------- every derivative cotangent, once
  reload(u);
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int s = 0; s < 2; ++s) {
#pragma unroll
      for (int c = 0; c < 9; ++c) {
        const int m = c / 3, nn = c % 3;
        T val = (s == 0 ? -hbat(c) : hbat(c)) * tw2[a];  // +-rbar am ap / at_a
        val += hbat(c) * yat(a, s, a) - hbat(3 * a + nn) * yat(a, s, m);
        if (m != a) val += T(2) * Pb[a][s] * u[a][s][c];
        val += wb[a][s][m] * sgat(3 * a + nn) + zzb[a][s] * sgat(c);
        if (m == nn) val += trb[a][s];
        if (nn == a) val += RbA[a][s][m];
        // the |rho| cross terms, summed over the other axis' sign
        if (a == 0 && m == 1) val += X01[s][0] * u[1][0][nn] + X01[s][1] * u[1][1][nn];
        if (a == 1 && m == 0) val += X01[0][s] * u[0][0][3 + nn] + X01[1][s] * u[0][1][3 + nn];
        if (D == 3) {
          if (a == 0 && m == 2) val += X20[0][s] * u[D - 1][0][nn] + X20[1][s] * u[D - 1][1][nn];
          if (a == 1 && m == 2) val += X12[s][0] * u[D - 1][0][3 + nn] + X12[s][1] * u[D - 1][1][3 + nn];
          if (a == D - 1 && m == 1) val += X12[0][s] * u[1][0][6 + nn] + X12[1][s] * u[1][1][6 + nn];
          if (a == D - 1 && m == 0) val += X20[s][0] * u[0][0][6 + nn] + X20[s][1] * u[0][1][6 + nn];
        }
        store(a, s, c, val);
        sgb[3 * a + nn] += wb[a][s][m] * u[a][s][c];
        sgb[c] += zzb[a][s] * u[a][s][c];
      }
    }
}

static __device__ __forceinline__ int wrapi(int p, int n) { return p < 0 ? p + n : (p >= n ? p - n : p); }

// Tile mapping shared by the two flux-adjoint kernels: blocks of TX x TY
// threads (TX = min(32, n) cells along the contiguous axis, TY = min(8, n)
// along the next one), one plane per block in 3D.  Hardware block b runs on
// XCD b % 8; logical tile L = (b % 8)(nb / 8) + b / 8 hands every XCD one
// contiguous range of tiles, so the stencil neighbours of a tile (the tiles
// above/below it and, in 3D, the neighbouring planes) are mostly in flight
// on the same XCD and served from its L2.
// ROWS (3D): logical tiles ordered (row tile, plane, column tile), column
// tile fastest: a row of tiles along the contiguous axis and its
// neighbouring planes are in flight together, so the halo columns and the
// slow axis's neighbours come from the XCD's L2 (the cell VJP's and the
// gather's order; the gather read 1.4x its bytes in the (column, plane)
// order); the default orders them (column, plane), planes fastest.
template <int D, bool ROWS = false>
__device__ __forceinline__ long long adj_tile_cell(int n, int co[3], long long str[3]) {
  const int TX = blockDim.x, TY = blockDim.y;
  const int nct = n / TX, nrt = n / TY;
  const int nb = (D == 3 ? n : 1) * nct * nrt;
  const int bl = blockIdx.x;
  const int L = (nb % 8 == 0) ? (bl % 8) * (nb / 8) + bl / 8 : bl;
  int px, tyz;
  if (D == 3 && ROWS) {
    const int rt = L / (n * nct), rem = L % (n * nct);
    px = rem / nct;
    tyz = rt * nct + rem % nct;
  } else {
    px = (D == 3) ? L % n : 0;
    tyz = (D == 3) ? L / n : L;
  }
  const int col = (tyz % nct) * TX + threadIdx.x, row = (tyz / nct) * TY + threadIdx.y;
  if (D == 3) {
    co[0] = px;
    co[1] = row;
    co[2] = col;
    str[0] = (long long)n * n;
    str[1] = n;
    str[2] = 1;
    return ((long long)px * n + row) * n + col;
  }
  co[0] = row;
  co[1] = col;
  co[2] = 0;
  str[0] = n;
  str[1] = 1;
  str[2] = 0;
  return (long long)row * n + col;
}

// Block = TX x TY cells of the two fastest axes (one plane in 3D), the tile
// order of adj_tile_cell.  The plane tile of all nine components with its
// 2-cell halo is staged in LDS with one batch of coalesced loads; in 3D each
// thread's four slow-axis neighbours come straight from global memory.
template <typename T, int D, int DYN, int TX, int TY>
__global__ void __launch_bounds__(TX* TY) k_flux_vjp_cell(VjpJob job) {
  constexpr int PW = TX + 4, TILE = PW * (TY + 4), NT = TX * TY, FILL = (TILE + NT - 1) / NT;
  // 3D: the lean reverse (cell_flux_vjp_lean) with its corner velocities,
  // Y, Hb and sigma in LDS (the plain form spills in 3D; in 2D it does not
  // and measures slightly faster)
  constexpr bool LEAN = (D == 3);
  // 3D grids stream the planes touched once (G, rbar, the cotangents out);
  // 2D ones (512^2 at most here) keep them in the infinity cache
  constexpr bool STRM = (D == 3);
  __shared__ T red[4];
  __shared__ T tile[9][TILE];
  __shared__ T vsh[LEAN ? 3 * (1 << D) : 1][LEAN ? NT : 1];
  __shared__ T ysh[LEAN ? 6 * D + 18 : 1][LEAN ? NT : 1];
  const int n = job.n;
  const long long cells = job.cells;
  int co[3];
  long long str[3];
  const long long ii = adj_tile_cell<D, true>(n, co, str);
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
  // tile origin (contiguous axis = last)
  const int col0 = co[D - 1] - tx, row0 = co[D - 2] - ty;
  const long long pbase = (D == 3) ? (long long)co[0] * n * n : 0;
  const T N = T(n);
  T xn[D == 3 ? 9 : 1][4];
  if constexpr (D == 3) {
    long long xo[4];
    for (int k = 0; k < 4; ++k) xo[k] = ii + (long long)(wrapi(co[0] + (k < 2 ? k - 2 : k - 1), n) - co[0]) * str[0];
    for (int c = 0; c < 9; ++c)
      for (int k = 0; k < 4; ++k) xn[c][k] = job.beta[c * cells + xo[k]];
  }
  {
    long long foff[FILL];
    bool fval[FILL];
    for (int k = 0; k < FILL; ++k) {
      const int e = tid + k * NT;
      fval[k] = e < TILE;
      const int er = e / PW, ec = e % PW;
      foff[k] = pbase + (long long)wrapi(row0 + er - 2, n) * n + wrapi(col0 + ec - 2, n);
    }
    T tf[9][FILL];
    for (int c = 0; c < 9; ++c)
      for (int k = 0; k < FILL; ++k) tf[c][k] = fval[k] ? job.beta[c * cells + foff[k]] : T(0);
    for (int c = 0; c < 9; ++c)
      for (int k = 0; k < FILL; ++k)
        if (fval[k]) tile[c][tid + k * NT] = tf[c][k];
  }
  __syncthreads();
  T u[D][2][9], fc[9];
  // minmod selection of this cell's own limiter site m_i, per axis and
  // component: 2 bits at 2 (9 a + c), 3 = the zero case
  unsigned long long code = 0ull;
  const int lr = ty + 2, lc = tx + 2;
  // Every component's five-point stencils along the in-plane axes, read from
  // the LDS tile in three groups of three components, each group's reads
  // issued before the arithmetic of the group before it (at one wave per
  // SIMD every exposed LDS latency is lost time; the scheduling barriers
  // keep the compiler from sinking the reads to their first use).
  // fn(c, s): s[0..4] along the contiguous axis, s[5, 6, 7, 8] at rows
  // -2, -1, +1, +2 of the next axis, s[2] the cell.
  auto stencils = [&](auto&& fn) __attribute__((always_inline)) {
    T sv[2][3][9];
    auto rd = [&](int g, T (&sq)[3][9]) __attribute__((always_inline)) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const T* tb = tile[3 * g + k];
#pragma unroll
        for (int o = 0; o < 5; ++o) sq[k][o] = tb[lr * PW + lc - 2 + o];
        sq[k][5] = tb[(lr - 2) * PW + lc];
        sq[k][6] = tb[(lr - 1) * PW + lc];
        sq[k][7] = tb[(lr + 1) * PW + lc];
        sq[k][8] = tb[(lr + 2) * PW + lc];
      }
    };
    rd(0, sv[0]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int g = 0; g < 3; ++g) {
      if (g < 2) rd(g + 1, sv[(g + 1) & 1]);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < 3; ++k) fn(3 * g + k, sv[g & 1][k]);
      __builtin_amdgcn_sched_barrier(0);
    }
  };
  // the five-point stencil of component c along axis a
  auto stencil5 = [&](int a, int c, const T* sq, const T (*xv)[4], T (&f)[5]) __attribute__((always_inline)) {
    if (a == D - 1) {
#pragma unroll
      for (int k = 0; k < 5; ++k) f[k] = sq[k];
    } else if (a == D - 2) {
      f[0] = sq[5];
      f[1] = sq[6];
      f[2] = sq[2];
      f[3] = sq[7];
      f[4] = sq[8];
    } else {
      f[0] = xv[c][0];
      f[1] = xv[c][1];
      f[2] = sq[2];
      f[3] = xv[c][2];
      f[4] = xv[c][3 % 4];
    }
  };
  stencils([&](int c, const T* sq) __attribute__((always_inline)) {
    fc[c] = sq[2];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      T f[5];
      stencil5(a, c, sq, xn, f);
      lderivs<T>(f, N, u[a][0][c], u[a][1][c]);
      const T Dm1 = f[2] - f[1], D0 = f[3] - f[2], D1 = f[4] - f[3];
      const int sel = mm3_sel(D1 - D0, (D1 - Dm1) * T(0.5), D0 - Dm1);
      code |= (unsigned long long)(sel < 0 ? 3 : sel) << (2 * (9 * a + c));
    }
  });
  job.codes[ii] = code;
  T sg[9];
  {
    const T tr = fc[0] + fc[4] + fc[8];
    const T mu = T(job.mu), lam = T(job.lam);
    const int SC[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
    for (int c = 0; c < 9; ++c) {
      const int i = c / 3, j = c % 3;
      sg[c] = ld_pol<STRM>(&job.G[SC[c] * cells + ii]) - mu * (fc[c] + fc[3 * j + i]) -
              (i == j ? lam * tr : T(0));
    }
  }
  // rbar scaled by rscale (times the step's dt when dtp is given)
  const T rs = T(job.rscale) * (job.dtp ? T(*job.dtp) : T(1));
  T rb[9];
  for (int c = 0; c < 9; ++c) rb[c] = rs * ld_pol<STRM>(&job.rbar[c * cells + ii]);
  T sgb[9], lamb;
  if constexpr (LEAN) {
    double* ubar = job.ubar;
    // the derivatives again for the final pass: the plane tile is still in
    // LDS, the slow-axis neighbours are read again (from L2, mostly)
    auto reload = [&](T (&uu)[D][2][9]) {
      long long xo[4];
      for (int k = 0; k < 4; ++k) xo[k] = ii + (long long)(wrapi(co[0] + (k < 2 ? k - 2 : k - 1), n) - co[0]) * str[0];
      T xr[D == 3 ? 9 : 1][4];
      if constexpr (D == 3) {
#pragma unroll
        for (int c = 0; c < 9; ++c)
#pragma unroll
          for (int k = 0; k < 4; ++k) xr[c][k] = job.beta[c * cells + xo[k]];
      }
      stencils([&](int c, const T* sq) __attribute__((always_inline)) {
#pragma unroll
        for (int a = 0; a < D; ++a) {
          T f[5];
          stencil5(a, c, sq, xr, f);
          lderivs<T>(f, N, uu[a][0][c], uu[a][1][c]);
        }
      });
    };
    cell_flux_vjp_lean<T, D, DYN>(
        u, sg, T(job.Lambda), rb, &vsh[0][tid], &ysh[0][tid], NT, sgb, lamb,
        [&](int a, int s, int c, T val) {
          st_pol<STRM>(&ubar[((a * 2 + s) * 9 + c) * cells + ii], N * val);
        },
        reload);
  } else {
    T ub[D][2][9];
    cell_flux_vjp<T, D, DYN>(u, sg, T(job.Lambda), rb, ub, sgb, lamb);
    for (int a = 0; a < D; ++a)
      for (int s = 0; s < 2; ++s)
        for (int c = 0; c < 9; ++c) job.ubar[((a * 2 + s) * 9 + c) * cells + ii] = N * ub[a][s][c];
  }
  // the symmetrised stress cotangent as six planes, xx yy zz xy xz yz
  {
    constexpr int SI[6] = {0, 1, 2, 0, 0, 1}, SJ[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int q = 0; q < 6; ++q)
      st_pol<STRM>(&job.sbar[q * cells + ii], T(0.5) * (sgb[3 * SI[q] + SJ[q]] + sgb[3 * SJ[q] + SI[q]]));
  }
  const int nw = (int)(blockDim.x * blockDim.y + 63) / 64;
  for (int off = 32; off > 0; off >>= 1) lamb += __shfl_xor(lamb, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = lamb;
  __syncthreads();
  if (tid == 0) {
    T t = T(0);
    for (int k = 0; k < nw; ++k) t += red[k];
    job.lpart[blockIdx.x] = t;
  }
  if (job.upart) {
    T uz = T(job.lam) * (sgb[0] + sgb[4] + sgb[8]) + T(2) * T(job.mu) * sgb[8];
    for (int off = 32; off > 0; off >>= 1) uz += __shfl_xor(uz, off, 64);
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = uz;
    __syncthreads();
    if (tid == 0) {
      T t = T(0);
      for (int k = 0; k < nw; ++k) t += red[k];
      job.upart[blockIdx.x] = t;
    }
  }
}

// ------------------------------------------------------------------ frozen branches
// k_flux_branch: the flux of cell_flux_vjp's forward recompute (the same
// operations as k_flux_vjp_cell, derivatives in reference units) with its
// branches -- the minmod selection of every limiter site and, per axis, the
// corners of the largest and smallest velocity and whether they pass zero --
//   MODE 0: decided at the state (the ordinary flux),
//   MODE 1: decided and written to job.sel (FrozenJob),
//   MODE 2: read from job.sel (the branch-frozen map, smooth in beta).
// V = Dual carries a tangent with every value (forward-mode differentiation:
// the derivative of each operation at its value, branches decided on the
// values), so MODE 0 with Dual gives the tangent-linear map of the flux,
// J_L(beta) dbeta + (dL/dLambda) dLambda, exactly.

// A value and its tangent.
struct Dual {
  double v, d;
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, double s) { return {a.v * s, a.d * s}; }
__device__ __forceinline__ Dual operator*(double s, Dual a) { return {a.v * s, a.d * s}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
__device__ __forceinline__ Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
__device__ __forceinline__ Dual& operator*=(Dual& a, Dual b) { return a = a * b; }

__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ double val(Dual x) { return x.v; }
template <typename V>
__device__ __forceinline__ V cst(double x) {
  if constexpr (sizeof(V) == sizeof(double)) return x;
  else return V{x, 0.0};
}
// s m + d (s a constant)
__device__ __forceinline__ double sfma(double s, double m, double d) { return fma(s, m, d); }
__device__ __forceinline__ Dual sfma(double s, Dual m, Dual d) { return {fma(s, m.v, d.v), s * m.d + d.d}; }
__device__ __forceinline__ Dual rsqrt_nr(Dual x) {
  const double y = rsqrt_nr(x.v);
  return {y, -0.5 * x.d * y * y * y};
}

template <typename V>
__device__ __forceinline__ V mm_pick(unsigned s, V a, V b, V c) {
  return s == 0u ? a : (s == 1u ? b : (s == 2u ? c : cst<V>(0.0)));
}

template <typename V, int D, int DYN, int TX, int TY, int MODE>
__global__ void __launch_bounds__(TX* TY) k_flux_branch(FrozenJob job) {
  constexpr int PW = TX + 4, TILE = PW * (TY + 4), NT = TX * TY, FILL = (TILE + NT - 1) / NT;
  constexpr int NCOR = 1 << D;
  constexpr bool TAN = sizeof(V) != sizeof(double);
  __shared__ double tile[9][TILE];
  __shared__ double dtile[TAN ? 9 : 1][TAN ? TILE : 1];
  const int n = job.n;
  const long long cells = job.cells;
  int co[3];
  long long str[3];
  const long long ii = adj_tile_cell<D>(n, co, str);
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
  const int col0 = co[D - 1] - tx, row0 = co[D - 2] - ty;
  const long long pbase = (D == 3) ? (long long)co[0] * n * n : 0;
  const double N = double(n);
  auto load = [&](int c, long long off) -> V {
    if constexpr (TAN) return V{job.beta[c * cells + off], job.dbeta[c * cells + off]};
    else return job.beta[c * cells + off];
  };
  V xn[D == 3 ? 9 : 1][4];
  if constexpr (D == 3) {
    for (int k = 0; k < 4; ++k) {
      const long long xo = ii + (long long)(wrapi(co[0] + (k < 2 ? k - 2 : k - 1), n) - co[0]) * str[0];
      for (int c = 0; c < 9; ++c) xn[c][k] = load(c, xo);
    }
  }
  for (int k = 0; k < FILL; ++k) {
    const int e = tid + k * NT;
    if (e < TILE) {
      const int er = e / PW, ec = e % PW;
      const long long off = pbase + (long long)wrapi(row0 + er - 2, n) * n + wrapi(col0 + ec - 2, n);
      for (int c = 0; c < 9; ++c) {
        tile[c][e] = job.beta[c * cells + off];
        if constexpr (TAN) dtile[c][e] = job.dbeta[c * cells + off];
      }
    }
  }
  __syncthreads();
  auto tv = [&](int c, int e) -> V {
    if constexpr (TAN) return V{tile[c][e], dtile[c][e]};
    else return tile[c][e];
  };
  // replay: this cell's words and, per axis, the site word of the cell before it
  unsigned long long own = 0ull, spd = 0ull, prev[3] = {0ull, 0ull, 0ull};
  if constexpr (MODE == 2) {
    own = job.sel[ii];
    spd = job.sel[cells + ii];
    for (int a = 0; a < D; ++a) prev[a] = job.sel[ii + (long long)(wrapi(co[a] - 1, n) - co[a]) * str[a]];
  }
  V u[D][2][9], fc[9];
  unsigned long long code = 0ull;
  const int lr = ty + 2, lc = tx + 2;
  for (int a = 0; a < D; ++a) {
    for (int c = 0; c < 9; ++c) {
      V f[5];
      if (a == D - 1) {
        for (int k = 0; k < 5; ++k) f[k] = tv(c, lr * PW + lc + k - 2);
      } else if (a == D - 2) {
        for (int k = 0; k < 5; ++k) f[k] = tv(c, (lr + k - 2) * PW + lc);
      } else {
        f[0] = xn[c][0];
        f[1] = xn[c][1];
        f[2] = tv(c, lr * PW + lc);
        f[3] = xn[c][2];
        f[4] = xn[c][3 % 4];
      }
      if (a == 0) fc[c] = f[2];
      const V Dm2 = f[1] - f[0], Dm1 = f[2] - f[1], D0 = f[3] - f[2], D1 = f[4] - f[3];
      const V a1 = D1 - D0, b1 = (D1 - Dm1) * 0.5, c1 = D0 - Dm1;
      const V a2 = D0 - Dm1, b2 = (D0 - Dm2) * 0.5, c2 = Dm1 - Dm2;
      unsigned s1, s2;
      if constexpr (MODE == 2) {
        const int sh = 2 * (9 * a + c);
        s1 = (unsigned)(own >> sh) & 3u;
        s2 = (unsigned)(prev[a] >> sh) & 3u;
      } else {
        const int k1 = mm3_sel(val(a1), val(b1), val(c1)), k2 = mm3_sel(val(a2), val(b2), val(c2));
        s1 = k1 < 0 ? 3u : (unsigned)k1;
        s2 = k2 < 0 ? 3u : (unsigned)k2;
        if constexpr (MODE == 1) code |= (unsigned long long)s1 << (2 * (9 * a + c));
      }
      const V m_i = mm_pick<V>(s1, a1, b1, c1), m_im1 = mm_pick<V>(s2, a2, b2, c2);
      u[a][0][c] = sfma(-0.5, m_i, D0) * N;
      u[a][1][c] = sfma(0.5, m_im1, Dm1) * N;
    }
  }
  V sg[9];
  {
    const V tr = fc[0] + fc[4] + fc[8];
    const double mu = job.mu, lam = job.lam;
    const int SC[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
    for (int c = 0; c < 9; ++c) {
      const int i = c / 3, j = c % 3;
      V g;
      if constexpr (TAN) g = V{job.G[SC[c] * cells + ii], job.dG[SC[c] * cells + ii]};
      else g = job.G[SC[c] * cells + ii];
      sg[c] = g - mu * (fc[c] + fc[3 * j + i]) - (i == j ? lam * tr : cst<V>(0.0));
    }
  }
  const double ME = kME;
  V Lambda;
  if constexpr (TAN) Lambda = V{job.Lambda, job.dLambda};
  else Lambda = job.Lambda;
  const bool climb_sup = (DYN == DYN_GLIDE_CLIMB) && (job.Lambda != 0.0);
  const double third = 1.0 / 3.0;
  const V st = climb_sup ? Lambda * (sg[0] + sg[4] + sg[8]) * third : cst<V>(0.0);
  const V jfac = -Lambda * third;
  V w[D][2][3], zz[D][2], trs[D][2], P[D][2];
  for (int a = 0; a < D; ++a)
    for (int s = 0; s < 2; ++s) {
      const V* ua = u[a][s];
      for (int l = 0; l < 3; ++l)
        w[a][s][l] = ua[3 * l] * sg[3 * a] + ua[3 * l + 1] * sg[3 * a + 1] + ua[3 * l + 2] * sg[3 * a + 2];
      V z = cst<V>(0.0);
      for (int c = 0; c < 9; ++c) z += ua[c] * sg[c];
      zz[a][s] = z;
      trs[a][s] = ua[0] + ua[4] + ua[8];
      V p = cst<V>(0.0);
      for (int m = 0; m < 3; ++m)
        if (m != a)
          for (int jj = 0; jj < 3; ++jj) p += ua[3 * m + jj] * ua[3 * m + jj];
      P[a][s] = p;
    }
  V v[NCOR][3], J[NCOR];
  for (int c = 0; c < NCOR; ++c) {
    int sg_[3] = {0, 0, 0};
    for (int a = 0; a < D; ++a) sg_[a] = (c >> (D - 1 - a)) & 1;
    V R[3], q[3], v0[3];
    for (int l = 0; l < 3; ++l) {
      V r = cst<V>(0.0), vv = cst<V>(0.0);
      for (int a = 0; a < D; ++a) {
        r += u[a][sg_[a]][3 * l + a];
        vv += w[a][sg_[a]][l];
      }
      R[l] = r;
      v0[l] = vv;
    }
    for (int a = 0; a < D; ++a) v0[a] -= zz[a][sg_[a]];
    for (int a = 0; a < 3; ++a) q[a] = (a < D ? trs[a % D][sg_[a]] : cst<V>(0.0)) - R[a];
    if (DYN == DYN_GLIDE_CLIMB) {
      if (climb_sup) {
        for (int a = 0; a < D; ++a) v0[a] += st * trs[a][sg_[a]];
        for (int l = 0; l < 3; ++l) v0[l] -= st * R[l];
      }
    } else {
      V rrr = cst<V>(ME);
      for (int a = 0; a < D; ++a) rrr += trs[a][sg_[a]] * trs[a][sg_[a]] - 2.0 * trs[a][sg_[a]] * R[a];
      for (int l = 0; l < 3; ++l) rrr += R[l] * R[l];
      const V rvv = cst<V>(ME) + v0[0] * q[0] + v0[1] * q[1] + v0[2] * q[2];
      const V fpr = rvv / rrr;
      for (int a = 0; a < 3; ++a) v0[a] += q[a] * fpr;
    }
    V r2 = cst<V>(ME);
    for (int a = 0; a < D; ++a) r2 += P[a][sg_[a]];
    V cross = cst<V>(0.0);
    for (int jj = 0; jj < 3; ++jj) {
      cross += u[0][sg_[0]][3 + jj] * u[1][sg_[1]][jj];
      if (D == 3) {
        cross += u[1][sg_[1]][6 + jj] * u[D - 1][sg_[D - 1]][3 + jj];
        cross += u[D - 1][sg_[D - 1]][jj] * u[0][sg_[0]][6 + jj];
      }
    }
    r2 -= 2.0 * cross;
    const V rinv = rsqrt_nr(r2);
    for (int l = 0; l < 3; ++l) v[c][l] = v0[l] * rinv;
    J[c] = climb_sup ? jfac * (v[c][0] * q[0] + v[c][1] * q[1] + v[c][2] * q[2]) : cst<V>(0.0);
  }
  V ap[3], am[3], at[3];
  unsigned long long spd_out = 0ull;
  for (int a = 0; a < D; ++a) {
    int cmax, cmin;
    bool pos, neg;
    if constexpr (MODE == 2) {
      const unsigned b = (unsigned)(spd >> (8 * a)) & 255u;
      cmax = (int)(b & 7u);
      cmin = (int)((b >> 3) & 7u);
      pos = (b >> 6) & 1u;
      neg = (b >> 7) & 1u;
    } else {
      cmax = cmin = 0;
      double mxv = val(v[0][a]), mnv = val(v[0][a]);
      for (int c = 1; c < NCOR; ++c) {
        if (val(v[c][a]) > mxv) {
          mxv = val(v[c][a]);
          cmax = c;
        }
        if (val(v[c][a]) < mnv) {
          mnv = val(v[c][a]);
          cmin = c;
        }
      }
      pos = mxv > 0.0;
      neg = mnv < 0.0;
      spd_out |= (unsigned long long)(cmax | (cmin << 3) | ((int)pos << 6) | ((int)neg << 7)) << (8 * a);
    }
    ap[a] = pos ? v[cmax][a] : cst<V>(0.0);
    am[a] = neg ? -v[cmin][a] : cst<V>(0.0);
    at[a] = ap[a] + am[a] + cst<V>(ME);
  }
  V Y[D][2][3], Jsum = cst<V>(0.0);
  for (int a = 0; a < D; ++a)
    for (int s = 0; s < 2; ++s)
      for (int l = 0; l < 3; ++l) Y[a][s][l] = cst<V>(0.0);
  for (int c = 0; c < NCOR; ++c) {
    V wc = cst<V>(1.0);
    for (int a = 0; a < D; ++a) wc *= ((c >> (D - 1 - a)) & 1) == 0 ? am[a] : ap[a];
    for (int a = 0; a < D; ++a)
      for (int l = 0; l < 3; ++l) Y[a][(c >> (D - 1 - a)) & 1][l] += wc * v[c][l];
    Jsum += wc * J[c];
  }
  V inv = cst<V>(1.0);
  for (int a = 0; a < D; ++a) inv *= at[a];
  inv = cst<V>(1.0) / inv;
  V twv[3];
  for (int a = 0; a < D; ++a) {
    V o = inv;
    for (int b = 0; b < D; ++b)
      if (b != a) o *= at[b];
    twv[a] = am[a] * ap[a] * o;
  }
  for (int i = 0; i < 3; ++i)
    for (int jj = 0; jj < 3; ++jj) {
      const int cc = 3 * i + jj;
      V h = (i == jj) ? Jsum : cst<V>(0.0);
      for (int a = 0; a < D; ++a)
        for (int s = 0; s < 2; ++s) h += u[a][s][cc] * Y[a][s][a];
      if (i < D)
        for (int s = 0; s < 2; ++s)
          for (int l = 0; l < 3; ++l) h -= u[i % D][s][3 * l + jj] * Y[i % D][s][l];
      V r = -h * inv;
      for (int a = 0; a < D; ++a) r += (u[a][0][cc] - u[a][1][cc]) * twv[a];
      job.rhs[cc * cells + ii] = val(r);
      if constexpr (TAN) job.drhs[cc * cells + ii] = r.d;
    }
  if constexpr (MODE == 1) {
    job.sel[ii] = code;
    job.sel[cells + ii] = spd_out;
  }
}

// ------------------------------------------------------------------ R2
// Transpose of the limited-derivative stencils, in site form.  With
// D_k = f_{k+1} - f_k, up_i = N (D_i - m_i / 2), um_i = N (D_{i-1} + m_{i-1} / 2)
// and the limiter site m_k = minmod(D_{k+1} - D_k, (D_{k+1} - D_{k-1}) / 2,
// D_k - D_{k-1}), the cotangents P_i = N upbar_i and Q_i = N umbar_i give
//   Mbar_k = -P_k / 2 + Q_{k+1} / 2                       (site k)
//   Dbar_k = P_k + Q_{k+1} + up(k-1) + mid(k) + dn(k+1)    (site terms)
// where site k passes (up, mid, dn) = Mbar_k x {(1, -1, 0), (1/2, 0, -1/2),
// (0, 1, -1), 0} for its selection {first, second, third, zero} to
// (D_{k+1}, D_k, D_{k-1}); then betabar_j = Dbar_{j-1} - Dbar_j.  The
// selections come from the codes the cell kernel stored, so no limiter is
// re-evaluated here.  Added: the local stress transpose
// -lam tr(sbar) d - 2 mu sbar (sbar symmetric), the spectral part ebar when
// given, and the Runge-Kutta combination sum_k c_k a_k (read at the thread's
// own cell only, so out may alias an a_k).
template <typename T>
__device__ __forceinline__ void site_coef(unsigned sel, T& up, T& mid, T& dn) {
  up = sel == 0u ? T(1) : (sel == 1u ? T(0.5) : T(0));
  mid = sel == 0u ? T(-1) : (sel == 2u ? T(1) : T(0));
  dn = sel == 1u ? T(-0.5) : (sel == 2u ? T(-1) : T(0));
}

template <typename T, int D>
__global__ void __launch_bounds__(256) k_flux_vjp_gather(GatherJob g) {
  constexpr bool STRM = (D == 3);  // as in k_flux_vjp_cell
  const int n = g.n;
  const long long cells = g.cells;
  const T* __restrict__ ubar = g.ubar;
  int co[3];
  long long str[3];
  const long long idx = adj_tile_cell<D, true>(n, co, str);
  T acc[9];
  for (int c = 0; c < 9; ++c) acc[c] = T(0);
  for (int a = 0; a < D; ++a) {
    long long off[5];  // cells j-2 .. j+2 along axis a
    for (int k = 0; k < 5; ++k) off[k] = idx + (long long)(wrapi(co[a] + k - 2, n) - co[a]) * str[a];
    unsigned long long code[4];  // sites j-2 .. j+1
    for (int s = 0; s < 4; ++s) code[s] = g.codes[off[s]];
    for (int c = 0; c < 9; ++c) {
      const T* Pp = ubar + ((a * 2 + 0) * 9 + c) * cells;
      const T* Qp = ubar + ((a * 2 + 1) * 9 + c) * cells;
      T P[4], Q[4];  // P at cells j-2..j+1, Q at cells j-1..j+2
      for (int s = 0; s < 4; ++s) {
        P[s] = Pp[off[s]];
        Q[s] = Qp[off[s + 1]];
      }
      T up[4], mid[4], dn[4];
      for (int s = 0; s < 4; ++s) {
        const T Mb = T(0.5) * (Q[s] - P[s]);
        site_coef<T>((unsigned)(code[s] >> (2 * (9 * a + c))) & 3u, up[s], mid[s], dn[s]);
        up[s] *= Mb;
        mid[s] *= Mb;
        dn[s] *= Mb;
      }
      const T Dbm1 = P[1] + Q[1] + up[0] + mid[1] + dn[2];
      const T Db0 = P[2] + Q[2] + up[1] + mid[2] + dn[3];
      acc[c] += Dbm1 - Db0;
    }
  }
  T s[6];  // sbar: six planes, xx yy zz xy xz yz
  for (int q = 0; q < 6; ++q) s[q] = ld_pol<STRM>(&g.sbar[q * cells + idx]);
  const T trs = s[0] + s[1] + s[2];
  const int SC[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
  for (int c = 0; c < 9; ++c) {
    const int i = c / 3, j = c % 3;
    T v = acc[c] - T(2) * T(g.mu) * s[SC[c]] - (i == j ? T(g.lam) * trs : T(0));
    if (g.ebar) v += ld_pol<STRM>(&g.ebar[SC[c] * cells + idx]);
    for (int k = 0; k < g.nterm; ++k) v += T(g.coef[k]) * ld_pol<STRM>(&g.term[k][c * cells + idx]);
    if (g.csrc) v += g.csrc[c];
    acc[c] = v;
  }
  if (g.uzz) acc[8] += T(g.uzz_scale) * g.uzz[0];
  for (int c = 0; c < 9; ++c) st_pol<STRM>(&g.out[c * cells + idx], acc[c]);
}

// sbar = S : sigma (the energy density's stress gradient, PlasticityState.py:
// 317-330: e = 1/2 S_ijkm sigma_ij sigma_km), per-block energy sums.
__global__ void __launch_bounds__(256) k_energy_grad(const double* G, const double* beta, long long cells, double mu,
                                                    double nu, double lam, double* sbar, double* partial) {
  __shared__ double red[4];
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (idx < cells) {
    double b[9], s[9];
    for (int c = 0; c < 9; ++c) b[c] = beta[c * cells + idx];
    const double tr = b[0] + b[4] + b[8];
    const int SC[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
    for (int c = 0; c < 9; ++c) {
      const int i = c / 3, j = c % 3;
      s[c] = G[SC[c] * cells + idx] - mu * (b[c] + b[3 * j + i]) - (i == j ? lam * tr : 0.0);
    }
    const double trs = s[0] + s[4] + s[8];
    double ss = 0.0, sst = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        ss += s[3 * i + j] * s[3 * i + j];
        sst += s[3 * i + j] * s[3 * j + i];
      }
    e = 0.5 * (2.0 * ss - sst - nu / (1.0 + nu) * trs * trs) / (2.0 * mu);
    // de/dsigma_ij = (2 s_ij - s_ji - nu/(1+nu) tr d_ij) / (2 mu), symmetrised
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const double g1 = (2.0 * s[3 * i + j] - s[3 * j + i] - (i == j ? nu / (1.0 + nu) * trs : 0.0)) / (2.0 * mu);
        const double g2 = (2.0 * s[3 * j + i] - s[3 * i + j] - (i == j ? nu / (1.0 + nu) * trs : 0.0)) / (2.0 * mu);
        sbar[(3 * i + j) * cells + idx] = 0.5 * (g1 + g2);
      }
  }
  for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < (int)(blockDim.x + 63) / 64; ++k) t += red[k];
    partial[blockIdx.x] = t;
  }
}

// Per-block sums of lam tr(sbar) + 2 mu sbar_zz: the cotangent of the
// uniaxial control's betaE_zz = rate t - (<beta_zz> - init)
// (RunLoadingMPI.py:68-83), whose stress enters every cell.
__global__ void __launch_bounds__(256) k_uniaxial_partial(const double* sbar, long long cells, double lam, double mu,
                                                         double* partial) {
  __shared__ double red[4];
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double u = 0.0;
  if (idx < cells)
    u = lam * (sbar[idx] + sbar[4 * cells + idx] + sbar[8 * cells + idx]) + 2.0 * mu * sbar[8 * cells + idx];
  for (int off = 32; off > 0; off >>= 1) u += __shfl_xor(u, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = u;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void k_add_uniform_zz(double* out, long long cells, const double* s, double scale) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < cells) out[8 * cells + idx] += scale * s[0];
}

hipError_t launch_uniaxial_partial(const double* sbar, long long cells, double lam, double mu, double* partial,
                                   int* nblocks, hipStream_t st) {
  const int nb = (int)((cells + 255) / 256);
  *nblocks = nb;
  hipLaunchKernelGGL(k_uniaxial_partial, dim3(nb), dim3(256), 0, st, sbar, cells, lam, mu, partial);
  return hipGetLastError();
}

hipError_t launch_add_uniform_zz(double* out, long long cells, const double* s, double scale, hipStream_t st) {
  hipLaunchKernelGGL(k_add_uniform_zz, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, st, out, cells, s, scale);
  return hipGetLastError();
}

// out = sum_k c[k] in[k] (k < nin), elementwise over len values
__global__ void k_lincomb(long long len, int nin, const double* a0, double c0, const double* a1, double c1,
                          const double* a2, double c2, const double* a3, double c3, double* out) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  double v = c0 * a0[i];
  if (nin > 1) v += c1 * a1[i];
  if (nin > 2) v += c2 * a2[i];
  if (nin > 3) v += c3 * a3[i];
  out[i] = v;
}

// Segment replay clock: mode 0 saves the plan's clock in clk[1] and sets up
// a replay of `len` recorded steps from time t0; mode 1 restores clk[1].
__global__ void k_clock_segment(Clock* clk, int mode, double t0, const double* dts, long long len) {
  if (threadIdx.x != 0) return;
  if (mode == 0) {
    clk[1] = clk[0];
    Clock c = clk[0];
    c.t = t0;
    c.t_end = INFINITY;
    c.done = 0;
    c.fixed_dt = 0.0;
    c.dt_sched = dts;
    c.sched_base = c.nsteps;
    c.sched_len = len;
    c.dt_log = nullptr;
    c.log_base = 0;
    c.log_cap = 0;
    c.slog = nullptr;
    c.slog_base = 0;
    c.slog_cap = 0;
    c.elog = nullptr;
    c.elog_base = 0;
    c.elog_cap = 0;
    c.vlog = nullptr;
    c.vlog_base = 0;
    c.vlog_cap = 0;
    c.status = 0;
    clk[0] = c;
  } else {
    clk[0] = clk[1];
  }
}

hipError_t launch_clock_segment(Clock* clk, int mode, double t0, const double* dts, long long len, hipStream_t st) {
  hipLaunchKernelGGL(k_clock_segment, dim3(1), dim3(64), 0, st, clk, mode, t0, dts, len);
  return hipGetLastError();
}

static dim3 adj_block(int n) { return dim3(n < 32 ? n : 32, n < 8 ? n : 8); }

int flux_vjp_nblocks(int d, int n) {
  const dim3 b = adj_block(n);
  return (d == 3 ? n : 1) * (n / (int)b.x) * (n / (int)b.y);
}

template <int TX, int TY>
static void launch_vjp_cell_t(const VjpJob& j, int nb, hipStream_t st) {
  const dim3 bs(TX, TY);
  if (j.d == 3) {
    if (j.dyn == DYN_GLIDE_ONLY) hipLaunchKernelGGL((k_flux_vjp_cell<double, 3, DYN_GLIDE_ONLY, TX, TY>), dim3(nb), bs, 0, st, j);
    else hipLaunchKernelGGL((k_flux_vjp_cell<double, 3, DYN_GLIDE_CLIMB, TX, TY>), dim3(nb), bs, 0, st, j);
  } else {
    if (j.dyn == DYN_GLIDE_ONLY) hipLaunchKernelGGL((k_flux_vjp_cell<double, 2, DYN_GLIDE_ONLY, TX, TY>), dim3(nb), bs, 0, st, j);
    else hipLaunchKernelGGL((k_flux_vjp_cell<double, 2, DYN_GLIDE_CLIMB, TX, TY>), dim3(nb), bs, 0, st, j);
  }
}

hipError_t launch_flux_vjp_cell(const VjpJob& j, int* nblocks, hipStream_t st) {
  const dim3 bs = adj_block(j.n);
  const int nb = flux_vjp_nblocks(j.d, j.n);
  *nblocks = nb;
  if (bs.x == 32 && bs.y == 8) launch_vjp_cell_t<32, 8>(j, nb, st);
  else if (bs.x == 16 && bs.y == 8) launch_vjp_cell_t<16, 8>(j, nb, st);
  else if (bs.x == 8 && bs.y == 8) launch_vjp_cell_t<8, 8>(j, nb, st);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

template <typename V, int MODE, int TX, int TY>
static void launch_branch_t(const FrozenJob& j, int nb, hipStream_t st) {
  const dim3 bs(TX, TY);
  if (j.d == 3) {
    if (j.dyn == DYN_GLIDE_ONLY)
      hipLaunchKernelGGL((k_flux_branch<V, 3, DYN_GLIDE_ONLY, TX, TY, MODE>), dim3(nb), bs, 0, st, j);
    else
      hipLaunchKernelGGL((k_flux_branch<V, 3, DYN_GLIDE_CLIMB, TX, TY, MODE>), dim3(nb), bs, 0, st, j);
  } else {
    if (j.dyn == DYN_GLIDE_ONLY)
      hipLaunchKernelGGL((k_flux_branch<V, 2, DYN_GLIDE_ONLY, TX, TY, MODE>), dim3(nb), bs, 0, st, j);
    else
      hipLaunchKernelGGL((k_flux_branch<V, 2, DYN_GLIDE_CLIMB, TX, TY, MODE>), dim3(nb), bs, 0, st, j);
  }
}

template <int TX, int TY>
static void launch_branch_tiles(const FrozenJob& j, int nb, hipStream_t st) {
  if (j.dbeta) launch_branch_t<Dual, 0, TX, TY>(j, nb, st);
  else if (j.record) launch_branch_t<double, 1, TX, TY>(j, nb, st);
  else launch_branch_t<double, 2, TX, TY>(j, nb, st);
}

hipError_t launch_flux_frozen(const FrozenJob& j, hipStream_t st) {
  const dim3 bs = adj_block(j.n);
  const int nb = flux_vjp_nblocks(j.d, j.n);
  if (bs.x == 32 && bs.y == 8) launch_branch_tiles<32, 8>(j, nb, st);
  else if (bs.x == 16 && bs.y == 8) launch_branch_tiles<16, 8>(j, nb, st);
  else if (bs.x == 8 && bs.y == 8) launch_branch_tiles<8, 8>(j, nb, st);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_flux_vjp_gather(const GatherJob& g, hipStream_t st) {
  if (g.nterm < 0 || g.nterm > 3) return hipErrorInvalidValue;
  const dim3 bs = adj_block(g.n);
  const int nb = flux_vjp_nblocks(g.d, g.n);
  if (g.d == 3)
    hipLaunchKernelGGL((k_flux_vjp_gather<double, 3>), dim3(nb), bs, 0, st, g);
  else
    hipLaunchKernelGGL((k_flux_vjp_gather<double, 2>), dim3(nb), bs, 0, st, g);
  return hipGetLastError();
}

hipError_t launch_energy_grad(const double* G, const double* beta, long long cells, double mu, double nu, double lam,
                              double* sbar, double* partial, int* nblocks, hipStream_t st) {
  const int nb = (int)((cells + 255) / 256);
  *nblocks = nb;
  hipLaunchKernelGGL(k_energy_grad, dim3(nb), dim3(256), 0, st, G, beta, cells, mu, nu, lam, sbar, partial);
  return hipGetLastError();
}

hipError_t launch_lincomb(long long len, int nin, const double* a0, double c0, const double* a1, double c1,
                          const double* a2, double c2, const double* a3, double c3, double* out, hipStream_t st) {
  const int bs = 256;
  hipLaunchKernelGGL(k_lincomb, dim3((unsigned)((len + bs - 1) / bs)), dim3(bs), 0, st, len, nin, a0, c0, a1, c1, a2, c2,
                     a3, c3, out);
  return hipGetLastError();
}

}  // namespace cdd
