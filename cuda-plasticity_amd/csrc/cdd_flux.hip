// Fused central-upwind Hamilton-Jacobi flux + TVD Runge-Kutta stage kernels.
//
// Replaces, per RK stage, CentralUpwindHJDynamics.CalculateFlux
// (Plasticity/FieldDynamics/CentralUpwindHJ3D.py:63-271) with the BetaP
// velocities/Hamiltonians of CentralUpwindHJBetaPDynamics3D.py:34-221
// (glide + climb, climb suppression Lambda) and
// CentralUpwindHJBetaPGlideOnlyDynamics.py:24-262 ('lvp' glide only), plus
// the stage combination of TVDRungeKutta_FieldMover
// (Plasticity/FieldMovers/FieldMover.py:115-135).
//
// One thread owns one cell.  It reads the stage field at the 4d+1 cells of
// the radius-2 cross stencil (optionally forming F0 = u + dt L0 on the fly,
// so the first RK stage update never makes its own HBM pass), the six
// non-local stress planes G and, for the combination, the step-start state
// at the cell, and writes nine output planes.
//
// The reference evaluates the corner velocity 2^d + d 2^d times per stage and
// the Hamiltonian once per corner with full 9x3 contractions.  Both are
// linear in per-axis quantities, which this kernel exploits:
//   v_l(corner)  = sum_a w_a^{s_a}[l] - [l<d] z_l^{s_l}  (+ Lambda / lvp terms)
//   w_a^s[l] = sum_n u_a^s[l,n] sigma_an,  z_a^s = sum_mn u_a^s[m,n] sigma_mn
//   |rho|^2 = ME + sum_a P_a^{s_a} - 2 sum_pairs C_ab^{s_a s_b}
//   sum_c w_c H_ij(c) = sum_a sum_s u_a^s[i,j] Y_a^s[a]
//                       - [i<d] sum_s sum_l u_i^s[l,j] Y_i^s[l] + d_ij sum_c w_c J_c
//   with Y_a^s[l] = sum over corners with s_a = s of w_c v_l(c).
// The rounding differs from the reference at the 1e-16 level only.
#include "cdd_kernels.h"

// Multiply-adds fuse only within one expression in this file (the build's
// default lets the compiler fuse across statements, choosing by the
// surrounding code): every kernel here that states the same arithmetic --
// the per-tile, multi-instance and row-marching fluxes share cell_flux --
// then rounds it the same way, so their states agree bit for bit.
#pragma clang fp contract(on)

namespace cdd {

// Two-argument minmod: min if both > 0, max if both <= 0, else 0.  That is
// the median of (a, c, 0), max(min(a, c), min(max(a, c), 0)): four min/max
// and no compares or selects.  min and max return one of their operands, so
// the value is exactly the minmod (only the sign of a zero result may differ
// from other formulations).
template <typename T>
__device__ __forceinline__ T minmod2(T a, T c) {
  return fmax(fmin(a, c), fmin(fmax(a, c), T(0)));
}

// One-sided limited derivatives of f at i from f[i-2..i+2]
// (FindDerivatives, CentralUpwindHJ3D.py:11-34, theta = 1, Dx = 1/N_last),
// returned in units of 1/N: the reference's u+ = N (D_i - m_i / 2) and
// u- = N (D_{i-1} + m_{i-1} / 2), with D_i = f_{i+1} - f_i and
//   m_i = minmod3(D_{i+1} - D_i, (D_{i+1} - D_{i-1}) / 2, D_i - D_{i-1})
// (NumericalMethods.py:367-378 MinMod3; the reference's 1/Dx^2 and Dx/2
// scalings are powers of two and commute with it).  With theta = 1 the
// middle argument is the mean of the other two, so it lies between them
// whenever they share a sign and minmod3 equals minmod2 of the outer two;
// the rounded mean can step outside by an ulp, so the two forms agree to
// the last bit except at such ties.  The factor N is applied once to the
// flux: every term of the right-hand side is homogeneous of degree one in
// the derivatives once the |rho| regulariser is taken in the same units
// (ME / N^2); for N a power of two the rescaling is exact and the result is
// bit-identical to scaling each derivative.
template <typename T>
__device__ __forceinline__ void limited_derivs(const T f[5], T& up, T& um) {
  const T Dm2 = f[1] - f[0], Dm1 = f[2] - f[1], D0 = f[3] - f[2], D1 = f[4] - f[3];
  const T a0 = D0 - Dm1;
  const T m_i = minmod2(D1 - D0, a0);
  const T m_im1 = minmod2(a0, Dm1 - Dm2);
  up = fma(T(-0.5), m_i, D0);
  um = fma(T(0.5), m_im1, Dm1);
}

// The RK stage combination cu U + cf (f + dt rhs) with its roundings fixed
// by explicit fused multiply-adds: left to contraction, the compiler may
// fuse either product into the addition depending on the surrounding code,
// and every kernel that forms a stage must give the same bits.
template <typename T>
__device__ __forceinline__ T rk_combine(T cu, T U, T cf, T f, T dt, T rhs) {
  return fma(cu, U, cf * fma(dt, rhs, f));
}

// Per-cell flux.  u[a][s][c]: derivative along axis a, s = 0 ('+') or 1
// ('-'), component c = 3 i + j, in units of 1/N (limited_derivs); rhs comes
// out in the reference's units.  sg: full (symmetric) stress.  vbuf: this
// thread's slot of an LDS array [4 * 2^D][blockDim] that parks the corner
// velocities between the speed pass and the Hamiltonian pass.
// PRE: keep the per-axis velocity numerators w, z in registers (fewer
// flops) or recompute them at every corner (fewer live registers).
// SC: also the rhs of a scalar carried by the same flux (the disorder field
// of a DisorderState: H_s = sum_a v_a s_a, no subtraction or climb term,
// CentralUpwindHJBetaPDynamics3D.py:186-188); usc[a][s] its derivatives in
// the units of u.
// CS: climb suppression decided at compile time (1: Lambda != 0, 0: Lambda
// == 0) or, for -1, from Lambda at run time; a run-time flag is applied as
// selects on every climb term.
// JY: form sum_c w_c J_c from the Y sums after the second pass instead of
// parking J per corner (vbuf then holds 3 * 2^D rows, not 4 * 2^D).  With
// J_c = jfac v(c).q(c), q_l(c) = [l<d] tr(u_l^{s_l}) - sum_a u_a^{s_a}[l,a],
//   sum_c w_c J_c = jfac sum_l ([l<d] sum_s tr(u_l^s) Y_l^s[l]
//                               - sum_a sum_s u_a^s[l,a] Y_a^s[l]).
// Glide + climb only (the woJzz climb direction is not of this form).
template <typename T, int D, int DYN, bool PRE, bool SC = false, int CS = -1, bool JY = false>
__device__ __forceinline__ void cell_flux(const T (&u)[D][2][9], const T (&sg)[9], T Lambda, T (&rhs)[9],
                                          T& speed, T* __restrict__ vbuf, int vstride, T nscale,
                                          const T (*usc)[2] = nullptr, T* rhs_sc = nullptr) {
  constexpr int NCOR = 1 << D;
  static_assert(!JY || DYN == DYN_GLIDE_CLIMB || DYN == DYN_GLIDE_ONLY, "JY: glide + climb form only");
  constexpr int VR = JY ? 3 : 4;  // vbuf rows per corner
  const T ME = T(kME);
  const T MEU = ME / (nscale * nscale);  // ME in the units of u^2
  const T third = T(1.0 / 3.0);
  // climb suppression: glide + climb uses Lambda tr(sigma)/3 and q; the woJzz
  // variant (2D) Lambda tr(sigma)/2 and w' (CentralUpwindHJBetaPDynamics_woJzz.py:71-75)
  const T st = (DYN == DYN_GLIDE_CLIMB) ? Lambda * (sg[0] + sg[4] + sg[8]) * third
               : (DYN == DYN_WOJZZ)    ? (sg[0] + sg[4] + sg[8]) * T(0.5) * Lambda
                                       : T(0);
  const T jfac = (DYN == DYN_WOJZZ) ? -Lambda * T(0.5) : -Lambda * third;
  const bool climb_sup =
      (DYN == DYN_GLIDE_CLIMB || DYN == DYN_WOJZZ) && (CS < 0 ? (Lambda != T(0)) : (CS == 1));
  // FOLD (glide + climb with climb suppression compiled in, J from the Y
  // sums): the corner velocity numerator
  //   v0_l = sum_a w_a[l] - [l<d] z_l + st ([l<d] tr_l - sum_a u_a[l,a])
  // is linear in per-axis quantities, so each axis / sign carries
  //   w''_a[l] = w_a[l] - st u_a[l,a] - [l==a] (z_a - st tr_a)
  // and a corner only adds three of them.
  constexpr bool FOLD = PRE && JY && !SC && DYN == DYN_GLIDE_CLIMB && CS == 1;
  // per axis / sign quantities
  T w[D][2][3], zz[D][2], tr[D][2], P[D][2];
#pragma unroll
  for (int a = 0; a < D; ++a) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const T* ua = u[a][s];
      if constexpr (PRE) {
#pragma unroll
        for (int l = 0; l < 3; ++l)
          w[a][s][l] = ua[3 * l] * sg[3 * a] + ua[3 * l + 1] * sg[3 * a + 1] + ua[3 * l + 2] * sg[3 * a + 2];
        T z = T(0);
#pragma unroll
        for (int c = 0; c < 9; ++c) z += ua[c] * sg[c];
        zz[a][s] = z;
      }
      tr[a][s] = ua[0] + ua[4] + ua[8];
      if constexpr (FOLD) {
#pragma unroll
        for (int l = 0; l < 3; ++l) w[a][s][l] = fma(-st, ua[3 * l + a], w[a][s][l]);
        w[a][s][a] -= fma(-st, tr[a][s], zz[a][s]);
      }
      T p = T(0);
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        if (m == a) continue;
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) p += ua[3 * m + jj] * ua[3 * m + jj];
      }
      P[a][s] = p;
    }
  }
  // the cross terms of |rho|^2 per pair of axes and their signs (each
  // corner's is the sum of three of them, 3D; one, 2D):
  //   X01[s0][s1] = sum_j u_0^s0[1,j] u_1^s1[0,j],  X12[s1][s2] = sum_j u_1^s1[2,j] u_2^s2[1,j],
  //   X20[s2][s0] = sum_j u_2^s2[0,j] u_0^s0[2,j]
  T X01[2][2], X12[2][2], X20[2][2];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      T a = T(0), b = T(0), e = T(0);
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) {
        a += u[0][s][3 + jj] * u[1][r][jj];
        if constexpr (D == 3) {
          b += u[1][s][6 + jj] * u[2][r][3 + jj];
          e += u[2][s][jj] * u[0][r][6 + jj];
        }
      }
      X01[s][r] = a;
      X12[s][r] = b;
      X20[s][r] = e;
    }
  // running max / min of the corner speeds, started at 0: the one-sided
  // speeds are max(0, max v) and -min(0, min v) (Max5/Max9 with 0)
  T amax[3], amin[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    amax[a] = T(0);
    amin[a] = T(0);
  }
#pragma unroll
  for (int c = 0; c < NCOR; ++c) {
    int sgn[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < D; ++a) sgn[a] = (c >> (D - 1 - a)) & 1;
    T R[3], v0[3], q[3];
    if constexpr (FOLD) {
#pragma unroll
      for (int l = 0; l < 3; ++l) v0[l] = (w[0][sgn[0]][l] + w[1][sgn[1]][l]) + w[2][sgn[2]][l];
    }
#pragma unroll
    for (int l = 0; l < 3 && !FOLD; ++l) {
      T r = T(0), vv = T(0);
#pragma unroll
      for (int a = 0; a < D; ++a) {
        r += u[a][sgn[a]][3 * l + a];
        if constexpr (PRE) {
          vv += w[a][sgn[a]][l];
        } else {
#pragma unroll
          for (int nn = 0; nn < 3; ++nn) vv += u[a][sgn[a]][3 * l + nn] * sg[3 * a + nn];
        }
      }
      R[l] = r;
      v0[l] = vv;
    }
#pragma unroll
    for (int a = 0; a < D && !FOLD; ++a) {
      if constexpr (PRE) {
        v0[a] -= zz[a][sgn[a]];
      } else {
        T z = T(0);
#pragma unroll
        for (int cc = 0; cc < 9; ++cc) z += u[a][sgn[a]][cc] * sg[cc];
        v0[a] -= z;
      }
    }
#pragma unroll
    for (int a = 0; a < 3 && !FOLD; ++a) {
      // q_a = tr(u_a) - sum_b u_b[a, b]; tr(u_a) = 0 for a >= d
      const T tra = (a < D) ? tr[a % D][sgn[a]] : T(0);
      q[a] = tra - R[a];
    }
    // woJzz climb direction w' = (ux_yy - uy_xy, uy_xx - ux_yx, -(ux_zx + uy_zy))
    T wz[3] = {T(0), T(0), T(0)};
    if (DYN == DYN_WOJZZ) {
      wz[0] = u[0][sgn[0]][4] - u[D - 1][sgn[D - 1]][1];
      wz[1] = u[D - 1][sgn[D - 1]][0] - u[0][sgn[0]][3];
      wz[2] = -(u[0][sgn[0]][6] + u[D - 1][sgn[D - 1]][7]);
    }
    if (DYN == DYN_GLIDE_CLIMB) {
      if (climb_sup && !FOLD) {
#pragma unroll
        for (int a = 0; a < D; ++a) v0[a] += st * tr[a][sgn[a]];
#pragma unroll
        for (int l = 0; l < 3; ++l) v0[l] -= st * R[l];
      }
    } else if (DYN == DYN_WOJZZ) {
      if (climb_sup) {
#pragma unroll
        for (int l = 0; l < 3; ++l) v0[l] += st * wz[l];
      }
    } else {
      // lvp projection (CentralUpwindHJBetaPGlideOnlyDynamics.py:170-195)
      T rr = MEU;
#pragma unroll
      for (int a = 0; a < D; ++a) rr += tr[a][sgn[a]] * tr[a][sgn[a]] - T(2) * tr[a][sgn[a]] * R[a];
#pragma unroll
      for (int l = 0; l < 3; ++l) rr += R[l] * R[l];
      const T rv = MEU + v0[0] * q[0] + v0[1] * q[1] + v0[2] * q[2];
      const T f = rv / rr;
#pragma unroll
      for (int a = 0; a < 3; ++a) v0[a] += q[a] * f;
    }
    // |rho|^2 (BetaP3D_RhoMod, CentralUpwindHJBetaPDynamics3D.py:45-58)
    T r2 = MEU;
#pragma unroll
    for (int a = 0; a < D; ++a) r2 += P[a][sgn[a]];
    T cross = X01[sgn[0]][sgn[1]];
    if constexpr (D == 3) cross += X12[sgn[1]][sgn[2]] + X20[sgn[2]][sgn[0]];
    r2 -= T(2) * cross;
    const T rinv = rsqrt_nr(r2);
    T vv[3];
#pragma unroll
    for (int l = 0; l < 3; ++l) vv[l] = v0[l] * rinv;
    const T J = (JY || !climb_sup)      ? T(0)
                : (DYN == DYN_WOJZZ) ? jfac * (vv[0] * wz[0] + vv[1] * wz[1] + vv[2] * wz[2])
                                     : jfac * (vv[0] * q[0] + vv[1] * q[1] + vv[2] * q[2]);
#pragma unroll
    for (int l = 0; l < 3; ++l) vbuf[(VR * c + l) * vstride] = vv[l];
    if constexpr (!JY) vbuf[(4 * c + 3) * vstride] = J;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      amax[a] = fmax(amax[a], vv[a]);
      amin[a] = fmin(amin[a], vv[a]);
    }
  }
  // local one-sided speeds (Max5/Max9 with 0, CentralUpwindHJ3D.py:114-241)
  T ap[3], am[3], at[3];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    ap[a] = amax[a];
    am[a] = -amin[a];
    at[a] = ap[a] + am[a] + ME;
  }
  T Y[D][2][3], Jsum = T(0);
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int l = 0; l < 3; ++l) Y[a][s][l] = T(0);
#pragma unroll
  for (int c = 0; c < NCOR; ++c) {
    T wc = T(1);
    int sgn[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < D; ++a) {
      sgn[a] = (c >> (D - 1 - a)) & 1;
      wc *= sgn[a] == 0 ? am[a] : ap[a];
    }
    T vv[3];
#pragma unroll
    for (int l = 0; l < 3; ++l) vv[l] = vbuf[(VR * c + l) * vstride];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int l = 0; l < 3; ++l) Y[a][sgn[a]][l] += wc * vv[l];
    if constexpr (!JY) {
      if (climb_sup) Jsum += wc * vbuf[(4 * c + 3) * vstride];
    }
  }
  if constexpr (JY) {
    if (climb_sup) {
      T js = T(0);
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        if (l < D) {
#pragma unroll
          for (int s = 0; s < 2; ++s) js = fma(tr[l % D][s], Y[l % D][s][l], js);
        }
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
          for (int s = 0; s < 2; ++s) js = fma(-u[a][s][3 * l + a], Y[a][s][l], js);
      }
      Jsum = jfac * js;
    }
  }
  T inv = T(1);
#pragma unroll
  for (int a = 0; a < D; ++a) inv *= at[a];
  inv = T(1) / inv;
  // 1 / at[a] from the one division: inv times the other axes' at
  T iat[D];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    T o = inv;
#pragma unroll
    for (int b = 0; b < D; ++b)
      if (b != a) o *= at[b];
    iat[a] = o;
  }
  speed = T(0);
#pragma unroll
  for (int a = 0; a < D; ++a) speed += at[a];
  T tw[D];
#pragma unroll
  for (int a = 0; a < D; ++a) tw[a] = am[a] * ap[a] * iat[a];
  if constexpr (FOLD) {
    // rhs = -inv h + sum_a (u_a^+ - u_a^-) tw_a with h linear in the same
    // derivatives: one coefficient per derivative of the cell's component,
    //   rhs = sum_{a,s} u_a^s[cc] k_a^s + sum_{s,l} u_i^s[l,jj] inv Y_i^s[l] - [i==jj] inv J,
    //   k_a^+ = tw_a - inv Y_a^+[a],  k_a^- = -tw_a - inv Y_a^-[a]
    T k[D][2], Yi[D][2][3];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      k[a][0] = fma(-inv, Y[a][0][a], tw[a]);
      k[a][1] = fma(-inv, Y[a][1][a], -tw[a]);
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int l = 0; l < 3; ++l) Yi[a][s][l] = inv * Y[a][s][l];
    }
    const T Jn = -inv * Jsum;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) {
        const int cc = 3 * i + jj;
        T r = (i == jj) ? Jn : T(0);
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
          for (int s = 0; s < 2; ++s) r = fma(u[a][s][cc], k[a][s], r);
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
          for (int l = 0; l < 3; ++l) r = fma(u[i][s][3 * l + jj], Yi[i][s][l], r);
        rhs[cc] = r * nscale;
      }
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int jj = 0; jj < 3; ++jj) {
      const int cc = 3 * i + jj;
      // the climb term joins H_xx, H_yy, H_zz (woJzz: H_xx, H_yy only)
      T h = (i == jj && (DYN != DYN_WOJZZ || i < 2)) ? Jsum : T(0);
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int s = 0; s < 2; ++s) h += u[a][s][cc] * Y[a][s][a];
      if (i < D) {
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
          for (int l = 0; l < 3; ++l) h -= u[i % D][s][3 * l + jj] * Y[i % D][s][l];
      }
      T r = -h * inv;
#pragma unroll
      for (int a = 0; a < D; ++a) r += (u[a][0][cc] - u[a][1][cc]) * tw[a];
      rhs[cc] = (DYN == DYN_WOJZZ && cc == 8) ? T(0) : r * nscale;  // woJzz holds beta_zz fixed
    }
  }
  if constexpr (SC) {
    T h = T(0);
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int s = 0; s < 2; ++s) h += usc[a][s] * Y[a][s][a];
    T r = -h * inv;
#pragma unroll
    for (int a = 0; a < D; ++a) r += (usc[a][0] - usc[a][1]) * tw[a];
    *rhs_sc = r * nscale;
  }
}

// ---------------------------------------------------------------------------
// The CUDA executable's numerics (plasticity_kernel.cu centralHJ :118-697 with
// the defaults of plasticity.h: LLF, theta = 1, DOUBLE): the same physics as
// cell_flux, with one local Lax-Friedrichs speed per cell, velocities divided
// by |rho| + sqrt(ME), and its own form of the limited derivatives.

// minmod3 of plasticity_kernel.cu:41-51: the smallest if all three are
// positive, the largest if all are negative, else 0.
template <typename T>
__host__ __device__ __forceinline__ T minmod3_cmp(T a, T b, T c) {
  if (a > T(0) && b > T(0) && c > T(0)) return fmin(fmin(a, b), c);
  if (a < T(0) && b < T(0) && c < T(0)) return fmax(fmax(a, b), c);
  return T(0);
}

// findDerivatives, theta = 1 (plasticity_kernel.cu:53-90): the one-sided
// limited derivatives from f[i-2..i+2], in the grid's units (times N).
template <typename T>
__host__ __device__ __forceinline__ void llf_derivs(const T f[5], T N, T& up, T& um) {
  const T vll = f[0], vl = f[1], v = f[2], vr = f[3], vrr = f[4];
  const T d2 = (vr + vl) - T(2) * v;
  const T mp = minmod3_cmp<T>((vrr + v) - T(2) * vr, d2, T(0.5) * (((vrr - vr) - v) + vl));
  const T mm = minmod3_cmp<T>((vll + v) - T(2) * vl, d2, T(0.5) * (((vr - v) - vl) + vll));
  up = ((vr - v) - T(0.5) * mp) * N;
  um = ((v - vl) + T(0.5) * mm) * N;
}

// Per-cell LLF flux.  u[a][s][c]: derivative along axis a, s = 0 ('+') or 1
// ('-'), in the grid's units.  Corner k takes the x sign k & 1, the y sign
// (k >> 1) & 1 and the z sign k >> 2.  With the single speed a = max over
// corners and axes of |v|, at = 2 a + ME:
//   rhs = -(a^d / at^d) sum_k H_k + sum_axes (u+ - u-) a^2 / at,
//   speed = d at.
// H_k is linear in the corner velocity v_k, so sum_k H_k is formed from
// Y[a][s][l] = sum over corners with sign s on axis a of v_k[l]
// (sum_k u_k[a][c] v_k[l] = sum_s u[a][s][c] Y[a][s][l]).  vbuf parks the
// 3 2^D corner velocities (stride vstride).
template <typename T, int D, int DYN>
__host__ __device__ __forceinline__ void cell_flux_llf(const T (&u)[D][2][9], const T (&sg)[9], T Lambda, T (&rhs)[9],
                                              T& speed, T* __restrict__ vbuf, int vstride) {
  constexpr int NCOR = 1 << D;
  const T ME = T(kME);
  const T sqME = sqrt(ME);
  // climb suppression lambda tr(sigma) / 3 (:297-300, :336-345, :371-386)
  const T st = ((sg[0] + sg[4]) + sg[8]) / T(3) * Lambda;
  T amax = T(0);
#pragma unroll
  for (int k = 0; k < NCOR; ++k) {
    const T* U0 = u[0][k & 1];
    const T* U1 = u[1][(k >> 1) & 1];
    const T* U2 = u[D - 1][(k >> 2) & 1];  // read in 3D only
    // velocity numerator (:289-456); the 3D v_z subtracts uz[y,n] sigma[z,n]
    // as the executable does
    T vx = U1[0] * sg[3] - U0[3] * sg[3] - U0[6] * sg[6] + U1[1] * sg[4] - U0[4] * sg[4] - U0[7] * sg[7] +
           U1[2] * sg[5] - U0[5] * sg[5] - U0[8] * sg[8];
    T vy = U0[3] * sg[0] - U1[0] * sg[0] - U1[6] * sg[6] + U0[4] * sg[1] - U1[1] * sg[1] - U1[7] * sg[7] +
           U0[5] * sg[2] - U1[2] * sg[2] - U1[8] * sg[8];
    T vz = U0[6] * sg[0] + U1[6] * sg[3] + U0[7] * sg[1] + U1[7] * sg[4] + U0[8] * sg[2] + U1[8] * sg[5];
    if constexpr (D == 3) {
      vx += U2[0] * sg[6] + U2[1] * sg[7] + U2[2] * sg[8];
      vy += U2[3] * sg[6] + U2[4] * sg[7] + U2[5] * sg[8];
      vz -= U2[0] * sg[0] + U2[1] * sg[1] + U2[2] * sg[2] + U2[3] * sg[6] + U2[4] * sg[7] + U2[5] * sg[8];
    }
    vx += st * ((U0[0] + U0[4]) + U0[8]);
    vy += st * ((U1[0] + U1[4]) + U1[8]);
    vx -= st * (U0[0] + U1[1]);
    vy -= st * (U0[3] + U1[4]);
    vz -= st * (U0[6] + U1[7]);
    if constexpr (D == 3) {
      vz += st * ((U2[0] + U2[4]) + U2[8]);
      vx -= st * U2[2];
      vy -= st * U2[5];
      vz -= st * U2[8];
    }
    if constexpr (DYN == DYN_GLIDE_ONLY) {
      // lvp projection (NEWGLIDEONLY, :389-452)
      const T txr = (U0[0] + U0[4]) + U0[8], tyr = (U1[0] + U1[4]) + U1[8];
      T rx = U0[0] + U1[1], ry = U0[3] + U1[4], rz = U0[6] + U1[7];
      T qx = (U0[4] + U0[8]) - U1[1], qy = (U1[0] + U1[8]) - U0[3], qz = -rz;
      T rr = ME + txr * txr + tyr * tyr;
      if constexpr (D == 3) {
        const T tzr = (U2[0] + U2[4]) + U2[8];
        rx += U2[2];
        ry += U2[5];
        rz += U2[8];
        qx -= U2[2];
        qy -= U2[5];
        qz = (U2[0] + U2[4]) - U0[6] - U1[7];
        rr += tzr * tzr - T(2) * tzr * rz;
      }
      rr += rx * rx + ry * ry + rz * rz - T(2) * txr * rx - T(2) * tyr * ry;
      const T rv = (ME + vx * qx + vy * qy + vz * qz) / rr;
      vx += qx * rv;
      vy += qy * rv;
      vz += qz * rv;
    }
    // |rho| without ME, negative values clamped (:232-287)
    T r = T(0);
#pragma unroll
    for (int c = 3; c < 9; ++c) r += U0[c] * U0[c];
#pragma unroll
    for (int c = 0; c < 9; ++c)
      if (c < 3 || c >= 6) r += U1[c] * U1[c];
#pragma unroll
    for (int j = 0; j < 3; ++j) r -= T(2) * U0[3 + j] * U1[j];
    if constexpr (D == 3) {
#pragma unroll
      for (int c = 0; c < 6; ++c) r += U2[c] * U2[c];
#pragma unroll
      for (int j = 0; j < 3; ++j) r -= T(2) * U0[6 + j] * U2[j] + T(2) * U1[6 + j] * U2[3 + j];
    }
    const T irho = T(1) / (sqrt(fmax(r, T(0))) + sqME);
    const T v[3] = {vx * irho, vy * irho, vz * irho};
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      vbuf[(3 * k + l) * vstride] = v[l];
      amax = fmax(amax, fabs(v[l]));
    }
  }
  T Y[D][2][3];
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int l = 0; l < 3; ++l) Y[a][s][l] = T(0);
#pragma unroll
  for (int k = 0; k < NCOR; ++k) {
    const int sgn[3] = {k & 1, (k >> 1) & 1, (k >> 2) & 1};
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const T v = vbuf[(3 * k + l) * vstride];
#pragma unroll
      for (int a = 0; a < D; ++a) Y[a][sgn[a]][l] += v;
    }
  }
  // Z(a, c, l) = sum_k u_k[a][c] v_k[l]
  auto Z = [&](int a, int c, int l) -> T { return u[a][0][c] * Y[a][0][l] + u[a][1][c] * Y[a][1][l]; };
  T Jsum = T(0);
  if constexpr (DYN == DYN_GLIDE_CLIMB) {
    // sum over corners of the climb rate q (:468-545)
    T Q = Z(0, 4, 0) + Z(0, 8, 0) - Z(1, 1, 0) + Z(1, 0, 1) + Z(1, 8, 1) - Z(0, 3, 1) - Z(0, 6, 2) - Z(1, 7, 2);
    if constexpr (D == 3) Q += -Z(2, 2, 0) - Z(2, 5, 1) + Z(2, 0, 2) + Z(2, 4, 2);
    Jsum = Q / T(3) * Lambda;
  }
  const T at = amax + amax + ME;
  T aD = amax * amax, atD = at * at;
  if constexpr (D == 3) {
    aD *= amax;
    atD *= at;
  }
  const T wH = aD / atD;
  const T wd = amax * amax / at;
  speed = at * T(D);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int cc = 3 * i + j;
      T h = (i == j) ? -Jsum : T(0);
#pragma unroll
      for (int a = 0; a < D; ++a) h += Z(a, cc, a);
      if (i < D) {
#pragma unroll
        for (int l = 0; l < 3; ++l) h -= Z(i % D, 3 * l + j, l);
      }
      T r = -h * wH;
#pragma unroll
      for (int a = 0; a < D; ++a) r += (u[a][0][cc] - u[a][1][cc]) * wd;
      rhs[cc] = r;
    }
  }
}

template <typename T>
__device__ __forceinline__ T block_max(T v, T* red) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int nt = blockDim.x * blockDim.y;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  const int wid = tid >> 6, lane = tid & 63;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  T m = -INFINITY;
  if (tid == 0) {
    for (int k = 0; k < (nt + 63) / 64; ++k) m = fmax(m, red[k]);
  }
  return m;
}

// Block = TC x TR cells of the two fastest axes (TC contiguous); 3D blocks
// sit at one x plane.  The block loads, in one batch, the plane tile of all
// nine components with a 2-cell halo (into LDS) and, in 3D, each thread's
// four slow-axis (x) neighbours (into registers, coalesced along the
// contiguous axis) -- one global-memory round trip per block.
template <int TC, int TR>
struct FluxTile {
  static constexpr int PW = TC + 4, PH = TR + 4;  // padded tile
  static constexpr int TILE = PW * PH;
  static constexpr int NT = TC * TR;               // threads
  static constexpr int FILL = (TILE + NT - 1) / NT; // fill loads per thread
};

// LLF: the CUDA executable's flux (llf_derivs, cell_flux_llf) in place of
// the central-upwind one.
// SC: the advected disorder scalar of FluxJob (Ad/Ud/outd) rides along.
template <typename T, int D, int DYN, int FTC, int FTR, bool PRE, bool LLF = false, bool SC = false>
__global__ void __launch_bounds__(FTC* FTR) k_flux(FluxJob job) {
  using FT = FluxTile<FTC, FTR>;
  constexpr int FPW = FT::PW, FTILE = FT::TILE, FNT = FT::NT, FFILL = FT::FILL;
  // the job's field pointers carry T data (float for fp32 plans)
  const T* __restrict__ A = reinterpret_cast<const T*>(job.A);
  const T* Gp = reinterpret_cast<const T*>(job.G);
  const T* Up = reinterpret_cast<const T*>(job.U);
  T* outp = reinterpret_cast<T*>(job.out);
  T* speedp = reinterpret_cast<T*>(job.speed);
  double* partialp = job.partial;
  if (blockIdx.y) {  // batched 2D: instance blockIdx.y
    const long long i = blockIdx.y;
    A += i * job.fstride;
    if (Up) Up += i * job.fstride;
    outp += i * job.fstride;
    Gp += i * job.gstride;
    if (partialp) partialp += i * job.pstride;
    if (speedp) speedp += i * (job.fstride / 9);
  }
  if (job.clk) job.clk += blockIdx.y;
  if (job.clk && job.clk->done) return;
  constexpr int NCOMP = SC ? 10 : 9;
  __shared__ T tile[NCOMP][FTILE];
  __shared__ T vsh[4 * (1 << D)][FNT];
  __shared__ T red[(FNT + 63) / 64];
  // 3D: per-wave LDS landing zone for G (planes 0-5) and U (planes 6-15),
  // filled by asynchronous global->LDS loads issued at kernel start
  __shared__ T ug[(D == 3 && sizeof(T) == 8) ? FNT / 64 : 1][16][64];
  const int n = job.n;
  const long long cells = job.cells;
  const int tc = threadIdx.x, trr = threadIdx.y;
  const int tid = trr * FTC + tc;
  // XCD-aware tile order.  Workgroups are dispatched round-robin over the
  // 8 XCDs (hardware block b runs on XCD b % 8); logical tile
  // L = (b % 8) * (nb / 8) + b / 8 gives every XCD one contiguous range of
  // L, and L runs x-fastest, so an XCD's CUs work on consecutive x planes
  // of the same (y, z) column at once and the slow-axis neighbour planes
  // are served from that XCD's L2.
  const int nct = n / FTC, nrt = n / FTR;
  const int nb = (D == 3 ? n : 1) * nct * nrt;
  const int bl = blockIdx.x;
  const int L = (nb % 8 == 0) ? (bl % 8) * (nb / 8) + bl / 8 : bl;
  int px, tyz;
  if constexpr (D == 3) {
    px = L % n;
    tyz = L / n;
  } else {
    px = 0;
    tyz = L;
  }
  const int c0 = (tyz % nct) * FTC, r0 = (tyz / nct) * FTR;
  // strides: col (contiguous) axis = last, row axis = second last
  const long long srow = n, splane = (long long)n * n;
  const long long pbase = (D == 3) ? (long long)px * splane : 0;
  const int mycol = c0 + tc, myrow = r0 + trr;
  const long long idx = pbase + (long long)myrow * srow + mycol;
  const T dt = job.clk ? T(job.clk->dt) : T(0);

  // One batch of loads for every component (a single memory round trip):
  // the plane tile with its 2-cell halo (periodic wrap) and, in 3D, the
  // four slow-axis neighbours of this thread's cell.
  long long foff[FFILL];
  bool fval[FFILL];
#pragma unroll
  for (int k = 0; k < FFILL; ++k) {
    const int e = tid + k * FNT;
    fval[k] = e < FTILE;
    const int er = e / FPW, ec = e % FPW;
    int gr = r0 + er - 2, gc = c0 + ec - 2;
    gr = gr < 0 ? gr + n : (gr >= n ? gr - n : gr);
    gc = gc < 0 ? gc + n : (gc >= n ? gc - n : gc);
    foff[k] = pbase + (long long)gr * srow + gc;
  }
  constexpr int NX = (D == 3) ? 4 : 1;
  long long xoff[NX];
#pragma unroll
  for (int k = 0; k < NX; ++k) {
    int p = px + (k < 2 ? k - 2 : k - 1);
    p = p < 0 ? p + n : (p >= n ? p - n : p);
    xoff[k] = (long long)p * splane + (long long)myrow * srow + mycol;
  }
  // 3D fp64: G and U land in LDS by asynchronous DMA (16 B = two cells per lane)
  constexpr bool DMA = (D == 3) && (sizeof(T) == 8);
  if constexpr (DMA) {
    // Each wave's 64 cells are contiguous in memory (TC = 64, or TC = n);
    // every lane moves 16 B (two cells) straight into LDS, no VGPRs used.
    // lane l moves cells (2 l', 2 l' + 1), l' = l mod 32, of the wave's 64
    // cells (tid order) for plane 2k + l / 32; rows of FTC cells, FTC even
    const int wv = tid >> 6, lane = tid & 63;
    const int half = lane >> 5, e = 64 * wv + 2 * (lane & 31);
    const long long cellp = pbase + (long long)(r0 + e / FTC) * srow + c0 + e % FTC;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const T* src = Gp + (2 * k + half) * cells + cellp;
      __builtin_amdgcn_global_load_lds((const void*)src, (__attribute__((address_space(3))) void*)&ug[wv][2 * k][0], 16,
                                       0, 0);
    }
    if (job.mode != 0) {
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        int c = 2 * k + half;
        c = c < 9 ? c : 8;  // the spare 16th plane receives a copy of plane 8
        const T* src = Up + c * cells + cellp;
        __builtin_amdgcn_global_load_lds((const void*)src,
                                         (__attribute__((address_space(3))) void*)&ug[wv][6 + 2 * k][0], 16, 0, 0);
      }
    }
  }
  T tf[NCOMP][FFILL];
  T xn[NCOMP][NX];
#pragma unroll
  for (int c = 0; c < NCOMP; ++c) {
    // component 9: the advected scalar's plane
    const T* Ac = (SC && c == 9) ? reinterpret_cast<const T*>(job.Ad) : A + c * cells;
#pragma unroll
    for (int k = 0; k < FFILL; ++k) tf[c][k] = fval[k] ? Ac[foff[k]] : T(0);
    if constexpr (D == 3) {
#pragma unroll
      for (int k = 0; k < NX; ++k) xn[c][k] = Ac[xoff[k]];
    }
  }
#pragma unroll
  for (int c = 0; c < NCOMP; ++c)
#pragma unroll
    for (int k = 0; k < FFILL; ++k)
      if (fval[k]) tile[c][tid + k * FNT] = tf[c][k];
  __syncthreads();

  T u[D][2][9];
  T fc[9];
  const int lr = trr + 2, lc = tc + 2;
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    const T* tb = tile[c];
    const T ctr = tb[lr * FPW + lc];
    fc[c] = ctr;
    {  // contiguous axis (last)
      T f[5] = {tb[lr * FPW + lc - 2], tb[lr * FPW + lc - 1], ctr, tb[lr * FPW + lc + 1], tb[lr * FPW + lc + 2]};
      if constexpr (LLF) llf_derivs<T>(f, T(n), u[D - 1][0][c], u[D - 1][1][c]);
      else limited_derivs<T>(f, u[D - 1][0][c], u[D - 1][1][c]);
    }
    {  // row axis (second last)
      T f[5] = {tb[(lr - 2) * FPW + lc], tb[(lr - 1) * FPW + lc], ctr, tb[(lr + 1) * FPW + lc],
                tb[(lr + 2) * FPW + lc]};
      if constexpr (LLF) llf_derivs<T>(f, T(n), u[D - 2][0][c], u[D - 2][1][c]);
      else limited_derivs<T>(f, u[D - 2][0][c], u[D - 2][1][c]);
    }
    if constexpr (D == 3) {
      T f[5] = {xn[c][0], xn[c][1], ctr, xn[c][2], xn[c][3 % NX]};
      if constexpr (LLF) llf_derivs<T>(f, T(n), u[0][0][c], u[0][1][c]);
      else limited_derivs<T>(f, u[0][0][c], u[0][1][c]);
    }
  }
  T usc[D][2], fsc = T(0);
  if constexpr (SC) {
    const T* tb = tile[9];
    fsc = tb[lr * FPW + lc];
    T f[5] = {tb[lr * FPW + lc - 2], tb[lr * FPW + lc - 1], fsc, tb[lr * FPW + lc + 1], tb[lr * FPW + lc + 2]};
    limited_derivs<T>(f, usc[D - 1][0], usc[D - 1][1]);
    T g[5] = {tb[(lr - 2) * FPW + lc], tb[(lr - 1) * FPW + lc], fsc, tb[(lr + 1) * FPW + lc], tb[(lr + 2) * FPW + lc]};
    limited_derivs<T>(g, usc[D - 2][0], usc[D - 2][1]);
    if constexpr (D == 3) {
      T h[5] = {xn[9][0], xn[9][1], fsc, xn[9][2], xn[9][3 % NX]};
      limited_derivs<T>(h, usc[0][0], usc[0][1]);
    }
  }
  // the asynchronous G/U loads must have landed before LDS is read
  if constexpr (DMA) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  auto Gv = [&](int c) -> T {
    if constexpr (DMA) return ug[tid >> 6][c][tid & 63];
    else return Gp[c * cells + idx];
  };
  auto Uv = [&](int c) -> T {
    if constexpr (DMA) return ug[tid >> 6][6 + c][tid & 63];
    else return Up[c * cells + idx];
  };
  // sigma = G - lam tr(beta) d - mu (beta + beta^T) at the cell
  T sg[9];
  {
    const T tr = fc[0] + fc[4] + fc[8];
    const T g0 = Gv(0), g1 = Gv(1), g2 = Gv(2);
    const T g3 = Gv(3), g4 = Gv(4), g5 = Gv(5);
    const T mu = T(job.mu), lam = T(job.lam);
    sg[0] = g0 - lam * tr - mu * (fc[0] + fc[0]);
    sg[4] = g1 - lam * tr - mu * (fc[4] + fc[4]);
    sg[8] = g2 - lam * tr - mu * (fc[8] + fc[8]);
    sg[1] = sg[3] = g3 - mu * (fc[1] + fc[3]);
    sg[2] = sg[6] = g4 - mu * (fc[2] + fc[6]);
    sg[5] = sg[7] = g5 - mu * (fc[5] + fc[7]);
  }
  if constexpr (sizeof(T) == 8) {
    // the disorder stress (non-symmetric), DisorderState.py:26-29
    if (job.S9) {
#pragma unroll
      for (int c = 0; c < 9; ++c) sg[c] += job.S9[c * cells + idx];
    }
  }
  T rhs[9], speed, rhs_sc = T(0);
  if constexpr (LLF) cell_flux_llf<T, D, DYN>(u, sg, T(job.Lambda), rhs, speed, &vsh[0][tid], FNT);
  else cell_flux<T, D, DYN, PRE, SC>(u, sg, T(job.Lambda), rhs, speed, &vsh[0][tid], FNT, T(n), usc, &rhs_sc);
  if constexpr (sizeof(T) == 8) {
    if (job.src) {
#pragma unroll
      for (int c = 0; c < 9; ++c) rhs[c] += T(job.src_rate) * job.src[c * cells + idx];
    }
  }
  if constexpr (SC) {
    double* od = job.outd;
    if (job.mode == 0) od[idx] = (double)rhs_sc;
    else od[idx] = rk_combine<double>(job.cu, job.Ud[idx], job.cf, (double)fsc, (double)dt, (double)rhs_sc);
  }
  if (job.mode == 0) {
#pragma unroll
    for (int c = 0; c < 9; ++c) outp[c * cells + idx] = rhs[c];
  } else {
    const T cu = T(job.cu), cf = T(job.cf);
    T ov[9], uv[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      uv[c] = Uv(c);
      ov[c] = rk_combine<T>(cu, uv[c], cf, fc[c], dt, rhs[c]);
      outp[c * cells + idx] = ov[c];
    }
    if constexpr (sizeof(T) == 8) {
      // vacancy coupling after the last stage: c += beta (tr F2 - tr u)
      // (OperatorSplittingTVDRK_FieldMover, FieldMover.py:351,375-376,
      // VacancyDynamics.py:37-39)
      if (job.vac) job.vac[idx] += job.vbeta * ((ov[0] + ov[4] + ov[8]) - (uv[0] + uv[4] + uv[8]));
    }
  }
  if (speedp) speedp[idx] = speed;
  if (partialp) {
    // a non-finite speed or rhs must reach the CFL reduction (fmax drops NaN,
    // also inside the speeds, so the rhs is checked too)
    bool fin = isfinite(speed) && isfinite(rhs_sc);
#pragma unroll
    for (int c = 0; c < 9; ++c) fin = fin && isfinite(rhs[c]);
    T m = block_max<T>(fin ? speed : T(INFINITY), red);
    if (tid == 0) partialp[blockIdx.x] = (double)m;
  }
}

// ---------------------------------------------------------------------------
// Persistent multi-instance 2D flux (k_flux_multi): the per-cell arithmetic
// of k_flux<double, 2, DYN, TC, TR, PRE = true> (same operands, same
// operations, so every output is bit-identical to it), with the work of
// every instance of every size group in one launch.  A fixed grid (per_cu
// blocks per CU) walks the flat item list (group, instance, tile) with a
// stride of the grid size; an instance whose clock is done is skipped.

// A work item's group entry (copied field by field with constant indices:
// a computed index into the kernel-argument table would move the whole
// table to scratch memory), instance and tile origin -- uniform values.
struct MultiItem {
  FluxGroup G;
  int tl, c0, r0;
  long long inst;
};

template <int TC, int TR>
__device__ __forceinline__ MultiItem multi_locate(const FluxMultiJob& job, int it, bool xperm) {
  // XCD-aware order: hardware block b runs on XCD b % 8 and, with a grid
  // that is a multiple of 8, takes the items it = b (mod 8), so logical
  // item (it % 8) (nitems / 8) + it / 8 gives every XCD one contiguous range
  // of tiles (neighbouring halos served from its L2)
  const int L = xperm ? (it % 8) * (job.nitems / 8) + it / 8 : it;
  MultiItem w;
  w.G = job.g[0];
#pragma unroll
  for (int k = 1; k < kFluxGroups; ++k) {
    if (k < job.ngroups && L >= job.g[k].item0) {
      w.G.n = job.g[k].n;
      w.G.ntiles = job.g[k].ntiles;
      w.G.item0 = job.g[k].item0;
      w.G.cells = job.g[k].cells;
      w.G.fstride = job.g[k].fstride;
      w.G.gstride = job.g[k].gstride;
      w.G.pstride = job.g[k].pstride;
      w.G.A = job.g[k].A;
      w.G.G = job.g[k].G;
      w.G.U = job.g[k].U;
      w.G.out = job.g[k].out;
      w.G.partial = job.g[k].partial;
      w.G.clk = job.g[k].clk;
    }
  }
  const int loc = L - w.G.item0;
  w.inst = loc / w.G.ntiles;
  w.tl = loc % w.G.ntiles;
  const int nct = w.G.n / TC;
  w.c0 = (w.tl % nct) * TC;
  w.r0 = (w.tl / nct) * TR;
  return w;
}

template <int DYN, int TC, int TR, bool COMB>
__global__ void __launch_bounds__(TC* TR, 2) k_flux_multi(FluxMultiJob job) {
  using FT = FluxTile<TC, TR>;
  constexpr int PW = FT::PW, TILE = FT::TILE, NT = FT::NT, FILL = FT::FILL;
  __shared__ double tile[9][TILE];
  __shared__ double vsh[16][NT];
  __shared__ double red[(NT + 63) / 64];
  const int tc = threadIdx.x, trr = threadIdx.y;
  const int tid = trr * TC + tc;
  const int lr = trr + 2, lc = tc + 2;
  const int nitems = job.nitems;
  const bool xperm = (nitems % 8) == 0 && (gridDim.x % 8) == 0;
  const double mu = job.mu, lam = job.lam;
  for (int it = blockIdx.x; it < nitems; it += (int)gridDim.x) {
    const MultiItem w = multi_locate<TC, TR>(job, it, xperm);
    const FluxGroup& G = w.G;
    const Clock* ck = G.clk + w.inst;
    if (ck->done) continue;  // uniform over the block
    const double dt = ck->dt;
    const int n = G.n;
    const long long cells = G.cells;
    const double* A = G.A + w.inst * G.fstride;
    const long long idx = (long long)(w.r0 + trr) * n + w.c0 + tc;
    // one batch of loads: the halo tile of all nine components, then this
    // cell's stress and U values
    double tf[9][FILL];
#pragma unroll
    for (int k = 0; k < FILL; ++k) {
      const int e = tid + k * NT;
      const bool ok = e < TILE;
      const int er = e / PW, ec = e % PW;
      int gr = w.r0 + er - 2, gc = w.c0 + ec - 2;
      gr = gr < 0 ? gr + n : (gr >= n ? gr - n : gr);
      gc = gc < 0 ? gc + n : (gc >= n ? gc - n : gc);
      const long long off = (long long)gr * n + gc;
#pragma unroll
      for (int c = 0; c < 9; ++c) tf[c][k] = ok ? A[c * cells + off] : 0.0;
    }
    const double* Gp = G.G + w.inst * G.gstride + idx;
    double gv[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) gv[c] = Gp[c * cells];
    // the previous item's reads of the tile are complete
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 9; ++c)
#pragma unroll
      for (int k = 0; k < FILL; ++k)
        if (tid + k * NT < TILE) tile[c][tid + k * NT] = tf[c][k];
    __syncthreads();
    double u[2][2][9], fc[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      const double* tb = tile[c];
      const double ctr = tb[lr * PW + lc];
      fc[c] = ctr;
      {  // contiguous axis
        double f[5] = {tb[lr * PW + lc - 2], tb[lr * PW + lc - 1], ctr, tb[lr * PW + lc + 1], tb[lr * PW + lc + 2]};
        limited_derivs<double>(f, u[1][0][c], u[1][1][c]);
      }
      {  // row axis
        double f[5] = {tb[(lr - 2) * PW + lc], tb[(lr - 1) * PW + lc], ctr, tb[(lr + 1) * PW + lc],
                       tb[(lr + 2) * PW + lc]};
        limited_derivs<double>(f, u[0][0][c], u[0][1][c]);
      }
    }
    double sg[9];
    {
      const double tr = fc[0] + fc[4] + fc[8];
      sg[0] = gv[0] - lam * tr - mu * (fc[0] + fc[0]);
      sg[4] = gv[1] - lam * tr - mu * (fc[4] + fc[4]);
      sg[8] = gv[2] - lam * tr - mu * (fc[8] + fc[8]);
      sg[1] = sg[3] = gv[3] - mu * (fc[1] + fc[3]);
      sg[2] = sg[6] = gv[4] - mu * (fc[2] + fc[6]);
      sg[5] = sg[7] = gv[5] - mu * (fc[5] + fc[7]);
    }
    double rhs[9], speed;
    cell_flux<double, 2, DYN, true>(u, sg, job.Lambda, rhs, speed, &vsh[0][tid], NT, double(n));
    double* o = G.out + w.inst * G.fstride + idx;
    if constexpr (!COMB) {
#pragma unroll
      for (int c = 0; c < 9; ++c) o[c * cells] = rhs[c];
    } else {
      const double* Up = G.U + w.inst * G.fstride + idx;
      const double cu = job.cu, cf = job.cf;
#pragma unroll
      for (int c = 0; c < 9; ++c) o[c * cells] = rk_combine<double>(cu, Up[c * cells], cf, fc[c], dt, rhs[c]);
    }
    if (G.partial) {
      // a non-finite speed or rhs must reach the CFL reduction
      bool fin = isfinite(speed);
#pragma unroll
      for (int c = 0; c < 9; ++c) fin = fin && isfinite(rhs[c]);
      const double m = block_max<double>(fin ? speed : INFINITY, red);
      if (tid == 0) G.partial[w.inst * G.pstride + w.tl] = m;
    }
  }
}

static void flux_multi_tile(int variant, int& tc, int& tr) {
  if (variant == 1) {
    tc = 64;
    tr = 4;
  } else if (variant == 2) {
    tc = 16;
    tr = 16;
  } else {
    tc = 32;
    tr = 8;
  }
}

bool flux_multi_ok(int n, int tile) {
  int tc, tr;
  flux_multi_tile(tile, tc, tr);
  return n >= tc && n >= tr && n % tc == 0 && n % tr == 0;
}

template <int TC, int TR>
static void launch_flux_multi_t(const FluxMultiJob& job, dim3 g, hipStream_t st) {
  const dim3 b(TC, TR);
#define CDD_MULTI(DYN_)                                                               \
  do {                                                                                \
    if (job.mode) hipLaunchKernelGGL((k_flux_multi<DYN_, TC, TR, true>), g, b, 0, st, job); \
    else hipLaunchKernelGGL((k_flux_multi<DYN_, TC, TR, false>), g, b, 0, st, job);         \
  } while (0)
  if (job.dyn == DYN_GLIDE_ONLY) CDD_MULTI(DYN_GLIDE_ONLY);
  else CDD_MULTI(DYN_GLIDE_CLIMB);
#undef CDD_MULTI
}

hipError_t launch_flux_multi(FluxMultiJob& job, int ngroups, hipStream_t st) {
  if (ngroups < 1 || ngroups > kFluxGroups) return hipErrorInvalidValue;
  if (job.dyn != DYN_GLIDE_ONLY && job.dyn != DYN_GLIDE_CLIMB) return hipErrorInvalidValue;
  int tc, tr;
  flux_multi_tile(job.tile, tc, tr);
  long long items = 0;
  for (int k = 0; k < ngroups; ++k) {
    FluxGroup& G = job.g[k];
    if (!flux_multi_ok(G.n, job.tile) || G.count < 1 || !G.clk || !G.A || !G.G || !G.out ||
        (job.mode && !G.U))
      return hipErrorInvalidValue;
    G.ntiles = (G.n / tc) * (G.n / tr);
    if (G.partial && G.pstride < G.ntiles) return hipErrorInvalidValue;
    G.item0 = (int)items;
    items += (long long)G.ntiles * G.count;
  }
  if (items > (1LL << 30)) return hipErrorInvalidValue;
  job.ngroups = ngroups;
  job.nitems = (int)items;
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  if (cus[dev] <= 0) {
    int c = 0;
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
    cus[dev] = c;
  }
  const int per_cu = job.per_cu == 1 ? 1 : 2;
  long long grid = (long long)per_cu * cus[dev];
  if (grid > items) grid = items;
  const dim3 g((unsigned)grid);
  if (job.tile == 1) launch_flux_multi_t<64, 4>(job, g, st);
  else if (job.tile == 2) launch_flux_multi_t<16, 16>(job, g, st);
  else launch_flux_multi_t<32, 8>(job, g, st);
  return hipGetLastError();
}

// Plane-marching variant of k_flux for the 3D fp64 stepper (no per-cell
// speed output, one instance), described at k_flux_march below.  The
// per-cell arithmetic is cell_flux, as in k_flux.
namespace march {
constexpr int TC = 32, TR = 8, NT = TC * TR;
// Cache policy of the loads of data read once per step (the stress planes G
// and the stage's base state U) and of the output stores: non-temporal, so
// they do not evict the state tiles whose halos the next rows and planes
// re-read from L2 (about 1% per step at 256^3 over the default policy).
constexpr int CPOL_STREAM = 2;
constexpr int PW = TC + 4, PH = TR + 4, TILE = PW * PH;  // 36 x 12 halo tile
// DMA geometry per element type: fp64 moves 16 bytes (two values of a tile
// row) per lane, fp32 4 bytes (one value).
template <typename T>
struct Dma {
  static constexpr int EPL = sizeof(T) == 8 ? 2 : 1;       // tile values per lane and DMA
  static constexpr int BYTES = EPL * (int)sizeof(T);       // 16 or 4
  static constexpr int NUNIT = 9 * TILE / EPL;             // lane units per tile
  static constexpr int NDMA = (NUNIT + NT - 1) / NT;       // tile DMA instructions per thread
  // The last DMA round is needed by waves 0..WLAST only; the waves above
  // repeat wave WLAST's (same cells to the same LDS words), so every wave
  // issues NDMA tile DMAs, and lanes past the tile fetch unit 0 into the
  // buffer's tail.
  static constexpr int WLAST = (NUNIT - 1 - (NDMA - 1) * NT) / 64;
  static constexpr int TBUF = EPL * ((NDMA - 1) * NT + (WLAST + 1) * 64);  // tile buffer (values)
  static_assert(PW % EPL == 0, "DMA units must not straddle tile rows");
  static_assert(NDMA >= 2 && WLAST >= 0 && WLAST < NT / 64, "tile DMA rounds");
  // Vector-memory instructions per wave and group, for the s_waitcnt vmcnt
  // counts of the pipeline.  Every group is issued whole, by every lane, on
  // every step (a conditional DMA makes the compiler's wait tracking merge
  // paths and wait for every DMA in flight).
  static constexpr int N_T = NDMA;                          // tile DMA
  static constexpr int N_9 = sizeof(T) == 8 ? 4 + 2 : 9;    // nine per-wave components
  static constexpr int N_GN = (sizeof(T) == 8 ? 3 : 6) + N_9;  // G + next plane
  static constexpr int N_ST = 9;                            // output stores
};
}  // namespace march

// s_waitcnt immediates (gfx9 encoding: vmcnt in bits 3:0 and 15:14, expcnt
// 6:4, lgkmcnt 11:8): wait until at most N vector-memory operations of this
// wave are outstanding; the compiler's own wait tracking sees these.
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}
// Workgroup barrier for LDS data only: every LDS access of this wave has
// completed; outstanding global stores and DMA stay in flight (their
// consumers wait on vmcnt themselves).  The memory clobber keeps the
// compiler from moving memory accesses across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// LDS byte address of a __shared__ object.
__device__ __forceinline__ unsigned lds_addr(const void* p) {
  return (unsigned)(size_t)(const __attribute__((address_space(3))) char*)p;
}
// Reads of the per-wave DMA buffers (N doubles, 512 bytes apart, from LDS
// byte address a), completed inside the statement (lgkmcnt 0).  They are
// written here because the compiler's own wait tracking for LDS DMA makes a
// plain read of these buffers wait for every DMA in flight, which would
// drain the fetch pipeline; the kernel's vmcnt waits order them instead.
__device__ __forceinline__ void lds_read6(unsigned a, double (&v)[6]) {
  asm volatile(
      "ds_read_b64 %0, %6\n\t"
      "ds_read_b64 %1, %6 offset:512\n\t"
      "ds_read_b64 %2, %6 offset:1024\n\t"
      "ds_read_b64 %3, %6 offset:1536\n\t"
      "ds_read_b64 %4, %6 offset:2048\n\t"
      "ds_read_b64 %5, %6 offset:2560\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4]), "=&v"(v[5])
      : "v"(a)
      : "memory");
}
__device__ __forceinline__ void lds_read9(unsigned a, double (&v)[9]) {
  asm volatile(
      "ds_read_b64 %0, %9\n\t"
      "ds_read_b64 %1, %9 offset:512\n\t"
      "ds_read_b64 %2, %9 offset:1024\n\t"
      "ds_read_b64 %3, %9 offset:1536\n\t"
      "ds_read_b64 %4, %9 offset:2048\n\t"
      "ds_read_b64 %5, %9 offset:2560\n\t"
      "ds_read_b64 %6, %9 offset:3072\n\t"
      "ds_read_b64 %7, %9 offset:3584\n\t"
      "ds_read_b64 %8, %9 offset:4096\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4]), "=&v"(v[5]), "=&v"(v[6]), "=&v"(v[7]),
        "=&v"(v[8])
      : "v"(a)
      : "memory");
}
// The next-plane cells (9 values from a9) and the stress planes (6 from a6)
// together, under one wait.
__device__ __forceinline__ void lds_read9_6(unsigned a9, unsigned a6, double (&v)[9], double (&g)[6]) {
  asm volatile(
      "ds_read_b64 %0, %15\n\t"
      "ds_read_b64 %1, %15 offset:512\n\t"
      "ds_read_b64 %2, %15 offset:1024\n\t"
      "ds_read_b64 %3, %15 offset:1536\n\t"
      "ds_read_b64 %4, %15 offset:2048\n\t"
      "ds_read_b64 %5, %15 offset:2560\n\t"
      "ds_read_b64 %6, %15 offset:3072\n\t"
      "ds_read_b64 %7, %15 offset:3584\n\t"
      "ds_read_b64 %8, %15 offset:4096\n\t"
      "ds_read_b64 %9, %16\n\t"
      "ds_read_b64 %10, %16 offset:512\n\t"
      "ds_read_b64 %11, %16 offset:1024\n\t"
      "ds_read_b64 %12, %16 offset:1536\n\t"
      "ds_read_b64 %13, %16 offset:2048\n\t"
      "ds_read_b64 %14, %16 offset:2560\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4]), "=&v"(v[5]), "=&v"(v[6]), "=&v"(v[7]),
        "=&v"(v[8]), "=&v"(g[0]), "=&v"(g[1]), "=&v"(g[2]), "=&v"(g[3]), "=&v"(g[4]), "=&v"(g[5])
      : "v"(a9), "v"(a6)
      : "memory");
}
// The same for fp32 buffers (N floats, 256 bytes apart).
__device__ __forceinline__ void lds_read6(unsigned a, float (&v)[6]) {
  asm volatile(
      "ds_read_b32 %0, %6\n\t"
      "ds_read_b32 %1, %6 offset:256\n\t"
      "ds_read_b32 %2, %6 offset:512\n\t"
      "ds_read_b32 %3, %6 offset:768\n\t"
      "ds_read_b32 %4, %6 offset:1024\n\t"
      "ds_read_b32 %5, %6 offset:1280\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4]), "=&v"(v[5])
      : "v"(a)
      : "memory");
}
__device__ __forceinline__ void lds_read9(unsigned a, float (&v)[9]) {
  asm volatile(
      "ds_read_b32 %0, %9\n\t"
      "ds_read_b32 %1, %9 offset:256\n\t"
      "ds_read_b32 %2, %9 offset:512\n\t"
      "ds_read_b32 %3, %9 offset:768\n\t"
      "ds_read_b32 %4, %9 offset:1024\n\t"
      "ds_read_b32 %5, %9 offset:1280\n\t"
      "ds_read_b32 %6, %9 offset:1536\n\t"
      "ds_read_b32 %7, %9 offset:1792\n\t"
      "ds_read_b32 %8, %9 offset:2048\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4]), "=&v"(v[5]), "=&v"(v[6]), "=&v"(v[7]),
        "=&v"(v[8])
      : "v"(a)
      : "memory");
}

// Plane-marching flux + stage kernel body (3D, fp64 or fp32, central upwind,
// glide + climb or glide only; the kernels k_flux_march / k_flux_march32
// below).  A block owns a TC x TR column of cells (fast axes) and
// marches a chunk of planes along the slow axis, one cell per thread.  Per
// plane it DMAs (global_load_lds) into LDS the halo tile of the stage field,
// the six non-local stress planes G and (stage combinations) the step-start
// state U, and carries the slow-axis differences in registers.
// The fetches are pipelined so that the memory pipe never drains between
// planes, in one of three schemes (NB tile buffers):
//  - 4 (fp64 stage combinations): a ring of four tile buffers.  Plane p reads
//    its stencil from tile(p), the slow axis's next cells f(p + 2) from the
//    centre of tile(p + 2), and fetches tile(p + 3) into the buffer plane
//    p - 1 used, so every plane of the stage field is fetched once.
//  - 3 (fp64 first stage, whose shorter plane leaves too little time for a
//    tile fetched mid-plane to land by the plane's end): three tile
//    buffers, tile(p + 2) fetched into plane p - 1's, and the next cells
//    f(p + 3) into a per-wave buffer of their own.
//  - 2 (fp32, two blocks per CU at 80 KB of LDS each): two tile buffers,
//    tile(p + 2) fetched into plane p's after a second barrier, and the next
//    cells as for 3.
// One barrier per plane for 4 and 3.  G(p + 1) goes out as soon as plane p has
// read G(p), U(p + 1) as soon as U(p) is read.
// CLIMB: Lambda != 0 (climb suppression terms compiled in); COMB: a stage
// combination (mode != 0, reads U) rather than the bare rhs; SPEED (rhs
// only): also the per-cell CFL speed plane (cdd_flux's speed output); LLF
// (fp64): the CUDA executable's flux (llf_derivs, cell_flux_llf), its
// slow-axis stencil carried as the four values around the plane.  The host
// picks the instance from the job; all are compile-time so that the
// compiler's wait tracking sees one vector-memory sequence per step.
template <typename T, int DYN, bool CLIMB, bool COMB, bool SPEED = false, bool LLF = false>
__device__ __forceinline__ void flux_march_body(const FluxJob& job, int chunk) {
  static_assert(!(COMB && SPEED), "the speed plane comes with the bare rhs");
  static_assert(!LLF || sizeof(T) == 8, "the LLF flux is float64");
  using namespace march;
  using DM = Dma<T>;
  constexpr int NDMA = DM::NDMA, WLAST = DM::WLAST, EPL = DM::EPL, NUNIT = DM::NUNIT;
  constexpr int N_T = DM::N_T, N_9 = DM::N_9, N_GN = DM::N_GN, N_ST = DM::N_ST + (SPEED ? 1 : 0);
  typedef __attribute__((address_space(3))) void lds_void;
  // the job's field pointers carry T data (float for fp32 plans)
  const T* __restrict__ A = reinterpret_cast<const T*>(job.A);
  const T* __restrict__ Gp = reinterpret_cast<const T*>(job.G);
  const T* __restrict__ Up = reinterpret_cast<const T*>(job.U);
  T* __restrict__ outp = reinterpret_cast<T*>(job.out);
  if (job.clk && job.clk->done) return;
  // NB 4: the ring as one array (155 KB of LDS in all); no tile DMA is in
  // flight while any of its buffers is read.  NB 3 and 2: the tile buffers
  // as separate arrays (plane k of the chunk in buffer k % NB), so the
  // compiler sees that reads of one never wait for the DMA into another.
  constexpr int NB = sizeof(T) == 8 ? (COMB ? 4 : 3) : 2;
  constexpr bool R4 = NB == 4;
  __shared__ T tring[R4 ? 4 : 1][R4 ? DM::TBUF : 1];
  __shared__ T tile0[R4 ? 1 : DM::TBUF];
  __shared__ T tile1[R4 ? 1 : DM::TBUF];
  __shared__ T tile2[NB == 3 ? DM::TBUF : 1];
  // fp32: the corner velocities are parked in LDS (two blocks per CU need
  // the registers); fp64 (one block per CU) keeps them in registers
  __shared__ T vsh[sizeof(T) == 8 ? 1 : 3 * 8][sizeof(T) == 8 ? 1 : NT];
  __shared__ T gb[NT / 64][6][64];                       // stress planes
  __shared__ T nbuf[R4 ? 1 : NT / 64][R4 ? 1 : 9][64];  // NB 3, 2: the slow axis's next plane
  __shared__ T ub[NT / 64][9][64];                       // U cells
  __shared__ T red[NT / 64];
  const int n = job.n;
  const long long cells = job.cells, splane = (long long)n * n;
  const int tid = threadIdx.x, tc = tid % TC, trr = tid / TC;
  // the wave index is uniform: in a scalar register, the per-wave LDS
  // addresses (the DMA destinations) are scalar arithmetic
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  // block -> (chunk of planes, column); each XCD (hardware block b runs on
  // XCD b % 8) takes one contiguous range of logical blocks, i.e. adjacent
  // columns at the same planes, so shared halo rows are served from its L2
  const int nct = n / TC, ncol = nct * (n / TR);
  const int nb = ncol * (n / chunk);
  const int b = blockIdx.x;
  const int L = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
  const int col = L % ncol, x0 = (L / ncol) * chunk;
  const int c0 = (col % nct) * TC, r0 = (col / nct) * TR;
  const T dt = job.clk ? T(job.clk->dt) : T(0);
  const T mu = T(job.mu), lam = T(job.lam), Lambda = T(job.Lambda);
  const T cu = T(job.cu), cf = T(job.cf);
  auto wrapp = [n](int p) { return p < 0 ? p + n : (p >= n ? p - n : p); };

  // per-lane element offsets inside one plane (plane base added per step):
  // halo-tile units (component, row, column; a fp64 unit is an even-column
  // pair, and n is even, so a pair never straddles the periodic wrap), the
  // cells of this lane in the per-wave buffers (fp64: its pair, and the
  // cell of its 4-byte word of component 8), and this thread's own cell
  unsigned toff[NDMA];
#pragma unroll
  for (int k = 0; k < NDMA; ++k) {
    const int wk = (k == NDMA - 1 && wv > WLAST) ? WLAST : wv;
    const int e = k * NT + wk * 64 + lane;
    const int ee = e < NUNIT ? e : 0;
    const int comp = ee / (TILE / EPL), rem = ee % (TILE / EPL);
    const int gr = wrapp(r0 - 2 + rem / (PW / EPL)), gc = wrapp(c0 - 2 + EPL * (rem % (PW / EPL)));
    toff[k] = (unsigned)(comp * cells + (long long)gr * n + gc);
  }
  const int half = lane >> 5, e2 = 64 * wv + 2 * (lane & 31);
  const unsigned poff = (unsigned)((r0 + e2 / TC) * n + c0 + e2 % TC);
  const unsigned woff = (unsigned)((r0 + 2 * wv) * n + c0 + (lane >> 1));  // + n: the wave's second row
  const unsigned xoff = (unsigned)((r0 + trr) * n + c0 + tc);
  const unsigned a_gb = lds_addr(&gb[wv][0][lane]), a_nb = lds_addr(&nbuf[R4 ? 0 : wv][0][lane]),
                 a_ub = lds_addr(&ub[wv][0][lane]);

  // Planes wrap periodically, so a fetch past the chunk reads valid cells
  // that are never used.
  auto dma_tile = [&](int p, T* dst) {
    const T* base = A + (long long)wrapp(p) * splane;
#pragma unroll
    for (int k = 0; k < NDMA; ++k) {
      const int wk = (k == NDMA - 1 && wv > WLAST) ? WLAST : wv;
      __builtin_amdgcn_global_load_lds((const void*)(base + toff[k]), (lds_void*)&dst[EPL * (k * NT + wk * 64)],
                                       DM::BYTES, 0, 0);
    }
  };
  // fp64: two components per 16-byte DMA, lanes 0-31 the even one, 32-63
  // the odd; fp32: one component per 4-byte DMA, each lane its own cell
  auto dma_g = [&](int p) {
    if constexpr (EPL == 2) {
      const T* base = Gp + (long long)wrapp(p) * splane + poff;
#pragma unroll
      for (int k = 0; k < 3; ++k)
        __builtin_amdgcn_global_load_lds((const void*)(base + (2 * k + half) * cells), (lds_void*)&gb[wv][2 * k][0],
                                         16, 0, march::CPOL_STREAM);
    } else {
      const T* base = Gp + (long long)wrapp(p) * splane + xoff;
#pragma unroll
      for (int c = 0; c < 6; ++c)
        __builtin_amdgcn_global_load_lds((const void*)(base + c * cells), (lds_void*)&gb[wv][c][0], 4, 0, march::CPOL_STREAM);
    }
  };
  // nine components; fp64: 0-7 in pairs, 8 as 4-byte words (lane l: word
  // l & 1 of cell l >> 1 of the wave's first, then second row); the cache
  // policy CP is an integral_constant (the DMA takes it as an immediate)
  auto dma_9 = [&](const T* src, int p, T (*dst)[64], auto cp) {
    constexpr int CP = decltype(cp)::value;
    const T* plane = src + (long long)wrapp(p) * splane;
    if constexpr (EPL == 2) {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        __builtin_amdgcn_global_load_lds((const void*)(plane + poff + (2 * k + half) * cells),
                                         (lds_void*)&dst[2 * k][0], 16, 0, CP);
      const char* w8 = (const char*)(plane + 8 * cells + woff) + 4 * (lane & 1);
      __builtin_amdgcn_global_load_lds((const void*)w8, (lds_void*)&dst[8][0], 4, 0, CP);
      __builtin_amdgcn_global_load_lds((const void*)(w8 + 8 * (size_t)n), (lds_void*)&dst[8][32], 4, 0, CP);
    } else {
#pragma unroll
      for (int c = 0; c < 9; ++c)
        __builtin_amdgcn_global_load_lds((const void*)(plane + xoff + c * cells), (lds_void*)&dst[c][0], 4, 0, CP);
    }
  };
  auto center = [&](int p, T (&v)[9]) {
    const T* base = A + (long long)wrapp(p) * splane + xoff;
#pragma unroll
    for (int c = 0; c < 9; ++c) v[c] = base[c * cells];
  };

  // Slow-axis derivatives are carried from plane to plane: with
  // D_k = f(k+1) - f(k), plane p needs D_{p+1} = f(p+2) - f(p+1), the
  // carried f(p+1), D_p, s_p = D_p - D_{p-1} and u-(p) = D_{p-1} + m_{p-1}/2;
  // it forms m_p = minmod(D_{p+1} - D_p, s_p) and u+(p) = D_p - m_p/2 and
  // hands on D_{p+1} - D_p and D_p + m_p/2.  Every value comes from the same
  // operands by the same operation as in limited_derivs, so the result is
  // bit-identical to it.
  if constexpr (R4) {
    dma_tile(x0, tring[0]);
    dma_tile(x0 + 1, tring[1]);
    dma_tile(x0 + 2, tring[2]);
  } else {
    dma_tile(x0, tile0);
    dma_tile(x0 + 1, tile1);
    dma_9(A, x0 + 2, nbuf[wv], std::integral_constant<int, 0>{});
  }
  dma_g(x0);
  if constexpr (COMB) dma_9(Up, x0, ub[wv], std::integral_constant<int, march::CPOL_STREAM>{});
  // LLF: f(p-2), f(p-1) and f(p+1) carried in xs, xd and xf
  T xf[9], xd[9], xs[9], xum[9];
  if constexpr (LLF) {
    center(x0 - 2, xs);
    center(x0 - 1, xd);
    center(x0 + 1, xf);
  } else {
    T fm2[9], fm1[9], f0[9];
    center(x0 - 2, fm2);
    center(x0 - 1, fm1);
    center(x0, f0);
    center(x0 + 1, xf);
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      const T Dm2 = fm1[c] - fm2[c], Dm1 = f0[c] - fm1[c];
      xd[c] = xf[c] - f0[c];
      xs[c] = xd[c] - Dm1;
      xum[c] = fma(T(0.5), minmod2(xs[c], Dm1 - Dm2), Dm1);
    }
  }
  wait_vmcnt<0>();
  T run_max = -INFINITY;
  const int lr = trr + 2, lc = tc + 2;

  // One plane, its stencil tile in tcur; tpre receives the tile of p + 3
  // (NB 4; tnx holds tile(p + 2)) or of p + 2 (NB 3, 2).  Per wave, the
  // vector-memory issue order of a step is
  //   NB 4:    [G(p+1)] [tile(p+3)] [stores(p)] [U(p+1)]
  //   NB 3, 2: [G(p+1), next(p+3)] [tile(p+2)] [stores(p)] [U(p+1)]
  // so U(p) (issued last on the step before) is complete once at most
  // N_G + N_T (N_GN + N_T) younger operations are outstanding.  A step ends
  // once G(p+1) and tile(p+3) (NB 4), or G(p+1), next(p+3) and tile(p+1)
  // (NB 3, 2; issued on the step before), are complete: at most N_ST
  // (N_T + N_ST) outstanding, + N_9 for U(p+1).
  constexpr int N_G = EPL == 2 ? 3 : 6;
  auto step = [&](int p, T* tcur, T* tpre, const T* tnx) __attribute__((always_inline)) {
    // tile(p), G(p) and (fp32) next(p+2), this wave's part, landed at the
    // wait that closed the previous step; the barrier publishes every wave's
    // part
    lds_barrier();
    T nv[9], gv[6];
    if constexpr (R4) {
      lds_read6(a_gb, gv);
#pragma unroll
      for (int c = 0; c < 9; ++c) nv[c] = tnx[c * TILE + lr * PW + lc];
    } else if constexpr (sizeof(T) == 8) {
      lds_read9_6(a_nb, a_gb, nv, gv);
    } else {
      lds_read9(a_nb, nv);
      lds_read6(a_gb, gv);
    }
    // this wave's G and next-plane cells are read (the reads complete inside
    // lds_read*): refill them for the next plane (per-wave buffers, no
    // barrier).  Without the combination, or with three tile buffers, the
    // refill goes out at once, under the plane's whole arithmetic; with the
    // combination and two buffers (U in flight too) after the derivatives,
    // which measured faster there.
    if constexpr (R4) {
      dma_g(p + 1);
    } else if constexpr (!COMB) {
      dma_g(p + 1);
      dma_9(A, p + 3, nbuf[wv], std::integral_constant<int, 0>{});
    }
    // The in-plane stencil values are read in three groups of three
    // components, each group's LDS reads issued before the arithmetic of the
    // group before it, so that one wave per SIMD does not sit out every read's
    // latency (the scheduling barriers keep the compiler from sinking the
    // reads to their first use).
    T u[3][2][9], fc[9];
    T sv[2][3][9];  // per component: z - 2 .. z + 2, then y - 2, y - 1, y + 1, y + 2
    auto read_group = [&](int g, T (&s)[3][9]) __attribute__((always_inline)) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const T* tb = tcur + (3 * g + k) * TILE;
#pragma unroll
        for (int o = 0; o < 5; ++o) s[k][o] = tb[lr * PW + lc - 2 + o];
        s[k][5] = tb[(lr - 2) * PW + lc];
        s[k][6] = tb[(lr - 1) * PW + lc];
        s[k][7] = tb[(lr + 1) * PW + lc];
        s[k][8] = tb[(lr + 2) * PW + lc];
      }
    };
    read_group(0, sv[0]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int g = 0; g < 3; ++g) {
      if (g < 2) read_group(g + 1, sv[(g + 1) & 1]);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int c = 3 * g + k;
        const T* s = sv[g & 1][k];
        fc[c] = s[2];
        if constexpr (LLF) {
          T fz[5] = {s[0], s[1], s[2], s[3], s[4]};
          llf_derivs<T>(fz, T(n), u[2][0][c], u[2][1][c]);
          T fy[5] = {s[5], s[6], s[2], s[7], s[8]};
          llf_derivs<T>(fy, T(n), u[1][0][c], u[1][1][c]);
          T fx[5] = {xs[c], xd[c], s[2], xf[c], nv[c]};
          llf_derivs<T>(fx, T(n), u[0][0][c], u[0][1][c]);
          xs[c] = xd[c];
          xd[c] = s[2];
          xf[c] = nv[c];
          continue;
        }
        {
          T f[5] = {s[0], s[1], s[2], s[3], s[4]};
          limited_derivs<T>(f, u[2][0][c], u[2][1][c]);
        }
        {
          T f[5] = {s[5], s[6], s[2], s[7], s[8]};
          limited_derivs<T>(f, u[1][0][c], u[1][1][c]);
        }
        {  // slow axis, from the carried differences
          const T D1 = nv[c] - xf[c];
          const T s1 = D1 - xd[c];
          const T m = minmod2(s1, xs[c]);
          u[0][0][c] = fma(T(-0.5), m, xd[c]);
          u[0][1][c] = xum[c];
          xum[c] = fma(T(0.5), m, xd[c]);
          xf[c] = nv[c];
          xs[c] = s1;
          xd[c] = D1;
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (COMB && !R4) {
      dma_g(p + 1);
      dma_9(A, p + 3, nbuf[wv], std::integral_constant<int, 0>{});
    }
    T sg[9];
    {
      const T tr = fc[0] + fc[4] + fc[8];
      sg[0] = gv[0] - lam * tr - mu * (fc[0] + fc[0]);
      sg[4] = gv[1] - lam * tr - mu * (fc[4] + fc[4]);
      sg[8] = gv[2] - lam * tr - mu * (fc[8] + fc[8]);
      sg[1] = sg[3] = gv[3] - mu * (fc[1] + fc[3]);
      sg[2] = sg[6] = gv[4] - mu * (fc[2] + fc[6]);
      sg[5] = sg[7] = gv[5] - mu * (fc[5] + fc[7]);
    }
    // materialise the derivatives here: left to itself the compiler sinks
    // their arithmetic past the barrier and keeps all 81 raw stencil values
    // per thread live instead of the 54 derivatives
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int c = 0; c < 9; ++c) asm volatile("" : "+v"(u[a][s][c]));
#pragma unroll
    for (int c = 0; c < 9; ++c) asm volatile("" : "+v"(sg[c]), "+v"(fc[c]));
    // NB 4, 3: every wave has finished plane p - 1 (the barrier that opened
    // this plane), so its buffer takes the tile of p + 3 (p + 2); NB 2: every
    // wave is done with tile(p) before tile(p + 2) refills its buffer.
    // Issued after the stencil reads, so that the DMA does not compete with
    // them for the LDS.
    if constexpr (R4) {
      dma_tile(p + 3, tpre);
    } else {
      if constexpr (NB == 2) lds_barrier();
      dma_tile(p + 2, tpre);
    }
    T rhs[9], speed;
    if constexpr (sizeof(T) == 8) {
      T vreg[3 * 8];
      if constexpr (LLF) cell_flux_llf<T, 3, DYN>(u, sg, Lambda, rhs, speed, vreg, 1);
      else cell_flux<T, 3, DYN, true, false, CLIMB ? 1 : 0, true>(u, sg, Lambda, rhs, speed, vreg, 1, T(n));
    } else {
      cell_flux<T, 3, DYN, true, false, CLIMB ? 1 : 0, true>(u, sg, Lambda, rhs, speed, &vsh[0][tid], NT, T(n));
    }
    bool fin = isfinite(speed);
#pragma unroll
    for (int c = 0; c < 9; ++c) fin = fin && isfinite(rhs[c]);
    run_max = fmax(run_max, fin ? speed : T(INFINITY));
    T* o = outp + (long long)p * splane + xoff;
    if constexpr (COMB) {
      wait_vmcnt<(R4 ? N_G : N_GN) + N_T>();  // U(p)
      T uv[9], ov[9];
      lds_read9(a_ub, uv);
#pragma unroll
      for (int c = 0; c < 9; ++c) {
        ov[c] = rk_combine<T>(cu, uv[c], cf, fc[c], dt, rhs[c]);
        __builtin_nontemporal_store(ov[c], &o[c * cells]);
      }
      // vacancy coupling after the last stage: c += beta (tr F2 - tr u)
      if (job.vac) {
        job.vac[(long long)p * splane + xoff] += job.vbeta * ((ov[0] + ov[4] + ov[8]) - (uv[0] + uv[4] + uv[8]));
        wait_vmcnt<0>();
      }
      dma_9(Up, p + 1, ub[wv], std::integral_constant<int, march::CPOL_STREAM>{});
      wait_vmcnt<(R4 ? 0 : N_T) + N_ST + N_9>();
    } else {
#pragma unroll
      for (int c = 0; c < 9; ++c) __builtin_nontemporal_store(rhs[c], &o[c * cells]);
      if constexpr (SPEED) reinterpret_cast<T*>(job.speed)[(long long)p * splane + xoff] = speed;
      wait_vmcnt<(R4 ? 0 : N_T) + N_ST>();
    }
  };
  if constexpr (R4) {
    for (int p = x0; p < x0 + chunk; ++p) {
      const int k = p - x0;
      step(p, tring[k & 3], tring[(k + 3) & 3], tring[(k + 2) & 3]);
    }
  } else if constexpr (NB == 3) {
    int p = x0;
    for (; p + 3 <= x0 + chunk; p += 3) {
      step(p, tile0, tile2, nullptr);
      step(p + 1, tile1, tile0, nullptr);
      step(p + 2, tile2, tile1, nullptr);
    }
    if (p < x0 + chunk) step(p, tile0, tile2, nullptr);
    if (p + 1 < x0 + chunk) step(p + 1, tile1, tile0, nullptr);
  } else {
    for (int p = x0; p < x0 + chunk; p += 2) {
      step(p, tile0, tile0, nullptr);
      step(p + 1, tile1, tile1, nullptr);
    }
  }
  // nothing may still write this block's LDS when it exits
  wait_vmcnt<0>();
  if (job.partial) {
    const T m = block_max<T>(run_max, red);
    if (tid == 0) job.partial[blockIdx.x] = (double)m;
  }
}

template <int DYN, bool CLIMB, bool COMB, bool SPEED = false, bool LLF = false>
__global__ void __launch_bounds__(march::NT) k_flux_march(FluxJob job, int chunk) {
  flux_march_body<double, DYN, CLIMB, COMB, SPEED, LLF>(job, chunk);
}
// fp32 plans: the same kernel on float fields.  Its registers (at most 256
// per lane) and LDS (80 KB) admit two blocks per CU, two waves per SIMD.
template <int DYN, bool CLIMB, bool COMB, bool SPEED = false>
__global__ void __launch_bounds__(march::NT, 2) k_flux_march32(FluxJob job, int chunk) {
  flux_march_body<float, DYN, CLIMB, COMB, SPEED>(job, chunk);
}

// ---------------------------------------------------------------------------
// Row-marching persistent multi-instance 2D flux (k_flux_rows), fp64: the
// batched stepper's flux as one launch over every instance of every size
// group whose side is a multiple of 128.  Every block takes one contiguous,
// equal share of the flat (group, instance, 128-column strip, row) space and
// marches down its rows: per row, the 128-cell segment with its 2-cell halo
// arrives in LDS by asynchronous DMA two rows ahead, the stress cells and
// the next row's own cells one row ahead (per-wave buffers), and the
// row-axis derivatives are carried from row to row in registers exactly as
// the 3D marching kernel carries its slow axis (bit-identical to
// limited_derivs).  The per-cell arithmetic is k_flux's (cell_flux with
// PRE, the same sigma and stage combination), so the states equal the
// per-tile kernels' bit for bit.  The CFL maximum of each instance goes to
// its first partial slot by an atomic maximum on the bit pattern (speeds
// are positive, a non-finite speed or rhs enters as +inf); the CFL
// reduction reads and clears that slot.
namespace rows {
constexpr int TC = 128, NT = TC;                                  // one row segment, one cell per thread
constexpr int PW = TC + 4;                                        // with its 2-cell halo
constexpr int EPL = 2;                                            // values per 16-byte DMA unit
constexpr int NUNIT = 9 * PW / EPL;                               // DMA units of a row tile
constexpr int NDMA = (NUNIT + NT - 1) / NT;                       // tile DMA instructions per thread
constexpr int WLAST = (NUNIT - 1 - (NDMA - 1) * NT) / 64;         // waves the last round needs
constexpr int TBUF = EPL * ((NDMA - 1) * NT + (WLAST + 1) * 64);  // tile buffer (values)
constexpr int N_T = NDMA, N_9 = 4 + 2, N_G = 3, N_GN = N_G + N_9, N_ST = 9;
static_assert(PW % EPL == 0 && NDMA >= 2 && WLAST >= 0 && WLAST < NT / 64, "row tile DMA rounds");
}  // namespace rows

bool flux_rows_ok(int n) { return n >= rows::TC && n % rows::TC == 0; }

template <int DYN, bool COMB>
__global__ void __launch_bounds__(rows::NT) k_flux_rows(FluxMultiJob job) {
  using namespace rows;
  typedef __attribute__((address_space(3))) void lds_void;
  __shared__ double tile0[TBUF];
  __shared__ double tile1[TBUF];
  __shared__ double gb[NT / 64][6][64];    // stress cells
  __shared__ double nbuf[NT / 64][9][64];  // the row axis's next row
  __shared__ double ub[NT / 64][9][64];    // U cells
  __shared__ double red[NT / 64];
  const int tid = threadIdx.x;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int half = lane >> 5;
  const unsigned a_gb = lds_addr(&gb[wv][0][lane]), a_nb = lds_addr(&nbuf[wv][0][lane]),
                 a_ub = lds_addr(&ub[wv][0][lane]);
  const double mu = job.mu, lam = job.lam, Lambda = job.Lambda, cu = job.cu, cf = job.cf;
  const long long total = job.nitems;
  const long long r_beg = total * blockIdx.x / gridDim.x, r_end = total * (blockIdx.x + 1) / gridDim.x;
  long long r = r_beg;
  while (r < r_end) {
    // this row-step's group (constant indices into the kernel-argument
    // table), instance, strip and row
    FluxGroup G = job.g[0];
#pragma unroll
    for (int k = 1; k < kFluxGroups; ++k) {
      if (k < job.ngroups && r >= job.g[k].item0) {
        G.n = job.g[k].n;
        G.ntiles = job.g[k].ntiles;
        G.item0 = job.g[k].item0;
        G.cells = job.g[k].cells;
        G.fstride = job.g[k].fstride;
        G.gstride = job.g[k].gstride;
        G.pstride = job.g[k].pstride;
        G.A = job.g[k].A;
        G.G = job.g[k].G;
        G.U = job.g[k].U;
        G.out = job.g[k].out;
        G.partial = job.g[k].partial;
        G.clk = job.g[k].clk;
      }
    }
    const int n = G.n;
    const long long loc = r - G.item0;
    const long long inst = loc / G.ntiles;
    const int rem = (int)(loc % G.ntiles);
    const int c0 = (rem / n) * TC, p0 = rem % n;
    const int len = (int)((r_end - r) < (long long)(n - p0) ? (r_end - r) : (long long)(n - p0));
    r += len;
    const Clock* ck = G.clk + inst;
    if (ck->done) continue;  // uniform over the block
    const double dt = ck->dt;
    const long long cells = G.cells;
    const double* __restrict__ A = G.A + inst * G.fstride;
    const double* __restrict__ Gp = G.G + inst * G.gstride;
    const double* __restrict__ Up = COMB ? G.U + inst * G.fstride : nullptr;
    double* __restrict__ outp = G.out + inst * G.fstride;
    auto wrapr = [n](int p) { return p < 0 ? p + n : (p >= n ? p - n : p); };
    // per-lane element offsets inside one row (row base added per step):
    // halo-segment units (component, even column pair; n is even, so a pair
    // never straddles the periodic wrap), the cells of this lane in the
    // per-wave buffers (its pair, and the cell of its 4-byte word of
    // component 8), and this thread's own cell
    unsigned toff[NDMA];
#pragma unroll
    for (int k = 0; k < NDMA; ++k) {
      const int wk = (k == NDMA - 1 && wv > WLAST) ? WLAST : wv;
      const int e = k * NT + wk * 64 + lane;
      const int ee = e < NUNIT ? e : 0;
      const int comp = ee / (PW / EPL), pr = ee % (PW / EPL);
      int gc = c0 - 2 + EPL * pr;
      gc = gc < 0 ? gc + n : (gc >= n ? gc - n : gc);
      toff[k] = (unsigned)(comp * cells + gc);
    }
    const unsigned poff = (unsigned)(c0 + 64 * wv + 2 * (lane & 31));
    const unsigned woff = (unsigned)(c0 + 64 * wv + (lane >> 1));  // + 32: the wave's second half
    const unsigned xoff = (unsigned)(c0 + tid);
    auto dma_tile = [&](int p, double* dst) {
      const double* base = A + (long long)wrapr(p) * n;
#pragma unroll
      for (int k = 0; k < NDMA; ++k) {
        const int wk = (k == NDMA - 1 && wv > WLAST) ? WLAST : wv;
        __builtin_amdgcn_global_load_lds((const void*)(base + toff[k]), (lds_void*)&dst[EPL * (k * NT + wk * 64)], 16,
                                         0, 0);
      }
    };
    auto dma_g = [&](int p) {
      const double* base = Gp + (long long)wrapr(p) * n + poff;
#pragma unroll
      for (int k = 0; k < 3; ++k)
        __builtin_amdgcn_global_load_lds((const void*)(base + (2 * k + half) * cells), (lds_void*)&gb[wv][2 * k][0], 16,
                                         0, march::CPOL_STREAM);
    };
    auto dma_9 = [&](const double* src, int p, double (*dst)[64], auto cp) {
      constexpr int CP = decltype(cp)::value;
      const double* row = src + (long long)wrapr(p) * n;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        __builtin_amdgcn_global_load_lds((const void*)(row + poff + (2 * k + half) * cells), (lds_void*)&dst[2 * k][0],
                                         16, 0, CP);
      const char* w8 = (const char*)(row + 8 * cells + woff) + 4 * (lane & 1);
      __builtin_amdgcn_global_load_lds((const void*)w8, (lds_void*)&dst[8][0], 4, 0, CP);
      __builtin_amdgcn_global_load_lds((const void*)(w8 + 8 * 32), (lds_void*)&dst[8][32], 4, 0, CP);
    };
    auto center = [&](int p, double (&v)[9]) {
      const double* base = A + (long long)wrapr(p) * n + xoff;
#pragma unroll
      for (int c = 0; c < 9; ++c) v[c] = base[c * cells];
    };
    // row-axis derivatives carried from row to row (as in the 3D marching
    // kernel's slow axis): f(p+2) arrives in nbuf one row ahead
    dma_tile(p0, tile0);
    dma_tile(p0 + 1, tile1);
    dma_g(p0);
    dma_9(A, p0 + 2, nbuf[wv], std::integral_constant<int, 0>{});
    if constexpr (COMB) dma_9(Up, p0, ub[wv], std::integral_constant<int, march::CPOL_STREAM>{});
    double xf[9], xd[9], xs[9], xum[9];
    {
      double fm2[9], fm1[9], f0[9];
      center(p0 - 2, fm2);
      center(p0 - 1, fm1);
      center(p0, f0);
      center(p0 + 1, xf);
#pragma unroll
      for (int c = 0; c < 9; ++c) {
        const double Dm2 = fm1[c] - fm2[c], Dm1 = f0[c] - fm1[c];
        xd[c] = xf[c] - f0[c];
        xs[c] = xd[c] - Dm1;
        xum[c] = fma(0.5, minmod2(xs[c], Dm1 - Dm2), Dm1);
      }
    }
    wait_vmcnt<0>();
    double run_max = -INFINITY;
    const int lc = tid + 2;
    // One row.  Per wave, the vector-memory issue order of a step is
    //   [G(p+1), next(p+3)] [tile(p+2)] [stores(p)] [U(p+1)]
    // so U(p) (issued last on the step before) is complete once at most
    // N_GN + N_T younger operations are outstanding, and G(p+1), next(p+3)
    // and tile(p+1) (issued on the step before) once at most N_T + N_ST
    // (+ N_9 for U(p+1)) are.
    auto step = [&](int p, double* tcur) __attribute__((always_inline)) {
      // tile(p), G(p) and next(p+2) (this wave's part) landed at the wait
      // that closed the previous step; the barrier publishes every wave's part
      lds_barrier();
      double nv[9], gv[6];
      lds_read9(a_nb, nv);
      lds_read6(a_gb, gv);
      if constexpr (!COMB) {  // the refill goes out under the whole row's arithmetic
        dma_g(p + 1);
        dma_9(A, p + 3, nbuf[wv], std::integral_constant<int, 0>{});
      }
      double u[2][2][9], fc[9];
#pragma unroll
      for (int c = 0; c < 9; ++c) {
        const double* tb = tcur + c * PW;
        const double ctr = tb[lc];
        fc[c] = ctr;
        {  // contiguous axis
          double f[5] = {tb[lc - 2], tb[lc - 1], ctr, tb[lc + 1], tb[lc + 2]};
          limited_derivs<double>(f, u[1][0][c], u[1][1][c]);
        }
        {  // row axis, from the carried differences
          const double D1 = nv[c] - xf[c];
          const double s1 = D1 - xd[c];
          const double m = minmod2(s1, xs[c]);
          u[0][0][c] = fma(-0.5, m, xd[c]);
          u[0][1][c] = xum[c];
          xum[c] = fma(0.5, m, xd[c]);
          xf[c] = nv[c];
          xs[c] = s1;
          xd[c] = D1;
        }
      }
      if constexpr (COMB) {
        dma_g(p + 1);
        dma_9(A, p + 3, nbuf[wv], std::integral_constant<int, 0>{});
      }
      double sg[9];
      {
        const double tr = fc[0] + fc[4] + fc[8];
        sg[0] = gv[0] - lam * tr - mu * (fc[0] + fc[0]);
        sg[4] = gv[1] - lam * tr - mu * (fc[4] + fc[4]);
        sg[8] = gv[2] - lam * tr - mu * (fc[8] + fc[8]);
        sg[1] = sg[3] = gv[3] - mu * (fc[1] + fc[3]);
        sg[2] = sg[6] = gv[4] - mu * (fc[2] + fc[6]);
        sg[5] = sg[7] = gv[5] - mu * (fc[5] + fc[7]);
      }
      // every wave is done with tile(p) before tile(p + 2) refills its buffer
      lds_barrier();
      dma_tile(p + 2, tcur);
      double rhs[9], speed, vreg[16];
      cell_flux<double, 2, DYN, true>(u, sg, Lambda, rhs, speed, vreg, 1, double(n));
      bool fin = isfinite(speed);
#pragma unroll
      for (int c = 0; c < 9; ++c) fin = fin && isfinite(rhs[c]);
      run_max = fmax(run_max, fin ? speed : INFINITY);
      double* o = outp + (long long)p * n + xoff;
      if constexpr (COMB) {
        wait_vmcnt<N_GN + N_T>();  // U(p)
        double uv[9];
        lds_read9(a_ub, uv);
#pragma unroll
        for (int c = 0; c < 9; ++c)
          __builtin_nontemporal_store(rk_combine<double>(cu, uv[c], cf, fc[c], dt, rhs[c]), &o[c * cells]);
        dma_9(Up, p + 1, ub[wv], std::integral_constant<int, march::CPOL_STREAM>{});
        wait_vmcnt<N_T + N_ST + N_9>();
      } else {
#pragma unroll
        for (int c = 0; c < 9; ++c) __builtin_nontemporal_store(rhs[c], &o[c * cells]);
        wait_vmcnt<N_T + N_ST>();
      }
    };
    const int pe = p0 + len;
    int p = p0;
    for (; p + 1 < pe; p += 2) {
      step(p, tile0);
      step(p + 1, tile1);
    }
    if (p < pe) step(p, tile0);
    // nothing may still write this block's LDS when the next segment starts
    wait_vmcnt<0>();
    if (G.partial) {
      const double m = block_max<double>(run_max, red);
      if (tid == 0)
        atomicMax(reinterpret_cast<unsigned long long*>(G.partial + inst * G.pstride),
                  (unsigned long long)__double_as_longlong(m));
    }
  }
}

hipError_t launch_flux_rows(FluxMultiJob& job, int ngroups, hipStream_t st) {
  if (ngroups < 1 || ngroups > kFluxGroups) return hipErrorInvalidValue;
  if (job.dyn != DYN_GLIDE_ONLY && job.dyn != DYN_GLIDE_CLIMB) return hipErrorInvalidValue;
  long long steps = 0;
  for (int k = 0; k < ngroups; ++k) {
    FluxGroup& G = job.g[k];
    if (!flux_rows_ok(G.n) || G.count < 1 || !G.clk || !G.A || !G.G || !G.out || (job.mode && !G.U) ||
        (G.partial && G.pstride < 1))
      return hipErrorInvalidValue;
    G.ntiles = (G.n / rows::TC) * G.n;  // row-steps per instance
    G.item0 = (int)steps;
    steps += (long long)G.ntiles * G.count;
  }
  if (steps > (1LL << 30)) return hipErrorInvalidValue;
  job.ngroups = ngroups;
  job.nitems = (int)steps;
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  if (cus[dev] <= 0) {
    int c = 0;
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
    cus[dev] = c;
  }
  const int per_cu = job.per_cu >= 1 && job.per_cu <= 4 ? job.per_cu : 2;
  long long grid = (long long)per_cu * cus[dev];
  // at least 16 rows per block
  if (grid > steps / 16) grid = steps / 16 > 0 ? steps / 16 : 1;
  const dim3 g((unsigned)grid), b(rows::NT);
  if (job.dyn == DYN_GLIDE_ONLY) {
    if (job.mode) hipLaunchKernelGGL((k_flux_rows<DYN_GLIDE_ONLY, true>), g, b, 0, st, job);
    else hipLaunchKernelGGL((k_flux_rows<DYN_GLIDE_ONLY, false>), g, b, 0, st, job);
  } else {
    if (job.mode) hipLaunchKernelGGL((k_flux_rows<DYN_GLIDE_CLIMB, true>), g, b, 0, st, job);
    else hipLaunchKernelGGL((k_flux_rows<DYN_GLIDE_CLIMB, false>), g, b, 0, st, job);
  }
  return hipGetLastError();
}

// Planes per marching block: the longest chunk (a power of two, at least 16
// planes) that still gives every CU of the current device per_cu blocks (the
// fp64 kernel's LDS admits one block per CU, the fp32 kernel's two), so
// longer chunks only save prologues.
// The CDD_FLUX_CHUNK environment variable overrides it (0 selects k_flux);
// the kernel marches planes in pairs, so an override must be even.
static int march_chunk(int n, int per_cu = 1) {
  static int env = -2;
  if (env == -2) {
    const char* s = getenv("CDD_FLUX_CHUNK");
    env = s ? atoi(s) : -1;
  }
  if (env == 0) return 0;
  if (env > 0) return (env <= n && n % env == 0 && env % 2 == 0) ? env : 0;
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  if (cus[dev] <= 0) {
    int c = 0;
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
    cus[dev] = c;
  }
  const int ncol = (n / march::TC) * (n / march::TR);
  int chunk = n;
  while (chunk > 16 && ncol * (n / chunk) < per_cu * cus[dev]) chunk /= 2;
  return chunk;
}

// (job.llf selects the LLF instances of the same kernel)
static bool march_ok(const FluxJob& job) {
  return job.d == 3 && !(job.llf && job.fp32) && !job.S9 && !job.Ad && !job.src && !(job.speed && job.mode) &&
         job.count <= 1 && job.tile == 0 && job.n >= 32 &&
         job.dyn != DYN_WOJZZ && (job.mode == 0 || job.U) && march_chunk(job.n) > 0;
}

// Tile shapes (cells of the contiguous x row axis): variant 0 = 32 x 8,
// 1 = 64 x 4, 2 = 16 x 16 where n allows; whole rows for small n.  A 1D
// grid per instance (grid.y = instances); the kernel maps block ids to tiles.
static void flux_tile(int n, int variant, int& tc, int& tr) {
  if (variant == 1 && n >= 64) {
    tc = 64;
    tr = 4;
  } else if (variant == 2 && n >= 16) {
    tc = 16;
    tr = 16;
  } else if (n >= 32) {
    tc = 32;
    tr = 8;
  } else {
    tc = n;
    tr = 256 / n < n ? 256 / n : n;
  }
}

static void flux_geometry(int d, int n, int variant, int count, dim3& g, dim3& b) {
  int tc, tr;
  flux_tile(n, variant, tc, tr);
  b = dim3(tc, tr, 1);
  g = dim3((d == 3 ? n : 1) * (n / tc) * (n / tr), count > 0 ? count : 1, 1);
}

int flux_nblocks_tile(int d, int n, int tile) {
  dim3 g, b;
  flux_geometry(d, n, tile, 1, g, b);
  return (int)g.x;
}

int flux_nblocks(int d, int n) { return flux_nblocks_tile(d, n, 0); }

template <typename T, int TC, int TR>
static void launch_flux_tt(const FluxJob& job, dim3 g, dim3 b) {
  if (job.dyn == DYN_WOJZZ) {  // 2D only (checked at plan creation)
    hipLaunchKernelGGL((k_flux<T, 2, DYN_WOJZZ, TC, TR, true>), g, b, 0, job.stream, job);
    return;
  }
  if (job.d == 3) {
    if (job.dyn == DYN_GLIDE_ONLY) hipLaunchKernelGGL((k_flux<T, 3, DYN_GLIDE_ONLY, TC, TR, true>), g, b, 0, job.stream, job);
    else hipLaunchKernelGGL((k_flux<T, 3, DYN_GLIDE_CLIMB, TC, TR, true>), g, b, 0, job.stream, job);
  } else {
    if (job.dyn == DYN_GLIDE_ONLY) hipLaunchKernelGGL((k_flux<T, 2, DYN_GLIDE_ONLY, TC, TR, true>), g, b, 0, job.stream, job);
    else hipLaunchKernelGGL((k_flux<T, 2, DYN_GLIDE_CLIMB, TC, TR, true>), g, b, 0, job.stream, job);
  }
}

template <int TC, int TR>
static void launch_flux_t(const FluxJob& job, dim3 g, dim3 b) {
  if (job.fp32) launch_flux_tt<float, TC, TR>(job, g, b);
  else launch_flux_tt<double, TC, TR>(job, g, b);
}

// The LLF flux: fp64, one instance, glide + climb or glide only, default tile.
template <int TC, int TR>
static void launch_flux_llf(const FluxJob& job, dim3 g, dim3 b) {
  if (job.d == 3) {
    if (job.dyn == DYN_GLIDE_ONLY)
      hipLaunchKernelGGL((k_flux<double, 3, DYN_GLIDE_ONLY, TC, TR, true, true>), g, b, 0, job.stream, job);
    else hipLaunchKernelGGL((k_flux<double, 3, DYN_GLIDE_CLIMB, TC, TR, true, true>), g, b, 0, job.stream, job);
  } else {
    if (job.dyn == DYN_GLIDE_ONLY)
      hipLaunchKernelGGL((k_flux<double, 2, DYN_GLIDE_ONLY, TC, TR, true, true>), g, b, 0, job.stream, job);
    else hipLaunchKernelGGL((k_flux<double, 2, DYN_GLIDE_CLIMB, TC, TR, true, true>), g, b, 0, job.stream, job);
  }
}

// The flux with the advected disorder scalar: fp64, one instance, default tile.
template <int TC, int TR>
static void launch_flux_sc(const FluxJob& job, dim3 g, dim3 b) {
  if (job.dyn == DYN_WOJZZ) {
    hipLaunchKernelGGL((k_flux<double, 2, DYN_WOJZZ, TC, TR, true, false, true>), g, b, 0, job.stream, job);
  } else if (job.d == 3) {
    if (job.dyn == DYN_GLIDE_ONLY)
      hipLaunchKernelGGL((k_flux<double, 3, DYN_GLIDE_ONLY, TC, TR, true, false, true>), g, b, 0, job.stream, job);
    else
      hipLaunchKernelGGL((k_flux<double, 3, DYN_GLIDE_CLIMB, TC, TR, true, false, true>), g, b, 0, job.stream, job);
  } else {
    if (job.dyn == DYN_GLIDE_ONLY)
      hipLaunchKernelGGL((k_flux<double, 2, DYN_GLIDE_ONLY, TC, TR, true, false, true>), g, b, 0, job.stream, job);
    else
      hipLaunchKernelGGL((k_flux<double, 2, DYN_GLIDE_CLIMB, TC, TR, true, false, true>), g, b, 0, job.stream, job);
  }
}

hipError_t launch_flux(const FluxJob& job, int* nblocks_out) {
  if (job.B) return hipErrorInvalidValue;  // stage inputs are materialised before the flux pass
  if ((job.S9 || job.Ad) && (job.fp32 || job.count > 1 || job.llf)) return hipErrorInvalidValue;
  if (job.Ad && (!job.outd || (job.mode != 0 && !job.Ud))) return hipErrorInvalidValue;
  if (job.Ad) {
    dim3 g, b;
    flux_geometry(job.d, job.n, 0, 1, g, b);
    if (nblocks_out) *nblocks_out = (int)g.x;
    if (b.x == 32 && b.y == 8) launch_flux_sc<32, 8>(job, g, b);
    else if (b.x == 16 && b.y == 16) launch_flux_sc<16, 16>(job, g, b);
    else if (b.x == 8 && b.y == 8) launch_flux_sc<8, 8>(job, g, b);
    else return hipErrorInvalidValue;
    return hipGetLastError();
  }
  if (job.dyn == DYN_WOJZZ && job.d != 2) return hipErrorInvalidValue;
  if (job.count > 1 && job.d != 2) return hipErrorInvalidValue;
  if (job.llf) {
    if (job.fp32 || job.count > 1 || job.dyn == DYN_WOJZZ) return hipErrorInvalidValue;
    if (march_ok(job)) {
      // 3D: the plane-marching kernel with the executable's flux
      const int chunk = march_chunk(job.n, 1);
      const int nb = (job.n / march::TC) * (job.n / march::TR) * (job.n / chunk);
      if (nblocks_out) *nblocks_out = nb;
      const bool comb = job.mode != 0;
#define CDD_MARCH_LLF(DYN_)                                                                                   \
  do {                                                                                                        \
    if (comb)                                                                                                 \
      hipLaunchKernelGGL((k_flux_march<DYN_, false, true, false, true>), dim3(nb), dim3(march::NT), 0,        \
                         job.stream, job, chunk);                                                             \
    else if (job.speed)                                                                                       \
      hipLaunchKernelGGL((k_flux_march<DYN_, false, false, true, true>), dim3(nb), dim3(march::NT), 0,        \
                         job.stream, job, chunk);                                                             \
    else                                                                                                      \
      hipLaunchKernelGGL((k_flux_march<DYN_, false, false, false, true>), dim3(nb), dim3(march::NT), 0,       \
                         job.stream, job, chunk);                                                             \
  } while (0)
      if (job.dyn == DYN_GLIDE_ONLY) CDD_MARCH_LLF(DYN_GLIDE_ONLY);
      else CDD_MARCH_LLF(DYN_GLIDE_CLIMB);
#undef CDD_MARCH_LLF
      return hipGetLastError();
    }
    dim3 g, b;
    flux_geometry(job.d, job.n, 0, 1, g, b);
    if (nblocks_out) *nblocks_out = (int)g.x;
    if (b.x == 32 && b.y == 8) launch_flux_llf<32, 8>(job, g, b);
    else if (b.x == 16 && b.y == 16) launch_flux_llf<16, 16>(job, g, b);
    else if (b.x == 8 && b.y == 8) launch_flux_llf<8, 8>(job, g, b);
    else return hipErrorInvalidValue;
    return hipGetLastError();
  }
  if (march_ok(job)) {
    const int chunk = march_chunk(job.n, job.fp32 ? 2 : 1);
    const int nb = (job.n / march::TC) * (job.n / march::TR) * (job.n / chunk);
    if (nblocks_out) *nblocks_out = nb;
    const bool comb = job.mode != 0;
#define CDD_MARCH(K_, DYN_, CLIMB_)                                                                              \
  do {                                                                                                           \
    if (comb)                                                                                                    \
      hipLaunchKernelGGL((K_<DYN_, CLIMB_, true>), dim3(nb), dim3(march::NT), 0, job.stream, job, chunk);        \
    else if (job.speed)                                                                                          \
      hipLaunchKernelGGL((K_<DYN_, CLIMB_, false, true>), dim3(nb), dim3(march::NT), 0, job.stream, job, chunk); \
    else                                                                                                         \
      hipLaunchKernelGGL((K_<DYN_, CLIMB_, false>), dim3(nb), dim3(march::NT), 0, job.stream, job, chunk);       \
  } while (0)
#define CDD_MARCH_T(K_)                                           \
  do {                                                            \
    if (job.dyn == DYN_GLIDE_ONLY) CDD_MARCH(K_, DYN_GLIDE_ONLY, false); \
    else if (job.Lambda != 0.0) CDD_MARCH(K_, DYN_GLIDE_CLIMB, true);    \
    else CDD_MARCH(K_, DYN_GLIDE_CLIMB, false);                          \
  } while (0)
    if (!job.fp32) CDD_MARCH_T(k_flux_march);
    else CDD_MARCH_T(k_flux_march32);
#undef CDD_MARCH_T
#undef CDD_MARCH
    return hipGetLastError();
  }
  dim3 g, b;
  flux_geometry(job.d, job.n, job.tile, job.count, g, b);
  if (nblocks_out) *nblocks_out = (int)g.x;
  if (b.x == 64 && b.y == 4) launch_flux_t<64, 4>(job, g, b);
  else if (b.x == 32 && b.y == 8) launch_flux_t<32, 8>(job, g, b);
  else if (b.x == 16 && b.y == 16) launch_flux_t<16, 16>(job, g, b);
  else if (b.x == 8 && b.y == 8) launch_flux_t<8, 8>(job, g, b);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

// dt = CFL / (max(gridShape) max speed), clamped to the end time
// (FieldMover.py:94-111), for one instance: a whole block of up to 1024
// threads reduces its nb speed partials (red: 16 doubles of LDS).
__device__ __forceinline__ void finalize_block(double* partial, int nb, Clock* clk, double nmax, double dt_bound,
                                               int clear, double* red) {
  if (clk->done) return;  // uniform over the block
  double m = -INFINITY;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) m = fmax(m, partial[k]);
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = -INFINITY;
    for (int k = 0; k < (int)(blockDim.x + 63) / 64; ++k) mm = fmax(mm, red[k]);
    clk->last_max = mm;
    if (clear) partial[0] = 0.0;  // an atomically maximised slot starts the next step at zero
    if (!(mm < INFINITY)) {  // a non-finite speed: stop before the state is updated
      clk->status = 1;
      clk->done = 1;
      clk->dt = 0.0;
      return;
    }
    double dt = clk->fixed_dt > 0.0 ? clk->fixed_dt : clk->dt_rule ? clk->cfl / mm / nmax : clk->cfl / (nmax * mm);
    // the operator-splitting mover's bound on the adaptive step
    // (FieldMover.py:359-360; the plain TVD mover leaves it commented out)
    if (clk->fixed_dt <= 0.0 && dt_bound > 0.0 && dt > dt_bound) dt = dt_bound;
    const long long si = clk->nsteps - clk->sched_base;
    if (clk->dt_sched && si >= 0 && si < clk->sched_len) dt = clk->dt_sched[si];
    if (clk->t + dt > clk->t_end) dt = clk->t_end - clk->t + (clk->dt_rule ? kME : 0.0);
    clk->dt = dt;
  }
}

// one block per instance
__global__ void __launch_bounds__(1024) k_finalize_dt(double* partial, int nb, Clock* clk, double nmax,
                                                     long long pstride, double dt_bound, int clear) {
  __shared__ double red[16];
  finalize_block(partial + blockIdx.x * pstride, nb, clk + blockIdx.x, nmax, dt_bound, clear, red);
}

// Several instance groups in one launch (a batch's size groups): block b is
// instance b - first[g] of the group g whose range holds it.
struct FinalizeGroups {
  int ng;
  int first[kFluxGroups + 1];
  double* partial[kFluxGroups];
  int nb[kFluxGroups], clear[kFluxGroups];
  long long pstride[kFluxGroups];
  Clock* clk[kFluxGroups];
  double nmax[kFluxGroups];
};
__global__ void __launch_bounds__(1024) k_finalize_dt_groups(FinalizeGroups f) {
  __shared__ double red[16];
  const int b = (int)blockIdx.x;
  int g = 0;
  while (g + 1 < f.ng && b >= f.first[g + 1]) ++g;
  const int i = b - f.first[g];
  finalize_block(f.partial[g] + i * f.pstride[g], f.nb[g], f.clk[g] + i, f.nmax[g], 0.0, f.clear[g], red);
}

// Energy record of the step-start state: the per-block energy partials of
// k_energy summed and appended to the clock's energy log.
__global__ void __launch_bounds__(1024) k_energy_log(const double* partial, int nb, Clock* clk) {
  if (clk->done || !clk->elog) return;
  __shared__ double red[16];
  double e = 0.0;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) e += partial[k];
  for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < (int)(blockDim.x + 63) / 64; ++k) t += red[k];
    const long long li = clk->nsteps - clk->elog_base;
    if (li >= 0 && li < clk->elog_cap) clk->elog[li] = t;
  }
}

hipError_t launch_energy_log(const double* partial, int nb, Clock* clk, hipStream_t st) {
  hipLaunchKernelGGL(k_energy_log, dim3(1), dim3(1024), 0, st, partial, nb, clk);
  return hipGetLastError();
}

// Large-k instability measure: the mode (N/2, ..., N/2) of the FFT of |rho|,
// which is sum over cells of (-1)^(i+j[+k]) |rho| (per-block partials).
template <typename T>
__global__ void __launch_bounds__(256) k_nyquist_mode(const T* rho, int d, int n, long long cells, double* partial) {
  __shared__ double red[4];
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (idx < cells) {
    double s = 0.0;
    for (int c = 0; c < 9; ++c) {
      const double r = (double)rho[c * cells + idx];
      s += r * r;
    }
    int par = (int)(idx % n) + (int)((idx / n) % n);
    if (d == 3) par += (int)(idx / ((long long)n * n));
    v = (par & 1) ? -sqrt(s) : sqrt(s);
  }
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

hipError_t launch_nyquist_mode(const void* rho, int d, int n, long long cells, int fp32, double* partial, int* nb,
                               hipStream_t st) {
  *nb = (int)((cells + 255) / 256);
  if (fp32)
    hipLaunchKernelGGL((k_nyquist_mode<float>), dim3(*nb), dim3(256), 0, st, static_cast<const float*>(rho), d, n,
                       cells, partial);
  else
    hipLaunchKernelGGL((k_nyquist_mode<double>), dim3(*nb), dim3(256), 0, st, static_cast<const double*>(rho), d, n,
                       cells, partial);
  return hipGetLastError();
}

__global__ void k_fixed_dt(Clock* clk, int count) {
  if ((int)threadIdx.x >= count) return;
  clk += threadIdx.x;
  if (clk->done) return;
  double dt = clk->fixed_dt;
  const long long si = clk->nsteps - clk->sched_base;
  if (clk->dt_sched && si >= 0 && si < clk->sched_len) dt = clk->dt_sched[si];
  if (clk->t + dt > clk->t_end) dt = clk->t_end - clk->t + (clk->dt_rule ? kME : 0.0);
  clk->dt = dt;
}

// t += dt after the step; stop once t >= endTime (FieldMover.py:30-36)
__device__ __forceinline__ void advance_one(Clock* clk) {
  if (clk->done) return;
  const long long li = clk->nsteps - clk->log_base;
  if (clk->dt_log && li >= 0 && li < clk->log_cap) clk->dt_log[li] = clk->dt;
  const long long vi = clk->nsteps - clk->vlog_base;
  if (clk->vlog && vi >= 0 && vi < clk->vlog_cap) clk->vlog[vi] = clk->last_max;
  clk->t += clk->dt;
  clk->nsteps += 1;
  if (clk->t >= clk->t_end) clk->done = 1;
}

__global__ void k_advance(Clock* clk, int count) {
  if ((int)threadIdx.x < count) advance_one(clk + threadIdx.x);
}

struct ClockGroups {
  int ng;
  int first[kFluxGroups + 1];
  Clock* clk[kFluxGroups];
};
__global__ void k_advance_groups(ClockGroups c) {
  const int t = (int)threadIdx.x;
  if (t >= c.first[c.ng]) return;
  int g = 0;
  while (g + 1 < c.ng && t >= c.first[g + 1]) ++g;
  advance_one(c.clk[g] + (t - c.first[g]));
}

hipError_t launch_finalize_dt(const double* partial, int nb, Clock* clk, double nmax, hipStream_t st, int count,
                              long long pstride, double dt_bound, int clear) {
  if (count < 1 || count > 1024) return hipErrorInvalidValue;
  if (clear && nb != 1) return hipErrorInvalidValue;
  if (partial)
    hipLaunchKernelGGL(k_finalize_dt, dim3(count), dim3(1024), 0, st, const_cast<double*>(partial), nb, clk, nmax,
                       pstride, dt_bound, clear);
  else hipLaunchKernelGGL(k_fixed_dt, dim3(1), dim3(count), 0, st, clk, count);
  return hipGetLastError();
}

// u += dt L over n values (simpleTVD's updateField, plasticity_kernel.cu:1105-1109,
// 1130), two values per thread
__global__ void __launch_bounds__(256) k_euler(double* __restrict__ u, const double* __restrict__ L, long long n,
                                              const Clock* clk) {
  if (clk->done) return;
  const double dt = clk->dt;
  const long long stride = 2LL * gridDim.x * blockDim.x;
  for (long long i = 2LL * ((long long)blockIdx.x * blockDim.x + threadIdx.x); i < n; i += stride) {
    if (i + 1 < n) {
      double2 a = *reinterpret_cast<const double2*>(u + i);
      const double2 l = *reinterpret_cast<const double2*>(L + i);
      a.x += dt * l.x;
      a.y += dt * l.y;
      *reinterpret_cast<double2*>(u + i) = a;
    } else {
      u[i] += dt * L[i];
    }
  }
}

// out = u + dt L (the advected disorder field's F0 = u + dt L0, formed over
// L0 in place)
__global__ void __launch_bounds__(256) k_axpy_dt(double* out, const double* u, const double* L, long long n,
                                                const Clock* clk) {
  if (clk->done) return;
  const double dt = clk->dt;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = u[i] + dt * L[i];
}

hipError_t launch_axpy_dt(double* out, const double* u, const double* L, long long n, const Clock* clk,
                          hipStream_t st) {
  if (n <= 0) return hipErrorInvalidValue;
  long long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_axpy_dt, dim3((unsigned)blocks), dim3(256), 0, st, out, u, L, n, clk);
  return hipGetLastError();
}

hipError_t launch_euler(double* u, const double* L, long long n, const Clock* clk, hipStream_t st) {
  if (n <= 0 || (reinterpret_cast<uintptr_t>(u) | reinterpret_cast<uintptr_t>(L)) % 16) return hipErrorInvalidValue;
  long long blocks = (n / 2 + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_euler, dim3((unsigned)blocks), dim3(256), 0, st, u, L, n, clk);
  return hipGetLastError();
}

// The executable's vacancy component under forward Euler (VACANCIES,
// plasticity_kernel.cu:692-697 and the axpy of simpleTVD): its rhs is the
// trace of the tensor rhs, c += vbeta dt (L_xx + L_yy + L_zz).
__global__ void __launch_bounds__(256) k_vacancy_euler(double* __restrict__ c, const double* __restrict__ L,
                                                       long long cells, double vbeta, const Clock* clk) {
  if (clk->done) return;
  const double dt = clk->dt;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < cells;
       i += (long long)gridDim.x * blockDim.x)
    c[i] += vbeta * (dt * ((L[i] + L[4 * cells + i]) + L[8 * cells + i]));
}

hipError_t launch_vacancy_euler(double* c, const double* L, long long cells, double vbeta, const Clock* clk,
                                hipStream_t st) {
  long long blocks = (cells + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_vacancy_euler, dim3((unsigned)blocks), dim3(256), 0, st, c, L, cells, vbeta, clk);
  return hipGetLastError();
}

hipError_t launch_advance(Clock* clk, hipStream_t st, int count) {
  if (count < 1 || count > 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(count), 0, st, clk, count);
  return hipGetLastError();
}

hipError_t launch_finalize_dt_groups(int ng, double* const* partial, const int* nb, const long long* pstride,
                                     Clock* const* clk, const double* nmax, const int* count, const int* clear,
                                     hipStream_t st) {
  for (int g0 = 0; g0 < ng; g0 += kFluxGroups) {
    FinalizeGroups f{};
    f.ng = ng - g0 < kFluxGroups ? ng - g0 : kFluxGroups;
    for (int k = 0; k < f.ng; ++k) {
      const int g = g0 + k;
      if (count[g] < 1 || !partial[g] || !clk[g] || (clear[g] && nb[g] != 1)) return hipErrorInvalidValue;
      f.first[k + 1] = f.first[k] + count[g];
      f.partial[k] = partial[g];
      f.nb[k] = nb[g];
      f.clear[k] = clear[g];
      f.pstride[k] = pstride[g];
      f.clk[k] = clk[g];
      f.nmax[k] = nmax[g];
    }
    hipLaunchKernelGGL(k_finalize_dt_groups, dim3(f.first[f.ng]), dim3(1024), 0, st, f);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_advance_groups(int ng, Clock* const* clk, const int* count, hipStream_t st) {
  for (int g0 = 0; g0 < ng; g0 += kFluxGroups) {
    ClockGroups c{};
    c.ng = ng - g0 < kFluxGroups ? ng - g0 : kFluxGroups;
    for (int k = 0; k < c.ng; ++k) {
      if (count[g0 + k] < 1 || !clk[g0 + k]) return hipErrorInvalidValue;
      c.first[k + 1] = c.first[k] + count[g0 + k];
      c.clk[k] = clk[g0 + k];
    }
    if (c.first[c.ng] > 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_advance_groups, dim3(1), dim3(c.first[c.ng]), 0, st, c);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Elastic energy density e = 1/2 S_ijkm sigma_ij sigma_km
// (PlasticityState.py:317-330), summed per block.
template <typename T>
__global__ void __launch_bounds__(256) k_energy(const T* G, const T* beta, long long cells, double mu, double nu,
                                               double lam, double* partial) {
  __shared__ double red[4];
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (idx < cells) {
    double b[9];
    for (int c = 0; c < 9; ++c) b[c] = (double)beta[c * cells + idx];
    const double tr = b[0] + b[4] + b[8];
    double s[9];
    const int SC[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
    for (int c = 0; c < 9; ++c) {
      const int i = c / 3, j = c % 3;
      s[c] = (double)G[SC[c] * cells + idx] - mu * (b[c] + b[3 * j + i]) - (i == j ? lam * tr : 0.0);
    }
    double ss = 0.0, sst = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        ss += s[3 * i + j] * s[3 * i + j];
        sst += s[3 * i + j] * s[3 * j + i];
      }
    const double trs = s[0] + s[4] + s[8];
    e = 0.5 * (2.0 * ss - sst - nu / (1.0 + nu) * trs * trs) / (2.0 * mu);
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

__global__ void __launch_bounds__(1024) k_reduce_sum(const double* partial, int nb, double* out, int accumulate) {
  __shared__ double red[16];
  double m = 0.0;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) m += partial[k];
  for (int off = 32; off > 0; off >>= 1) m += __shfl_xor(m, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < (int)(blockDim.x + 63) / 64; ++k) t += red[k];
    *out = accumulate ? *out + t : t;
  }
}

hipError_t launch_energy(const double* G, const double* beta, long long cells, double mu, double nu, double lam,
                         double* partial, int* nblocks, hipStream_t st, int fp32) {
  const int nb = (int)((cells + 255) / 256);
  *nblocks = nb;
  if (fp32)
    hipLaunchKernelGGL((k_energy<float>), dim3(nb), dim3(256), 0, st, reinterpret_cast<const float*>(G),
                       reinterpret_cast<const float*>(beta), cells, mu, nu, lam, partial);
  else
    hipLaunchKernelGGL((k_energy<double>), dim3(nb), dim3(256), 0, st, G, beta, cells, mu, nu, lam, partial);
  return hipGetLastError();
}

hipError_t launch_reduce_sum(const double* partial, int n, double* out, hipStream_t st, int accumulate) {
  hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(1024), 0, st, partial, n, out, accumulate);
  return hipGetLastError();
}

}  // namespace cdd
