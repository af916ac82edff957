// Spectral passes of the stress solve (and the Nye tensor) on gfx950.
//
// Replaces PlasticityState.CalculateSigma(type='k', source='betaE')
// (Plasticity/PlasticityStates/PlasticityState.py:85-127, 357-390) and
// CalculateRhoFourier (:264-304), i.e. 9 numpy rfftn + 81 array expressions
// + 9 irfftn per call.  Here one stress solve is a chain of HBM passes over
// a half spectrum that only ever holds the six symmetric components
// e_mn = (beta_mn + beta_nm) / 2 (the Green's tensor T_ijmn is symmetric in
// m, n, so the antisymmetric part never contributes):
//
//   3D (N^3, spectrum [c][kz][x][y], kz in 0..N/2):
//     r2c_last   : R2C along z of two real components at once, store
//                  transposed so every later pass is unit-stride along y
//     fft_cols   : complex FFT along x (columns of 16 consecutive y)
//     contract   : FFT along y, Green's contraction (+ k = 0 load mode),
//                  inverse FFT along y -- one pass, data stays in LDS/VGPRs
//     fft_cols   : inverse FFT along x
//     c2r_last   : C2R along z, transposed load, writes 6 real planes G
//   2D (N^2, spectrum [c][ky][x]): r2c_last, contract (along x), c2r_last.
//
// G_ij = IFFT(lam d_ij T_ll(k) + mu (T_ij + T_ji)) is the non-local part of
// sigma; the local part -lam tr(beta) d_ij - mu (beta_ij + beta_ji) is added
// by whoever reads beta at the cell anyway (the flux kernel, or
// k_sigma_finalize for the standalone CalculateSigma).
//
// Every pass touches each of its input and output values once, so their
// global loads and stores take the non-temporal cache policy: the lines
// stream through instead of displacing others in L2 (3-4% of a 256^3 RK3
// step: the passes themselves and the flux kernel that follows run faster).
#include "cdd_kernels.h"

namespace cdd {

// ------------------------------------------------------------- tile sizes
template <int N>
struct Tiles {
  static constexpr int TT = FFTPlan<N>::T;
  static constexpr int P = LdsRow<N>::P;
  // rows per r2c/c2r tile (transposed store: TY consecutive complex values)
  // largest power of two <= min(16, 4096 / N, N), so TY divides N
  static constexpr int TY0 = (4096 / N) < 16 ? (4096 / N) : 16;
  static constexpr int TY1 = TY0 >= 16 ? 16 : (TY0 >= 8 ? 8 : (TY0 >= 4 ? 4 : (TY0 >= 2 ? 2 : 1)));
  static constexpr int TY = TY1 < N ? TY1 : N;
  static_assert(N % TY == 0, "row tile must divide N");
  // columns per strided tile
  static constexpr int TW = TY;
};

template <int N, int NC>
struct ContractTiles {
  static constexpr int TT = FFTPlan<N>::T;
  static constexpr int TR0 = 256 / (NC * TT);
  static constexpr int TR = TR0 >= 16 ? 16 : (TR0 >= 8 ? 8 : (TR0 >= 4 ? 4 : (TR0 >= 2 ? 2 : 1)));
};

// Source components of transform component c: SYM (the symmetric part
// e_mn = (beta_mn + beta_nm) / 2, order xx yy zz xy xz yz) averages two beta
// components for the off-diagonal ones; otherwise c itself.  -1: none.
template <bool SYM>
__device__ __forceinline__ void src_comps(int c, int& k0, int& k1) {
  if constexpr (SYM) {
    constexpr int K0[6] = {0, 4, 8, 1, 2, 5}, K1[6] = {-1, -1, -1, 3, 6, 7};
    k0 = K0[c];
    k1 = K1[c];
  } else {
    k0 = c;
    k1 = -1;
  }
}

// --------------------------------------------------------------- r2c_last
// grid: (N / TY, NO, npairs); block: TY * TT.  NO = N in 3D (outer x), 1 in 2D.
// (two waves per SIMD: the LDS tile admits two blocks per CU, and the
// register budget is held to match)
// TY: rows per block (Tiles<N>::TY, or a smaller power of two for more,
// smaller blocks).
template <typename T, int N, bool SYM, int TY = Tiles<N>::TY, bool STRM = false>
__global__ void __launch_bounds__(TY * Tiles<N>::TT, 2)
k_r2c_last(SrcField<T> src, cx<T>* __restrict__ spec, const cx<T>* __restrict__ tw, int NO, int NC,
           long long sstride) {
  using TL = Tiles<N>;
  constexpr int TT = TL::TT, P = TL::P, V = FFTPlan<N>::V, H = N / 2 + 1;
  // 3D: blockIdx.y is the outer x index; batched 2D: it is the instance
  const int inst = (NO == 1) ? (int)blockIdx.y : 0;
  const int o = (NO == 1) ? 0 : (int)blockIdx.y;
  if (inst) {
    src.A += inst * src.istride;
    if (src.B) src.B += inst * src.istride;
    if (src.W) src.W += inst * src.istride;
    spec += inst * sstride;
  }
  if (src.clk) src.clk += inst;
  if (src.clk && src.clk->done) return;
  __shared__ cx<T> lds[TY * P];
  const int tid = threadIdx.x;
  const int r = tid / TT, j = tid % TT;
  const int r0 = blockIdx.x * TY;
  // transform pairs: SYM pairs a diagonal with an off-diagonal component,
  // (xx, yz), (yy, xz), (zz, xy), so every block reads three beta components
  const int ca = SYM ? (int)blockIdx.z : 2 * (int)blockIdx.z;
  const int cb = SYM ? 5 - (int)blockIdx.z : 2 * (int)blockIdx.z + 1;
  const bool hasb = cb < NC;
  const T dt = (src.B && src.clk) ? T(src.clk->dt) : T(0);

  // Source slots: SYM reads the diagonal component ca and the two beta
  // components averaged into the off-diagonal cb; otherwise ca and cb.
  // Every load comes before the W stores (W may alias B: F0 formed in place
  // over L0), so the loads of all slots can be in flight together; where
  // that would not fit the register budget (V >= 24) the slots are loaded
  // one after another.  Each source component is read (and W written) by
  // exactly one pair block of a row tile.
  constexpr int NS = SYM ? 3 : 2;
  constexpr bool STAGED = 2 * NS * V > 96;
  const long long rowbase = ((long long)o * N + (r0 + r)) * N;
  int ks[NS];
  if constexpr (SYM) {
    int unused;
    src_comps<true>(ca, ks[0], unused);
    src_comps<true>(cb, ks[1], ks[2]);
  } else {
    ks[0] = ca;
    ks[1] = hasb ? cb : -1;
  }
  T a[NS][V];
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    if (ks[q] < 0) continue;
    const T* Ak = src.A + ks[q] * src.cs + rowbase;
#pragma unroll
    for (int i = 0; i < V; ++i) a[q][i] = ld_pol<STRM>(&Ak[fft_in_index<N>(j, i)]);
    if (src.B) {
      const T* Bk = src.B + ks[q] * src.cs + rowbase;
      T b[V];
#pragma unroll
      for (int i = 0; i < V; ++i) b[i] = ld_pol<STRM>(&Bk[fft_in_index<N>(j, i)]);
#pragma unroll
      for (int i = 0; i < V; ++i) a[q][i] += dt * b[i];
    }
    if constexpr (STAGED) asm volatile("" ::: "memory");
  }
  if (src.W) {
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      if (ks[q] < 0) continue;
      T* Wk = src.W + ks[q] * src.cs + rowbase;
#pragma unroll
      for (int i = 0; i < V; ++i) st_pol<STRM>(&Wk[fft_in_index<N>(j, i)], a[q][i]);
    }
  }
  cx<T> v[V];
#pragma unroll
  for (int i = 0; i < V; ++i) {
    v[i].x = a[0][i];
    if constexpr (SYM) v[i].y = T(0.5) * (a[1][i] + a[2][i]);
    else v[i].y = ks[1] >= 0 ? a[1][i] : T(0);
  }
  RowExchange<T> ex{lds + r * P};
  fft_forward<T, N>(v, j, tw, ex);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < V; ++i) lds[r * P + lds_pad(fft_out_index<N>(j, i))] = v[i];
  __syncthreads();
  // separate the two real transforms and store [c][k][o][r] (r contiguous)
  const int rr = tid % TY, g = tid / TY;
#pragma unroll
  for (int m = 0; m < (H + TT - 1) / TT; ++m) {
    const int k = g + m * TT;
    if (k >= H) break;
    cx<T> z1 = lds[rr * P + lds_pad(k)];
    cx<T> z2 = conj(lds[rr * P + lds_pad((N - k) % N)]);
    const long long oa = (((long long)ca * H + k) * NO + o) * N + r0 + rr;
    st_pol<STRM>(&spec[oa], (z1 + z2) * T(0.5));
    if (hasb) {
      const long long ob = (((long long)cb * H + k) * NO + o) * N + r0 + rr;
      st_pol<STRM>(&spec[ob], mul_ni(z1 - z2) * T(0.5));
    }
  }
}

// --------------------------------------------------------------- c2r_last
// Inverse of k_r2c_last: reads [c][k][o][r], writes real out[c][o][r][e].
template <typename T, int N, int TY = Tiles<N>::TY, bool STRM = false>
__global__ void __launch_bounds__(TY * Tiles<N>::TT)
k_c2r_last(const cx<T>* __restrict__ spec, T* __restrict__ out, long long cs, const cx<T>* __restrict__ tw,
           int NO, int NC, const Clock* clk, long long sstride, long long ostride, const T* __restrict__ vac,
           T valpha, const T* __restrict__ valpha_f) {
  using TL = Tiles<N>;
  constexpr int TT = TL::TT, P = TL::P, V = FFTPlan<N>::V, H = N / 2 + 1;
  const int inst = (NO == 1) ? (int)blockIdx.y : 0;
  const int o = (NO == 1) ? 0 : (int)blockIdx.y;
  spec += inst * sstride;
  out += inst * ostride;
  if (clk) clk += inst;
  if (clk && clk->done) return;
  __shared__ cx<T> lds[TY * P];
  const int tid = threadIdx.x;
  const int r0 = blockIdx.x * TY;
  const int ca = 2 * blockIdx.z, cb = 2 * blockIdx.z + 1;
  const bool hasb = cb < NC;
  {
    // all loads of the transposed tile first (one memory round trip)
    constexpr int NK = (H + TT - 1) / TT;
    const int rr = tid % TY, g = tid / TY;
    cx<T> Av[NK], Bv[NK];
#pragma unroll
    for (int m = 0; m < NK; ++m) {
      const int k = g + m * TT;
      const bool in = k < H;
      Av[m] = in ? ld_pol<STRM>(&spec[(((long long)ca * H + k) * NO + o) * N + r0 + rr]) : cx<T>{T(0), T(0)};
      Bv[m] = (in && hasb) ? ld_pol<STRM>(&spec[(((long long)cb * H + k) * NO + o) * N + r0 + rr]) : cx<T>{T(0), T(0)};
    }
#pragma unroll
    for (int m = 0; m < NK; ++m) {
      const int k = g + m * TT;
      if (k >= H) break;
      cx<T> A = Av[m], B = Bv[m];
      if (k == 0 || k == N / 2) {  // irfft ignores the imaginary part here
        A.y = T(0);
        B.y = T(0);
      }
      lds[rr * P + lds_pad(k)] = cx<T>{A.x - B.y, A.y + B.x};
      if (k > 0 && k < N / 2) lds[rr * P + lds_pad(N - k)] = cx<T>{A.x + B.y, B.x - A.y};
    }
  }
  __syncthreads();
  const int r = tid / TT, j = tid % TT;
  cx<T> v[V];
#pragma unroll
  for (int i = 0; i < V; ++i) v[i] = conj(lds[r * P + lds_pad(fft_in_index<N>(j, i))]);
  RowExchange<T> ex{lds + r * P};
  fft_forward<T, N>(v, j, tw, ex);
  const long long rowbase = ((long long)o * N + (r0 + r)) * N;
  if (vac && (ca < 3 || (hasb && cb < 3))) {
    // vacancy coupling of the stepper's stress solves: the diagonal stress
    // planes (xx, yy, zz) get -alpha c (VacancyDynamics.py:24-28), alpha a
    // constant or a field
    T cv[V];
#pragma unroll
    for (int i = 0; i < V; ++i) {
      const long long idx = rowbase + fft_out_index<N>(j, i);
      cv[i] = (valpha_f ? valpha_f[idx] : valpha) * vac[idx];
    }
#pragma unroll
    for (int i = 0; i < V; ++i) {
      const long long idx = rowbase + fft_out_index<N>(j, i);
      out[ca * cs + idx] = ca < 3 ? v[i].x - cv[i] : v[i].x;
      if (hasb) out[cb * cs + idx] = cb < 3 ? -v[i].y - cv[i] : -v[i].y;
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < V; ++i) {
    const long long idx = rowbase + fft_out_index<N>(j, i);
    st_pol<STRM>(&out[ca * cs + idx], v[i].x);  // conj(FFT(conj z)).x
    if (hasb) st_pol<STRM>(&out[cb * cs + idx], -v[i].y);
  }
}

// --------------------------------------------------------------- fft_cols
// In-place complex FFT along x of spec[c][kz][x][y]; grid (N / TW, H, NC).
template <typename T, int N, bool INV, bool STRM = false>
__global__ void __launch_bounds__(Tiles<N>::TW * Tiles<N>::TT)
k_fft_cols(cx<T>* __restrict__ spec, const cx<T>* __restrict__ tw, const Clock* clk) {
  using TL = Tiles<N>;
  constexpr int TT = TL::TT, TW = TL::TW, V = FFTPlan<N>::V, H = N / 2 + 1;
  if (clk && clk->done) return;
  __shared__ cx<T> lds[TW * N];
  const int tid = threadIdx.x;
  const int t = tid % TW, j = tid / TW;
  const int y0 = blockIdx.x * TW;
  const long long base = ((long long)blockIdx.z * H + blockIdx.y) * N * N + y0 + t;
  cx<T> v[V];
#pragma unroll
  for (int i = 0; i < V; ++i) {
    cx<T> a = ld_pol<STRM>(&spec[base + (long long)fft_in_index<N>(j, i) * N]);
    v[i] = INV ? conj(a) : a;
  }
  ColExchange<T, TW> ex{lds + t};
  fft_forward<T, N>(v, j, tw, ex);
#pragma unroll
  for (int i = 0; i < V; ++i) {
    cx<T> a = v[i];
    st_pol<STRM>(&spec[base + (long long)fft_out_index<N>(j, i) * N], INV ? conj(a) : a);
  }
}

// --------------------------------------------------------------- contract
__device__ __forceinline__ double kwrap(int i, int n) {
  // 2 pi n_wrapped with the Nyquist index kept positive
  // (Plasticity/GridArray/FourierSpaceTools.py:5-7, 29-30, 47-49)
  return 6.283185307179586 * (double)(i <= n / 2 ? i : i - n);
}

// Symmetric component c = 0..5 (xx yy zz xy xz yz) -> (i, j).
__device__ __forceinline__ constexpr int sym_i(int c) { return c < 3 ? c : (c == 5 ? 1 : 0); }
__device__ __forceinline__ constexpr int sym_j(int c) { return c < 3 ? c : (c == 3 ? 1 : 2); }

// Green's contraction at one k, all six symmetric output components.
// e[6] = FFT of (xx, yy, zz, xy, xz, yz) of the symmetric part of beta.
template <typename T>
__device__ __forceinline__ void green_sigma(const cx<T>* e, cx<T>* out, T kx, T ky, T kz, const GreenParams& gp) {
  const T kSq = kx * kx + ky * ky + kz * kz;
  const T K2 = kSq + T(kME);
  const T K4 = kSq * kSq + T(kME);
  const cx<T> t = e[0] + e[1] + e[2];
  cx<T> p[3];
  p[0] = e[0] * kx + e[3] * ky + e[4] * kz;
  p[1] = e[3] * kx + e[1] * ky + e[5] * kz;
  p[2] = e[4] * kx + e[5] * ky + e[2] * kz;
  const cx<T> q = p[0] * kx + p[1] * ky + p[2] * kz;
  const T iK2 = T(1) / K2, iK4 = T(1) / K4;
  const cx<T> a1 = t * (T(gp.A) * iK2) - q * (T(gp.C) * iK4);
  const cx<T> Tll = a1 * kSq + q * (T(2) * iK2);
  const T k3[3] = {kx, ky, kz};
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const int i = sym_i(c), j = sym_j(c);
    // S_ij = T_ij + T_ji
    const cx<T> S = a1 * (T(2) * k3[i] * k3[j]) + (p[j] * k3[i] + p[i] * k3[j]) * (T(2) * iK2);
    cx<T> o = S * T(gp.mu);
    if (i == j) o = o + Tll * T(gp.lam);
    out[c] = o * T(gp.norm);
  }
}

// The whole stress of CalculateSigma(source='betaP') (PlasticityState.py:
// 151-182) at one k != 0, times 1/N^d: the M tensor summed over the
// symmetric part e of beta,
//   sigma_ij = a1 ((d_ij q + k_i k_j t) / K2 - d_ij t) - 2 mu e_ij
//              - a2 k_i k_j q / K4 + 2 mu (k_i p_j + k_j p_i) / K2,
// t = tr e, p = e k, q = k.e.k, a1 = 2 mu nu / (1 - nu) = 2 mu A,
// a2 = 2 mu / (1 - nu) = 2 mu C, K2 = kSq + ME, K4 = kSqSq + ME.  Unlike the
// fixed-volume form it keeps the local terms, so where the ME
// regularisation dominates (the sinktools' Nyquist wavenumbers, N sin(pi))
// the two differ as the reference's sources do.
template <typename T>
__device__ __forceinline__ void green_sigma_M(const cx<T>* e, cx<T>* out, T kx, T ky, T kz, const GreenParams& gp) {
  const T kSq = kx * kx + ky * ky + kz * kz;
  const T iK2 = T(1) / (kSq + T(kME)), iK4 = T(1) / (kSq * kSq + T(kME));
  const cx<T> t = e[0] + e[1] + e[2];
  cx<T> p[3];
  p[0] = e[0] * kx + e[3] * ky + e[4] * kz;
  p[1] = e[3] * kx + e[1] * ky + e[5] * kz;
  p[2] = e[4] * kx + e[5] * ky + e[2] * kz;
  const cx<T> q = p[0] * kx + p[1] * ky + p[2] * kz;
  const T mu = T(gp.mu), a1 = T(2) * mu * T(gp.A), a2 = T(2) * mu * T(gp.C);
  const T k3[3] = {kx, ky, kz};
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const int i = sym_i(c), j = sym_j(c);
    const T kk = k3[i] * k3[j];
    cx<T> o = (t * kk) * (a1 * iK2) - e[c] * (T(2) * mu) - q * (a2 * kk * iK4) +
              (p[j] * k3[i] + p[i] * k3[j]) * (T(2) * mu * iK2);
    if (i == j) o = o + (q * iK2 - t) * a1;
    out[c] = o * T(gp.norm);
  }
}

// Transpose of green_sigma w.r.t. the Frobenius product of symmetric
// matrices: with Gb (the 6 components of a symmetric cotangent), tg = tr Gb,
// gk = Gb k, qg = k.Gb.k,
//   (K^T Gb)_ij = d_ij (lam tg kSq A/K2 + 2 mu qg A/K2)
//               + k_i k_j (-lam tg kSq C/K4 + 2 lam tg/K2 - 2 mu qg C/K4)
//               + 2 mu/K2 (gk_i k_j + k_i gk_j).
// (It equals green_sigma up to the ME regularisation of K2, K4.)
template <typename T>
__device__ __forceinline__ void green_sigma_T(const cx<T>* g, cx<T>* out, T kx, T ky, T kz, const GreenParams& gp) {
  const T kSq = kx * kx + ky * ky + kz * kz;
  const T K2 = kSq + T(kME);
  const T K4 = kSq * kSq + T(kME);
  const T iK2 = T(1) / K2, iK4 = T(1) / K4;
  const cx<T> tg = g[0] + g[1] + g[2];
  cx<T> gk[3];
  gk[0] = g[0] * kx + g[3] * ky + g[4] * kz;
  gk[1] = g[3] * kx + g[1] * ky + g[5] * kz;
  gk[2] = g[4] * kx + g[5] * ky + g[2] * kz;
  const cx<T> qg = gk[0] * kx + gk[1] * ky + gk[2] * kz;
  const T lam = T(gp.lam), mu = T(gp.mu), A = T(gp.A), C = T(gp.C);
  const cx<T> kk = tg * (-lam * kSq * C * iK4 + T(2) * lam * iK2) - qg * (T(2) * mu * C * iK4);
  const cx<T> dg = tg * (lam * kSq * A * iK2) + qg * (T(2) * mu * A * iK2);
  const T k3[3] = {kx, ky, kz};
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const int i = sym_i(c), j = sym_j(c);
    cx<T> o = kk * (k3[i] * k3[j]) + (gk[i] * k3[j] + gk[j] * k3[i]) * (T(2) * mu * iK2);
    if (i == j) o = o + dg;
    out[c] = o * T(gp.norm);
  }
}

// The sinktools' wavenumber of index i (FourierSpaceTools(gridShape,
// func=sin), PlasticityState.py:35-36): N sin(2 pi i / N), the argument
// rounded as numpy forms it.
__device__ __forceinline__ double ksin(int i, int n) {
  return (double)n * sin((6.283185307179586 * (double)i) / (double)n);
}

// Free-boundary stress of a Nye tensor (CalculateSigma(source='rho'),
// PlasticityState.py:128-150), times 1/N^d, six symmetric outputs:
//   sigma_ij = (-i mu / K2) [X_ij + X_ji + 2C (k_i k_j / K2 - d_ij) tr X],
//   X_ab = sum_ns eps_bns rho_an k_s (row a of rho crossed with k),
// K2 = kSq + ME, 2C = 2 / (1 - nu): the reference's
// sum_mn sum_s (-i mu k_s / K2)(eps_ins d_jm + eps_jns d_im
//   + 2 eps_mns / (1 - nu)(k_i k_j / K2 - d_ij)) rho_mn, summed per term.
template <typename T>
__device__ __forceinline__ void green_sigma_rho(const cx<T>* r, cx<T>* out, T kx, T ky, T kz, const GreenParams& gp) {
  const T kSq = kx * kx + ky * ky + kz * kz;
  const T iK2 = T(1) / (kSq + T(kME));
  const T k3[3] = {kx, ky, kz};
  cx<T> X[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    X[a][0] = r[3 * a + 1] * kz - r[3 * a + 2] * ky;
    X[a][1] = r[3 * a + 2] * kx - r[3 * a + 0] * kz;
    X[a][2] = r[3 * a + 0] * ky - r[3 * a + 1] * kx;
  }
  const cx<T> trX = X[0][0] + X[1][1] + X[2][2];
  const T c2 = T(2) * T(gp.C);
  const T f = T(gp.mu) * iK2 * T(gp.norm);
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const int i = sym_i(c), j = sym_j(c);
    const T g = k3[i] * k3[j] * iK2 - (i == j ? T(1) : T(0));
    const cx<T> s = X[i][j] + X[j][i] + trX * (c2 * g);
    out[c] = mul_pi(s) * (-f);  // -i mu / K2
  }
#pragma unroll
  for (int c = 6; c < 9; ++c) out[c] = cx<T>{T(0), T(0)};
}

// Nye tensor rho_ij = -eps_ilm d_l beta_mj in k space (times 1/N^d), all
// nine components.
template <typename T>
__device__ __forceinline__ void green_rho(const cx<T>* b, cx<T>* out, T kx, T ky, T kz, T norm) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    out[j] = mul_pi(b[3 + j] * kz - b[6 + j] * ky) * norm;
    out[3 + j] = mul_pi(b[6 + j] * kx - b[j] * kz) * norm;
    out[6 + j] = mul_pi(b[j] * ky - b[3 + j] * kx) * norm;
  }
}

// The disorder stress of DisorderState.CalculateSigma (DisorderState.py:
// 26-29): sum_{l,t} eps_lit k_l F_tj (times 1/N^d) -- the Nye tensor's
// contraction without its factor i.  With the inverse passes reading the
// kz = 0 and Nyquist planes as real (k_c2r_last), as numpy's irfftn does,
// this is the reference's irfftn(k_l rfftn(rho'_tj)) for every l, t.
template <typename T>
__device__ __forceinline__ void green_curl(const cx<T>* b, cx<T>* out, T kx, T ky, T kz, T norm) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    out[j] = (b[3 + j] * kz - b[6 + j] * ky) * norm;
    out[3 + j] = (b[6 + j] * kx - b[j] * kz) * norm;
    out[6 + j] = (b[j] * ky - b[3 + j] * kx) * norm;
  }
}

// rho'_ij = d rho_ij / rhomod, rhomod = ME + sum_mn 2 rho_mn rho_mn in
// component order (DisorderState.py:19-25), in place.
__global__ void __launch_bounds__(256) k_disorder_weight(double* __restrict__ rho, const double* __restrict__ d,
                                                        long long cells) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= cells) return;
  double r[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) r[c] = rho[c * cells + idx];
  double m = kME;
#pragma unroll
  for (int c = 0; c < 9; ++c) m = m + 2.0 * r[c] * r[c];
  const double dv = d[idx];
#pragma unroll
  for (int c = 0; c < 9; ++c) rho[c * cells + idx] = dv * r[c] / m;
}

hipError_t launch_disorder_weight(double* rho, const double* d, long long cells, hipStream_t st) {
  hipLaunchKernelGGL(k_disorder_weight, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, st, rho, d, cells);
  return hipGetLastError();
}

// Rows of spec laid out as (c * NR + row) * N + e.  grid.x = ceil(NR / TR).
// 3D: row = kz * N + x, e = y;  2D: row = ky, e = x.
// Forward FFT of every component row (one transform per TT threads), then
// the contraction with each thread owning whole k points (the k-dependent
// terms once per k, all outputs at once, in place in LDS), then the
// inverse FFT of every output row.
template <typename T, int N, int NC, int D, bool TRANS = false, bool STRM = false>
__global__ void __launch_bounds__((NC * ContractTiles<N, NC>::TR * ContractTiles<N, NC>::TT))
k_contract(cx<T>* __restrict__ spec, const cx<T>* __restrict__ tw, GreenParams gp, long long sstride) {
  using CT = ContractTiles<N, NC>;
  constexpr int TT = CT::TT, TR = CT::TR, P = LdsRow<N>::P, V = FFTPlan<N>::V, H = N / 2 + 1;
  constexpr int NTH = NC * TR * TT;
  constexpr long long NR = (D == 3) ? (long long)H * N : (long long)H;
  spec += (long long)blockIdx.y * sstride;  // batched 2D instance
  if (gp.clk) gp.clk += blockIdx.y;
  if (gp.clk && gp.clk->done) return;
  __shared__ cx<T> lds[NC * TR * P];
  const int tid = threadIdx.x;
  const int j = tid % TT;
  const int r = (tid / TT) % TR;
  const int c = tid / (TT * TR);
  const long long row = (long long)blockIdx.x * TR + r;
  const bool valid = row < NR;
  const long long rbase = ((long long)c * NR + row) * N;

  cx<T> v[V];
#pragma unroll
  for (int i = 0; i < V; ++i) v[i] = valid ? ld_pol<STRM>(&spec[rbase + fft_in_index<N>(j, i)]) : cx<T>{T(0), T(0)};
  cx<T>* myrow = lds + (c * TR + r) * P;
  RowExchange<T> ex{myrow};
  fft_forward<T, N>(v, j, tw, ex);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < V; ++i) myrow[lds_pad(fft_out_index<N>(j, i))] = v[i];
  __syncthreads();

  const double tnow = gp.clk ? gp.clk->t : gp.t;
  for (int kp = tid; kp < TR * N; kp += NTH) {
    const int rr = kp / N, e = kp % N;
    const long long rowk = (long long)blockIdx.x * TR + rr;
    if (rowk >= NR) continue;
    T kx, ky, kz;
    if constexpr (D == 3) {
      if (gp.sink) {
        kz = T(ksin((int)(rowk / N), N));
        kx = T(ksin((int)(rowk % N), N));
        ky = T(ksin(e, N));
      } else {
        kz = T(6.283185307179586 * (double)(rowk / N));
        kx = T(kwrap((int)(rowk % N), N));
        ky = T(kwrap(e, N));
      }
    } else {
      if (gp.sink) {
        ky = T(ksin((int)rowk, N));
        kx = T(ksin(e, N));
      } else {
        ky = T(6.283185307179586 * (double)rowk);
        kx = T(kwrap(e, N));
      }
      kz = T(0);
    }
    cx<T> in[NC], o[NC];
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) in[cc] = lds[(cc * TR + rr) * P + lds_pad(e)];
    if constexpr (NC == 6) {
      const bool k0 = (rowk == 0) && (e == 0);
      if (gp.mform) {
        if (k0) {
          // the reference subtracts lam tr<beta> d_ij + mu (<beta_ij> + <beta_ji>)
          const cx<T> t = in[0] + in[1] + in[2];
#pragma unroll
          for (int cc = 0; cc < 6; ++cc) {
            cx<T> v = in[cc] * T(-2.0 * gp.mu);
            if (cc < 3) v = v - t * T(gp.lam);
            o[cc] = v * T(gp.norm);
          }
        } else {
          green_sigma_M<T>(in, o, kx, ky, kz, gp);
        }
      } else if (TRANS) {
        if (k0) {
#pragma unroll
          for (int cc = 0; cc < 6; ++cc) o[cc] = cx<T>{T(0), T(0)};
        } else {
          green_sigma_T<T>(in, o, kx, ky, kz, gp);
        }
      } else if (!k0) {
        green_sigma<T>(in, o, kx, ky, kz, gp);
      } else {
        // k = 0: kmask removes the Green's term; the uniform load enters here
        double s[9];
        const double mean_bzz = (double)in[2].x * gp.norm;
        load_shift(gp.load, gp.el, tnow, mean_bzz, s);
        if (gp.log_stress && gp.clk && gp.clk->slog) {
          // stress-strain record of the step start: t, mean sigma (load
          // included; the non-local part has zero mean), mean beta_zz
          const Clock* ck = gp.clk;
          const long long li = ck->nsteps - ck->slog_base;
          if (li >= 0 && li < ck->slog_cap) {
            double em[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) em[q] = (double)in[q].x * gp.norm;
            const double tr = em[0] + em[1] + em[2], lam = gp.el.lam, mu2 = 2.0 * gp.el.mu;
            double* rec = ck->slog + 8 * li;
            rec[0] = tnow;
            rec[1] = s[0] - lam * tr - mu2 * em[0];
            rec[2] = s[4] - lam * tr - mu2 * em[1];
            rec[3] = s[8] - lam * tr - mu2 * em[2];
            rec[4] = s[1] - mu2 * em[3];
            rec[5] = s[2] - mu2 * em[4];
            rec[6] = s[5] - mu2 * em[5];
            rec[7] = mean_bzz;
          }
        }
        constexpr int CS9[6] = {0, 4, 8, 1, 2, 5};
#pragma unroll
        for (int cc = 0; cc < 6; ++cc) o[cc] = cx<T>{T(s[CS9[cc]]), T(0)};
      }
    } else if constexpr (NC == 9) {
      if (gp.rho_sigma) {
        if ((rowk == 0) && (e == 0)) {  // the ktools' kmask
#pragma unroll
          for (int cc = 0; cc < 9; ++cc) o[cc] = cx<T>{T(0), T(0)};
        } else {
          green_sigma_rho<T>(in, o, kx, ky, kz, gp);
        }
      } else if (gp.curl) {
        green_curl<T>(in, o, kx, ky, kz, T(gp.norm));
      } else {
        green_rho<T>(in, o, kx, ky, kz, T(gp.norm));
      }
    } else {
      // diffusion filter of the vacancy concentration (VacancyDynamics.py:
      // 40-70): kc exp(kSq * -gamma * dt), normalised for the inverse pass
      const T kSq = kx * kx + ky * ky + kz * kz;
      const double dtv = gp.clk ? gp.clk->dt : 0.0;
      o[0] = in[0] * (T(exp(((double)kSq * -gp.vgamma) * dtv)) * T(gp.norm));
    }
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) lds[(cc * TR + rr) * P + lds_pad(e)] = conj(o[cc]);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < V; ++i) v[i] = myrow[lds_pad(fft_in_index<N>(j, i))];
  fft_forward<T, N>(v, j, tw, ex);
  if (valid) {
#pragma unroll
    for (int i = 0; i < V; ++i) st_pol<STRM>(&spec[rbase + fft_out_index<N>(j, i)], conj(v[i]));
  }
}

// sigma_ij = G_(ij) - lam tr(beta) d_ij - mu (beta_ij + beta_ji)
template <typename T>
__global__ void k_sigma_finalize(const T* __restrict__ G, const T* __restrict__ beta, T* __restrict__ sig,
                                 long long cells, T mu, T lam) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= cells) return;
  T b[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) b[c] = beta[c * cells + idx];
  const T tr = b[0] + b[4] + b[8];
  constexpr int SC[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    const int i = c / 3, jj = c % 3;
    T s = G[SC[c] * cells + idx] - mu * (b[c] + b[3 * jj + i]);
    if (i == jj) s -= lam * tr;
    sig[c * cells + idx] = s;
  }
}

// The six symmetric planes (xx yy zz xy xz yz) in planes 0-5 of s,
// expanded in place into the nine components c = 3 i + j.
__global__ void __launch_bounds__(256) k_sym6_to_9(double* __restrict__ s, long long cells) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= cells) return;
  double v[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) v[c] = s[c * cells + idx];
  constexpr int SC[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
#pragma unroll
  for (int c = 0; c < 9; ++c) s[c * cells + idx] = v[SC[c]];
}

hipError_t launch_sym6_to_9(double* s, long long cells, hipStream_t st) {
  hipLaunchKernelGGL(k_sym6_to_9, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, st, s, cells);
  return hipGetLastError();
}

// CalculateRhoSymmetric (PlasticityState.py:232-262): the Nye tensor from
// SymmetricDerivative (NumericalMethods.py:6-13, order 1),
// D_a f = 0.5 N (f(i+1) - f(i-1)) along axis a, N = the last grid side;
// rho_xj = D_y b_zj - D_z b_yj, rho_yj = D_z b_xj - D_x b_zj,
// rho_zj = D_x b_yj - D_y b_xj (2D: D_z = 0).
__global__ void __launch_bounds__(256) k_rho_symmetric(const double* __restrict__ b, double* __restrict__ rho, int d,
                                                      int n, long long cells) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= cells) return;
  long long st[3];
  int co[3];
  if (d == 3) {
    st[0] = (long long)n * n;
    st[1] = n;
    st[2] = 1;
    co[0] = (int)(idx / st[0]);
    co[1] = (int)((idx / n) % n);
    co[2] = (int)(idx % n);
  } else {
    st[0] = n;
    st[1] = 1;
    st[2] = 0;
    co[0] = (int)(idx / n);
    co[1] = (int)(idx % n);
    co[2] = 0;
  }
  const double h = 0.5 * (double)n;
  // offsets of the +1 and -1 neighbours along each axis (periodic)
  long long op[3], om[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (a >= d) {
      op[a] = om[a] = 0;
      continue;
    }
    op[a] = (co[a] == n - 1 ? -(long long)(n - 1) : 1) * st[a];
    om[a] = (co[a] == 0 ? (long long)(n - 1) : -1) * st[a];
  }
  auto D = [&](int a, int c) -> double {
    if (a >= d) return 0.0;
    const double* f = b + (long long)c * cells + idx;
    return h * (f[op[a]] - f[om[a]]);
  };
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double rx, ry, rz;
    if (d == 3) {
      rx = D(2, 3 + j) - D(1, 6 + j);
      ry = D(0, 6 + j) - D(2, j);
      rz = D(1, j) - D(0, 3 + j);
    } else {
      rx = -D(1, 6 + j);
      ry = D(0, 6 + j);
      rz = D(1, j) - D(0, 3 + j);
    }
    rho[(long long)j * cells + idx] = rx;
    rho[(long long)(3 + j) * cells + idx] = ry;
    rho[(long long)(6 + j) * cells + idx] = rz;
  }
}

hipError_t launch_rho_symmetric(const double* b, double* rho, int d, int n, long long cells, hipStream_t st) {
  hipLaunchKernelGGL(k_rho_symmetric, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, st, b, rho, d, n, cells);
  return hipGetLastError();
}

// ------------------------------------------------------------- launchers
// The job's SrcField<double> carries the source pointers; for fp32 jobs they
// hold float data and are reinterpreted here.
template <typename T>
static SrcField<T> src_as(const SrcField<double>& s) {
  return SrcField<T>{reinterpret_cast<const T*>(s.A), reinterpret_cast<const T*>(s.B), s.clk, s.cs,
                     reinterpret_cast<T*>(s.W), s.istride};
}

template <typename T, int N, bool STRM>
static hipError_t launch_spectral_tp(const SpectralJob& job) {
  using TL = Tiles<N>;
  const SrcField<T> src = src_as<T>(job.src);
  T* out = reinterpret_cast<T*>(job.out);
  const int d = job.d;
  const int NO = d == 3 ? N : 1;
  const int NC = job.nc;
  const int npairs = (NC + 1) / 2;
  const long long cs = job.cells;
  const int H = N / 2 + 1;
  hipStream_t st = job.stream;
  const cx<T>* tw = reinterpret_cast<const cx<T>*>(job.tw);
  cx<T>* spec = reinterpret_cast<cx<T>*>(job.spec);

  const int count = job.count > 0 ? job.count : 1;
  if (count > 1 && d != 2) return hipErrorInvalidValue;
  dim3 g1(N / TL::TY, d == 3 ? NO : count, npairs), b1(TL::TY * TL::TT);
  auto hook = [&](int kind) {
    if (job.hook) job.hook(job.hook_ctx, kind);
  };
  // row-tile variant 1: half the rows per R2C / C2R block (2D fp64 stress
  // solves of sides >= 256; twice the blocks, half the LDS each)
  constexpr int TYH = TL::TY >= 2 ? TL::TY / 2 : 1;
  constexpr bool HALF_OK = N >= 256 && sizeof(T) == 8 && TL::TY >= 2;
  const bool half = HALF_OK && job.stile == 1 && d == 2 && NC == 6 && !job.src_sym6;
  dim3 g1h(N / TYH, count, npairs), b1h(TYH * TL::TT);
  if (half) {
    if constexpr (HALF_OK)
      hipLaunchKernelGGL((k_r2c_last<T, N, true, TYH, STRM>), g1h, b1h, 0, st, src, spec, tw, NO, NC, job.sstride);
  } else if (NC == 6 && !job.src_sym6) {
    hipLaunchKernelGGL((k_r2c_last<T, N, true, TL::TY, STRM>), g1, b1, 0, st, src, spec, tw, NO, NC, job.sstride);
  } else {
    hipLaunchKernelGGL((k_r2c_last<T, N, false, TL::TY, STRM>), g1, b1, 0, st, src, spec, tw, NO, NC, job.sstride);
  }
  hook(KK_R2C);
  if (d == 3) {
    dim3 g2(N / TL::TW, H, NC), b2(TL::TW * TL::TT);
    hipLaunchKernelGGL((k_fft_cols<T, N, false, STRM>), g2, b2, 0, st, spec, tw, job.gp.clk);
    hook(KK_COLS_FWD);
  }
  const long long NR = d == 3 ? (long long)H * N : H;
  if (NC == 6) {
    using CT = ContractTiles<N, 6>;
    dim3 g3((unsigned)((NR + CT::TR - 1) / CT::TR), count), b3(6 * CT::TR * CT::TT);
    if (job.trans) {
      if (d == 3) hipLaunchKernelGGL((k_contract<T, N, 6, 3, true, STRM>), g3, b3, 0, st, spec, tw, job.gp, job.sstride);
      else hipLaunchKernelGGL((k_contract<T, N, 6, 2, true, STRM>), g3, b3, 0, st, spec, tw, job.gp, job.sstride);
    } else if (d == 3) {
      hipLaunchKernelGGL((k_contract<T, N, 6, 3, false, STRM>), g3, b3, 0, st, spec, tw, job.gp, job.sstride);
    } else {
      hipLaunchKernelGGL((k_contract<T, N, 6, 2, false, STRM>), g3, b3, 0, st, spec, tw, job.gp, job.sstride);
    }
  } else if (NC == 9) {
    using CT = ContractTiles<N, 9>;
    dim3 g3((unsigned)((NR + CT::TR - 1) / CT::TR), count), b3(9 * CT::TR * CT::TT);
    if (d == 3) hipLaunchKernelGGL((k_contract<T, N, 9, 3, false, STRM>), g3, b3, 0, st, spec, tw, job.gp, job.sstride);
    else hipLaunchKernelGGL((k_contract<T, N, 9, 2, false, STRM>), g3, b3, 0, st, spec, tw, job.gp, job.sstride);
  } else if (NC == 1) {
    using CT = ContractTiles<N, 1>;
    dim3 g3((unsigned)((NR + CT::TR - 1) / CT::TR), count), b3(CT::TR * CT::TT);
    if (d == 3) hipLaunchKernelGGL((k_contract<T, N, 1, 3, false, STRM>), g3, b3, 0, st, spec, tw, job.gp, job.sstride);
    else hipLaunchKernelGGL((k_contract<T, N, 1, 2, false, STRM>), g3, b3, 0, st, spec, tw, job.gp, job.sstride);
  } else {
    return hipErrorInvalidValue;
  }
  hook(KK_CONTRACT);
  if (d == 3) {
    dim3 g2(N / TL::TW, H, NC), b2(TL::TW * TL::TT);
    hipLaunchKernelGGL((k_fft_cols<T, N, true, STRM>), g2, b2, 0, st, spec, tw, job.gp.clk);
    hook(KK_COLS_INV);
  }
  if (job.vac && (NC != 6 || sizeof(T) != 8)) return hipErrorInvalidValue;
  if (half) {
    if constexpr (HALF_OK)
      hipLaunchKernelGGL((k_c2r_last<T, N, TYH, STRM>), g1h, b1h, 0, st, spec, out, cs, tw, NO, NC, job.gp.clk,
                         job.sstride, job.ostride, reinterpret_cast<const T*>(job.vac), T(job.valpha),
                         reinterpret_cast<const T*>(job.valpha_f));
  } else {
    hipLaunchKernelGGL((k_c2r_last<T, N, TL::TY, STRM>), g1, b1, 0, st, spec, out, cs, tw, NO, NC, job.gp.clk, job.sstride,
                       job.ostride, reinterpret_cast<const T*>(job.vac), T(job.valpha),
                       reinterpret_cast<const T*>(job.valpha_f));
  }
  hook(KK_C2R);
  return hipGetLastError();
}

// A launch whose working set (its spectrum over all components and
// instances, or job.ws_bytes) exceeds the 256 MB infinity cache streams its
// passes (non-temporal loads and stores: 3-4% of a 256^3 RK3 step); smaller
// ones keep the default policy, under which a pass re-reads its
// predecessor's output from that cache (streaming cost 2D 512^2 steps 8%).
template <typename T, int N>
static hipError_t launch_spectral_t(const SpectralJob& job) {
  if constexpr (N >= 128) {
    long long sb = (long long)job.nc * (N / 2 + 1) * (long long)sizeof(cx<T>) * (job.count > 0 ? job.count : 1);
    for (int a = 1; a < job.d; ++a) sb *= N;
    if (job.ws_bytes > 0) sb = job.ws_bytes;
    if (sb > kStreamBytes) return launch_spectral_tp<T, N, true>(job);
  }
  return launch_spectral_tp<T, N, false>(job);
}

template <int N>
static hipError_t launch_spectral_n(const SpectralJob& job) {
  return job.fp32 ? launch_spectral_t<float, N>(job) : launch_spectral_t<double, N>(job);
}

hipError_t launch_spectral(const SpectralJob& job) {
  switch (job.n) {
    case 8: return launch_spectral_n<8>(job);
    case 16: return launch_spectral_n<16>(job);
    case 32: return launch_spectral_n<32>(job);
    case 64: return launch_spectral_n<64>(job);
    case 128: return launch_spectral_n<128>(job);
    case 256: return launch_spectral_n<256>(job);
    case 384: return job.d == 2 ? launch_spectral_n<384>(job) : hipErrorInvalidValue;
    case 512: return launch_spectral_n<512>(job);
    case 768: return job.d == 2 ? launch_spectral_n<768>(job) : hipErrorInvalidValue;
    case 1024: return job.d == 2 ? launch_spectral_n<1024>(job) : hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_sigma_finalize(const double* G, const double* beta, double* sig, long long cells, double mu,
                                 double lam, hipStream_t st, int fp32) {
  const int bs = 256;
  const dim3 g((unsigned)((cells + bs - 1) / bs));
  if (fp32)
    hipLaunchKernelGGL((k_sigma_finalize<float>), g, dim3(bs), 0, st, reinterpret_cast<const float*>(G),
                       reinterpret_cast<const float*>(beta), reinterpret_cast<float*>(sig), cells, (float)mu,
                       (float)lam);
  else
    hipLaunchKernelGGL((k_sigma_finalize<double>), g, dim3(bs), 0, st, G, beta, sig, cells, mu, lam);
  return hipGetLastError();
}

}  // namespace cdd
