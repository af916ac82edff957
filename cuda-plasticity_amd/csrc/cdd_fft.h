// Hand-written complex FFT building blocks for gfx950.
//
// One length-N transform is spread over FFTPlan<N>::T threads of a
// workgroup; each thread keeps V = N / T complex values in registers.  The
// transform is a Stockham auto-sort decomposition N = R_0 R_1 ... R_{s-1}
// (Govindaraju et al., "High performance discrete Fourier transforms on
// graphics processors", SC'08): stage s with span L = R_0 ... R_{s-1}
// reads x[J + m N/R], multiplies by exp(-2 pi i m k / (L R)) with
// k = J mod L, does an in-register radix-R DFT and writes
// y[(J - k) R + k + m L].  Stages exchange data through LDS via a
// caller-supplied exchange object (so each pass chooses its own LDS layout),
// the first stage reads straight from the caller's registers and the last
// stage leaves its results in registers:
//   entry: thread j holds x[j + T q + m N/R_0]     at v[q R_0 + m]
//   exit:  thread j holds X[j + T q + m N/R_last]  at v[q R_last + m]
// Inverse transforms are conj(FFT(conj(x))), applied by the callers.
#pragma once
#include "cdd_common.h"

namespace cdd {

template <int N>
struct FFTPlan;
template <>
struct FFTPlan<8> {
  static constexpr int NS = 1, T = 1, V = 8;
  static constexpr int R[1] = {8};
};
template <>
struct FFTPlan<16> {
  static constexpr int NS = 1, T = 1, V = 16;
  static constexpr int R[1] = {16};
};
template <>
struct FFTPlan<32> {
  static constexpr int NS = 2, T = 4, V = 8;
  static constexpr int R[2] = {8, 4};
};
template <>
struct FFTPlan<64> {
  static constexpr int NS = 2, T = 8, V = 8;
  static constexpr int R[2] = {8, 8};
};
template <>
struct FFTPlan<128> {
  static constexpr int NS = 2, T = 8, V = 16;
  static constexpr int R[2] = {16, 8};
};
template <>
struct FFTPlan<256> {
  static constexpr int NS = 2, T = 16, V = 16;
  static constexpr int R[2] = {16, 16};
};
// 384 and 768 keep 12 values per thread (radix 3, then radix 4 and 2): the
// 2D contraction then fits 256 registers, two waves per SIMD (24 values per
// thread held one 3-wave block per CU; single 384^2 / 768^2 plans step 3%
// faster).  1024 stays at 16 values: 8 values per thread on 128 threads
// spilled in the contraction and measured slower.
template <>
struct FFTPlan<384> {
  static constexpr int NS = 5, T = 32, V = 12;
  static constexpr int R[5] = {3, 4, 4, 4, 2};
};
template <>
struct FFTPlan<512> {
  static constexpr int NS = 3, T = 32, V = 16;
  static constexpr int R[3] = {16, 8, 4};
};
template <>
struct FFTPlan<768> {
  static constexpr int NS = 5, T = 64, V = 12;
  static constexpr int R[5] = {3, 4, 4, 4, 4};
};
template <>
struct FFTPlan<1024> {
  static constexpr int NS = 3, T = 64, V = 16;
  static constexpr int R[3] = {16, 16, 4};
};

// LDS index padding: one spare slot every 16 complex values breaks the
// power-of-two strides of the Stockham exchanges across LDS banks.
__device__ __forceinline__ constexpr int lds_pad(int e) { return e + (e >> 4); }
template <int N>
struct LdsRow {
  static constexpr int P = N + (N >> 4) + 1;
};

// ---------------------------------------------------------------- radix DFTs
template <typename T, int R>
struct DFT;

template <typename T>
struct DFT<T, 2> {
  __device__ __forceinline__ static void run(cx<T>* a) {
    cx<T> t = a[0];
    a[0] = t + a[1];
    a[1] = t - a[1];
  }
};

template <typename T>
struct DFT<T, 3> {
  __device__ __forceinline__ static void run(cx<T>* a) {
    const T c = T(0.86602540378443864676);  // sqrt(3)/2
    cx<T> t = a[1] + a[2];
    cx<T> s = a[1] - a[2];
    cx<T> m = a[0] - t * T(0.5);
    a[0] = a[0] + t;
    cx<T> is = mul_ni(s) * c;  // -i c s
    a[1] = m + is;
    a[2] = m - is;
  }
};

template <typename T>
struct DFT<T, 4> {
  __device__ __forceinline__ static void run(cx<T>* a) {
    cx<T> t0 = a[0] + a[2], t1 = a[0] - a[2];
    cx<T> t2 = a[1] + a[3], t3 = a[1] - a[3];
    a[0] = t0 + t2;
    a[2] = t0 - t2;
    a[1] = t1 + mul_ni(t3);
    a[3] = t1 + mul_pi(t3);
  }
};

template <typename T>
struct DFT<T, 8> {
  __device__ __forceinline__ static void run(cx<T>* a) {
    const T h = T(0.70710678118654752440);
    cx<T> e[4] = {a[0], a[2], a[4], a[6]};
    cx<T> o[4] = {a[1], a[3], a[5], a[7]};
    DFT<T, 4>::run(e);
    DFT<T, 4>::run(o);
    // o[k] *= W8^k
    o[1] = cx<T>{(o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h};
    o[2] = mul_ni(o[2]);
    o[3] = cx<T>{(o[3].y - o[3].x) * h, -(o[3].x + o[3].y) * h};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a[k] = e[k] + o[k];
      a[k + 4] = e[k] - o[k];
    }
  }
};

template <typename T>
struct DFT<T, 16> {
  __device__ __forceinline__ static void run(cx<T>* a) {
    const T c1 = T(0.92387953251128675613);  // cos(pi/8)
    const T s1 = T(0.38268343236508977173);  // sin(pi/8)
    const T h = T(0.70710678118654752440);
    // n = n1 + 4 n2, k = k2 + 4 k1
    cx<T> b[4][4];
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) {
      cx<T> t[4] = {a[n1], a[n1 + 4], a[n1 + 8], a[n1 + 12]};
      DFT<T, 4>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) b[n1][k2] = t[k2];
    }
    // twiddles W16^(n1 k2)
    b[1][1] = cmul(b[1][1], cx<T>{c1, -s1});
    b[1][2] = cmul(b[1][2], cx<T>{h, -h});
    b[1][3] = cmul(b[1][3], cx<T>{s1, -c1});
    b[2][1] = cmul(b[2][1], cx<T>{h, -h});
    b[2][2] = mul_ni(b[2][2]);
    b[2][3] = cmul(b[2][3], cx<T>{-h, -h});
    b[3][1] = cmul(b[3][1], cx<T>{s1, -c1});
    b[3][2] = cmul(b[3][2], cx<T>{-h, -h});
    b[3][3] = cmul(b[3][3], cx<T>{-c1, s1});
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
      cx<T> t[4] = {b[0][k2], b[1][k2], b[2][k2], b[3][k2]};
      DFT<T, 4>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) a[k2 + 4 * k1] = t[k1];
    }
  }
};

// ------------------------------------------------------------ stage driver
template <int N, int S>
struct StageSpan {  // L = R_0 ... R_{S-1}
  static constexpr int L = StageSpan<N, S - 1>::L * FFTPlan<N>::R[S - 1];
};
template <int N>
struct StageSpan<N, 0> {
  static constexpr int L = 1;
};

template <typename T, int N, int S>
__device__ __forceinline__ void fft_stage_compute(cx<T>* v, int j, const cx<T>* __restrict__ tw) {
  constexpr int R = FFTPlan<N>::R[S];
  constexpr int L = StageSpan<N, S>::L;
  constexpr int TT = FFTPlan<N>::T;
  constexpr int V = FFTPlan<N>::V;
  constexpr int TS = N / (L * R);
#pragma unroll
  for (int q = 0; q < V / R; ++q) {
    if constexpr (L > 1) {
      const int J = j + TT * q;
      const int k = J % L;
#pragma unroll
      for (int m = 1; m < R; ++m) v[q * R + m] = cmul(v[q * R + m], tw[m * k * TS]);
    }
    DFT<T, R>::run(&v[q * R]);
  }
}

// Runs stages S..NS-1.  Ex must provide sync(), put(e, val), get(e).
template <typename T, int N, int S, class Ex>
__device__ __forceinline__ void fft_stages(cx<T>* v, int j, const cx<T>* __restrict__ tw, Ex& ex) {
  fft_stage_compute<T, N, S>(v, j, tw);
  if constexpr (S + 1 < FFTPlan<N>::NS) {
    constexpr int R = FFTPlan<N>::R[S];
    constexpr int L = StageSpan<N, S>::L;
    constexpr int R2 = FFTPlan<N>::R[S + 1];
    constexpr int TT = FFTPlan<N>::T;
    constexpr int V = FFTPlan<N>::V;
    ex.sync();
#pragma unroll
    for (int q = 0; q < V / R; ++q) {
      const int J = j + TT * q;
      const int k = J % L;
      const int base = (J - k) * R + k;
#pragma unroll
      for (int m = 0; m < R; ++m) ex.put(base + m * L, v[q * R + m]);
    }
    ex.sync();
#pragma unroll
    for (int q = 0; q < V / R2; ++q) {
#pragma unroll
      for (int m = 0; m < R2; ++m) v[q * R2 + m] = ex.get(j + TT * q + m * (N / R2));
    }
    fft_stages<T, N, S + 1>(v, j, tw, ex);
  }
}

// Element index held in v[i] at entry / exit of the full transform.
template <int N>
__device__ __forceinline__ int fft_in_index(int j, int i) {
  constexpr int R0 = FFTPlan<N>::R[0];
  return j + FFTPlan<N>::T * (i / R0) + (i % R0) * (N / R0);
}
template <int N>
__device__ __forceinline__ int fft_out_index(int j, int i) {
  constexpr int RL = FFTPlan<N>::R[FFTPlan<N>::NS - 1];
  return j + FFTPlan<N>::T * (i / RL) + (i % RL) * (N / RL);
}

template <typename T, int N, class Ex>
__device__ __forceinline__ void fft_forward(cx<T>* v, int j, const cx<T>* __restrict__ tw, Ex& ex) {
  fft_stages<T, N, 0>(v, j, tw, ex);
}

// LDS exchange over a row-major tile: element e of transform `row` lives at
// base[row * P + lds_pad(e)].
template <typename T>
struct RowExchange {
  cx<T>* base;
  __device__ __forceinline__ void sync() { __syncthreads(); }
  __device__ __forceinline__ void put(int e, cx<T> v) { base[lds_pad(e)] = v; }
  __device__ __forceinline__ cx<T> get(int e) { return base[lds_pad(e)]; }
};

// LDS exchange over a column tile: element e of column `col` lives at
// base[e * W + col] (consecutive columns are consecutive in LDS).
template <typename T, int W>
struct ColExchange {
  cx<T>* base;  // already offset by the column
  __device__ __forceinline__ void sync() { __syncthreads(); }
  __device__ __forceinline__ void put(int e, cx<T> v) { base[e * W] = v; }
  __device__ __forceinline__ cx<T> get(int e) { return base[e * W]; }
};

}  // namespace cdd
