// Layout conversion for the reference's on-disk state formats.
//
// The reference's native executable and its initial-condition files store
// every component with all grid axes reversed (Python axis 0 fastest):
// Plasticity/RunCUDA.py:9-22 writes state.betaP[c].transpose() per component
// into the .ics file, and plasticity.cu:239,270 appends [t, 9 N^d] records in
// that layout to the .plas output (assist.h:164-192 ContinueWriteMatrix).
// k_reverse_axes converts between that layout and the plan's NumPy (C
// order) layout; the reversal is an involution, so the same kernel reads and
// writes.
#include "cdd_kernels.h"

namespace cdd {

// For each component and each index j of the middle axis (3D), the (i, k)
// plane is transposed through an LDS tile: reads are coalesced along k
// (contiguous in the input), writes along i (contiguous in the output).
// 2D is the same with no middle axis.  grid: (n / TS, n / TS, ncomp * nj),
// block: (TS, TR); every index stays inside n^d since TS divides n.
template <typename T, int TS, int TR>
__global__ void __launch_bounds__(TS* TR) k_reverse_axes(const T* __restrict__ in, T* __restrict__ out, int n,
                                                        int nj, long long cells) {
  __shared__ T tile[TS][TS + 1];
  const int comp = blockIdx.z / nj, j = blockIdx.z % nj;
  const int k0 = blockIdx.x * TS, i0 = blockIdx.y * TS;
  const long long base = (long long)comp * cells;
  const long long sj = nj > 1 ? (long long)n : 1;  // 3D: (i n + j) n + k;  2D: i n + k
  const long long si = nj > 1 ? (long long)n * n : (long long)n;
  for (int r = threadIdx.y; r < TS; r += TR)
    tile[r][threadIdx.x] = in[base + (long long)(i0 + r) * si + (long long)j * sj + k0 + threadIdx.x];
  __syncthreads();
  for (int r = threadIdx.y; r < TS; r += TR)
    out[base + (long long)(k0 + r) * si + (long long)j * sj + i0 + threadIdx.x] = tile[threadIdx.x][r];
}

template <typename T>
static hipError_t launch_reverse_t(const void* in, void* out, int d, int n, int ncomp, hipStream_t st) {
  const long long cells = d == 3 ? (long long)n * n * n : (long long)n * n;
  const int nj = d == 3 ? n : 1;
  const T* a = static_cast<const T*>(in);
  T* b = static_cast<T*>(out);
  if (n >= 32) {
    dim3 g(n / 32, n / 32, ncomp * nj);
    hipLaunchKernelGGL((k_reverse_axes<T, 32, 8>), g, dim3(32, 8), 0, st, a, b, n, nj, cells);
  } else if (n == 16) {
    dim3 g(1, 1, ncomp * nj);
    hipLaunchKernelGGL((k_reverse_axes<T, 16, 8>), g, dim3(16, 8), 0, st, a, b, n, nj, cells);
  } else if (n == 8) {
    dim3 g(1, 1, ncomp * nj);
    hipLaunchKernelGGL((k_reverse_axes<T, 8, 8>), g, dim3(8, 8), 0, st, a, b, n, nj, cells);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_reverse_axes(const void* in, void* out, int d, int n, int ncomp, int fp32, hipStream_t st) {
  if (in == out || n % 8 != 0 || (n < 32 && n != 8 && n != 16)) return hipErrorInvalidValue;
  return fp32 ? launch_reverse_t<float>(in, out, d, n, ncomp, st) : launch_reverse_t<double>(in, out, d, n, ncomp, st);
}

}  // namespace cdd
