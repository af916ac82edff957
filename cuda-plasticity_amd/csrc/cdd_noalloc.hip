// The stress solve without a plan and without any device allocation, for
// callers that cannot allocate (code under HIP graph capture, serving
// loops): cdd_sigma_noalloc (include/cdd.h).
//
// Its workspace is `scratch`, a static device array of kScratchBytes that
// every GPU the library runs on holds until the process ends:
//   [twiddle tables of every supported side | 6-component half spectrum |
//    6 non-local stress planes]
// The first call on a device looks up that device's address of the array
// (hipGetSymbolAddress) and fills the twiddle tables with one synchronous
// copy; after that, host code uses only the looked-up address, never the
// array's name.  Nothing in the array has to start at zero: the R2C pass
// writes the whole spectrum and the C2R pass the whole stress planes before
// anything reads them.  Calls that overlap on one GPU share the array and
// race.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "../../include/cdd.h"
#include "cdd_kernels.h"

using namespace cdd;

int cdd_internal_fail(int code, const char* msg);

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;
constexpr int kSides[] = {8, 16, 32, 64, 128, 256, 384, 512, 768, 1024};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// byte offset of side n's twiddle table (2 n doubles) in the array
size_t twiddle_offset(int n) {
  size_t off = 0;
  for (int s : kSides) {
    if (s == n) return off;
    off += align256(16 * (size_t)s);
  }
  return SIZE_MAX;
}

size_t twiddle_bytes() {
  size_t off = 0;
  for (int s : kSides) off += align256(16 * (size_t)s);
  return off;
}

std::mutex g_setup;
unsigned char* g_base[kMaxDevices];  // per device: the array's address

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Look up the current device's copy of `scratch` and fill its twiddle
// tables, once per device.
static int scratch_base(int dev, unsigned char** base) {
  if (dev < 0 || dev >= kMaxDevices) return cdd_internal_fail(CDD_ERR_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_setup);
  if (!g_base[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return cdd_internal_fail(CDD_ERR_HIP, "hipGetSymbolAddress of the no-allocation scratch failed");
    std::vector<double> h(twiddle_bytes() / sizeof(double), 0.0);
    for (int s : kSides) fill_twiddles(h.data() + twiddle_offset(s) / sizeof(double), s);
    e = hipMemcpy(ptr, h.data(), twiddle_bytes(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return cdd_internal_fail(CDD_ERR_HIP, "twiddle upload to the no-allocation scratch failed");
    g_base[dev] = static_cast<unsigned char*>(ptr);
  }
  *base = g_base[dev];
  return CDD_OK;
}

extern "C" int cdd_sigma_noalloc(const cdd_config* cfg, const void* beta_v, double t, void* sigma_v, void* stream) {
  const double* beta = static_cast<const double*>(beta_v);
  double* sigma = static_cast<double*>(sigma_v);
  if (!cfg || !beta || !sigma) return cdd_internal_fail(CDD_ERR_ARG, "null argument");
  const int d = cfg->ndim, n = cfg->n;
  bool side_ok = false;
  for (int s : kSides) side_ok |= s == n;
  if ((d != 2 && d != 3) || !side_ok || (d == 3 && (n == 384 || n == 768 || n == 1024)))
    return cdd_internal_fail(CDD_ERR_UNSUPPORTED, "unsupported grid dimension/size");
  if (cfg->precision != CDD_FP64) return cdd_internal_fail(CDD_ERR_UNSUPPORTED, "cdd_sigma_noalloc is float64 only");
  if (!(cfg->mu > 0) || !(cfg->nu < 0.5)) return cdd_internal_fail(CDD_ERR_ARG, "bad mu/nu");
  if (cfg->load_kind < 0 || cfg->load_kind > 4) return cdd_internal_fail(CDD_ERR_ARG, "bad load_kind");
  if (cfg->load_mu < 0 || (cfg->load_mu > 0 && !(cfg->load_nu < 0.5)))
    return cdd_internal_fail(CDD_ERR_ARG, "bad load_mu/load_nu");
  const long long cells = (long long)n * n * (d == 3 ? n : 1);
  const long long H = n / 2 + 1;
  const size_t spec_bytes = align256(16 * 6 * (size_t)H * (size_t)(d == 3 ? (long long)n * n : n));
  const size_t g_bytes = 8 * 6 * (size_t)cells;
  const size_t spec_off = twiddle_bytes();
  const size_t g_off = spec_off + spec_bytes;
  if (g_off + g_bytes > kScratchBytes) {
    char msg[200];
    snprintf(msg, sizeof msg,
             "grid %dD n=%d needs %zu bytes of workspace; the no-allocation scratch holds %zu (use cdd_sigma with a plan)",
             d, n, g_off + g_bytes, kScratchBytes);
    return cdd_internal_fail(CDD_ERR_UNSUPPORTED, msg);
  }
  if (hipSetDevice(cfg->device) != hipSuccess) return cdd_internal_fail(CDD_ERR_HIP, "hipSetDevice failed");
  unsigned char* base = nullptr;
  int rc = scratch_base(cfg->device, &base);
  if (rc) return rc;
  double* G = reinterpret_cast<double*>(base + g_off);
  hipStream_t st = static_cast<hipStream_t>(stream);
  SpectralJob j;
  memset(&j, 0, sizeof j);
  j.d = d;
  j.n = n;
  j.nc = 6;
  j.cells = cells;
  j.src = SrcField<double>{beta, nullptr, nullptr, cells, nullptr, 0};
  j.spec = base + spec_off;
  j.tw = base + twiddle_offset(n);
  j.out = G;
  j.gp = green_params(*cfg);
  j.gp.clk = nullptr;
  j.gp.t = t;
  j.stream = st;
  hipError_t e = launch_spectral(j);
  if (e == hipSuccess) e = launch_sigma_finalize(G, beta, sigma, cells, j.gp.mu, j.gp.lam, st, 0);
  if (e != hipSuccess) return cdd_internal_fail(CDD_ERR_HIP, hipGetErrorString(e));
  return CDD_OK;
}
