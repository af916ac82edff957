// C ABI of the CDD stepper: plan ownership, state transfer and the RK
// stage sequence (declared in include/cdd.h).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/cdd.h"
#include "cdd_kernels.h"

using namespace cdd;

static thread_local std::string g_err;

static int fail(int code, const char* msg) {
  g_err = msg;
  return code;
}

int cdd_internal_fail(int code, const char* msg) { return fail(code, msg); }

#define HIPCHK(expr)                                                               \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      char b_[256];                                                                \
      snprintf(b_, sizeof b_, "%s failed: %s", #expr, hipGetErrorString(e_));      \
      return fail(CDD_ERR_HIP, b_);                                                \
    }                                                                              \
  } while (0)

#define RCHK(x)          \
  do {                   \
    int rc_ = (x);       \
    if (rc_) return rc_; \
  } while (0)

struct cdd_plan {
  cdd_config cfg;
  int d, n;
  int fp32;   // fields stored and computed in float (CDD_FP32)
  size_t es;  // bytes per field value
  long long cells;
  long long spec_elems;  // complex elements for 9 components
  double* state[2];
  int cur;               // buffer holding the state at steps_base
  long long steps_base;
  long long steps_launched;
  double* L0;
  double* buf;
  double* G;
  double* work9;
  double* partial;
  double* scalar;
  void* spec;
  void* tw;
  Clock* clk;
  hipStream_t own;
  hipStream_t stream;
  int nblocks;
  int npartial;
  GreenParams gp;
  // adjoint workspace (allocated on first use): 6 fields, the 2d*9
  // derivative-cotangent planes, the stress cotangent, per-block partials
  double* adj;
  double* ubar;
  double* sbar;
  double* lpart;
  double* upart;  // uniaxial control: per-block cotangent of betaE_zz
  unsigned long long* codes;
  // segment reverse pass: per-step stage states and stress planes, and the
  // per-block Lambda cotangents of every stage of the segment
  double* seg;
  size_t seg_bytes;
  double* lseg;
  size_t lseg_len;
  // pinned host staging of one .plas record (allocated on first use)
  void* host;
  // per-step energy record: k_energy partials of the step-start state
  double* epart;
  bool elog_on;
  // vacancy coupling (cdd_set_vacancy): the concentration field and the
  // operator-splitting mover's parameters
  double* vac;
  double* valpha_f;  // optional alpha field (cdd_vacancy_alpha)
  int vacancy;
  double valpha, vbeta, vgamma, dt_bound;
  // disorder (cdd_set_disorder): the field d, advected by the flux
  // (DisorderState) or fixed (BetaP_DisorderDynamics); F0 and F1 stage
  // planes of the advected field; the nine planes of the disorder stress
  int disorder, dis_adv;
  double* dis_u;
  double* dis_L;
  double* dis_F1;
  double* dsig;
  // dynamic nucleation (cdd_set_source): a fixed source added to every flux
  // evaluation and the cap on the adaptive time step
  double* nuc_src;
  double nuc_rate, nuc_cap;
  // HIP-event timing of every kernel launched by cdd_step (optional)
  bool timing;
  std::vector<hipEvent_t>* ev;
  std::vector<int>* evkind;
  int nev;
};

static void timing_hook(void* ctx, int kind) {
  cdd_plan* p = (cdd_plan*)ctx;
  if (!p->timing || p->nev >= (int)p->ev->size()) return;
  if (hipEventRecord((*p->ev)[p->nev], p->stream) == hipSuccess) {
    (*p->evkind)[p->nev] = kind;
    p->nev++;
  }
}

static bool valid_n(int d, int n) {
  const int ok3[] = {8, 16, 32, 64, 128, 256, 512};
  const int ok2[] = {8, 16, 32, 64, 128, 256, 384, 512, 768, 1024};
  if (d == 3) {
    for (int v : ok3)
      if (v == n) return true;
  } else if (d == 2) {
    for (int v : ok2)
      if (v == n) return true;
  }
  return false;
}

namespace cdd {
void fill_twiddles(double* h, int n) {
  for (int e = 0; e < n; ++e) {
    long double th = 2.0L * 3.14159265358979323846264338327950288L * (long double)e / (long double)n;
    h[2 * e] = (double)cosl(th);
    h[2 * e + 1] = (double)(-sinl(th));
  }
}

GreenParams green_params(const cdd_config& cfg) {
  GreenParams g;
  memset(&g, 0, sizeof g);
  const double mu = cfg.mu, nu = cfg.nu;
  const double lam = 2.0 * mu * nu / (1.0 - 2.0 * nu);
  g.A = lam / (lam + 2.0 * mu);
  g.C = 2.0 * (lam + mu) / (lam + 2.0 * mu);
  g.mu = mu;
  g.lam = lam;
  const long long cells = (long long)cfg.n * cfg.n * (cfg.ndim == 3 ? cfg.n : 1);
  g.norm = 1.0 / (double)cells;
  g.el = Elastic{mu, nu, lam};
  g.load.kind = cfg.load_kind;
  for (int c = 0; c < 9; ++c) {
    g.load.init[c] = cfg.load_init[c];
    g.load.rate[c] = cfg.load_rate[c];
  }
  g.load.t0 = cfg.load_t0;
  g.load.init_strain = cfg.load_init_strain;
  g.load.uniaxial_rate = cfg.load_uniaxial_rate;
  const double smu = cfg.load_mu > 0 ? cfg.load_mu : mu;
  const double snu = cfg.load_mu > 0 ? cfg.load_nu : nu;
  g.load.smu = smu;
  g.load.slam = 2.0 * smu * snu / (1.0 - 2.0 * snu);
  return g;
}
}  // namespace cdd

static int cur_state(cdd_plan* p) {
  // RK2 ping-pongs between state[0] and state[1]; RK3 and Euler update in place
  if (p->cfg.rk_order != 2) return p->cur;
  return p->cur ^ (int)((p->steps_launched - p->steps_base) & 1);
}

// Re-synchronise the host's view of the ping-pong buffer with the steps the
// device actually executed (steps after `done` are skipped on the device).
static int sync_clock(cdd_plan* p, Clock* host) {
  HIPCHK(hipStreamSynchronize(p->stream));
  Clock c;
  HIPCHK(hipMemcpy(&c, p->clk, sizeof(Clock), hipMemcpyDeviceToHost));
  if (p->cfg.rk_order == 2) {
    long long executed = c.nsteps;
    p->cur = p->cur ^ (int)((executed - p->steps_base) & 1);
    p->steps_base = executed;
    p->steps_launched = executed;
  }
  if (host) *host = c;
  return 0;
}

extern "C" {

const char* cdd_last_error(void) { return g_err.c_str(); }

int cdd_plan_create(const cdd_config* cfg, cdd_plan** out) {
  if (!cfg || !out) return fail(CDD_ERR_ARG, "null argument");
  if (!valid_n(cfg->ndim, cfg->n)) return fail(CDD_ERR_UNSUPPORTED, "unsupported grid dimension/size");
  if (cfg->scheme != CDD_SCHEME_CENTRAL_UPWIND && cfg->scheme != CDD_SCHEME_LLF) return fail(CDD_ERR_ARG, "bad scheme");
  if (cfg->scheme == CDD_SCHEME_CENTRAL_UPWIND && cfg->rk_order != 2 && cfg->rk_order != 3)
    return fail(CDD_ERR_ARG, "rk_order must be 2 or 3");
  if (cfg->scheme == CDD_SCHEME_LLF) {
    // the CUDA executable's simpleTVD (1) and TVD3rd (3), float64 (DOUBLE)
    if (cfg->rk_order != 1 && cfg->rk_order != 3) return fail(CDD_ERR_ARG, "the LLF scheme's rk_order is 1 or 3");
    if (cfg->precision != CDD_FP64) return fail(CDD_ERR_UNSUPPORTED, "the LLF scheme is float64 only");
    if (cfg->dynamics == CDD_GLIDE_CLIMB_WOJZZ)
      return fail(CDD_ERR_UNSUPPORTED, "the LLF scheme supports the glide+climb and glide-only dynamics");
  }
  if (cfg->dynamics != CDD_GLIDE_CLIMB && cfg->dynamics != CDD_GLIDE_ONLY && cfg->dynamics != CDD_GLIDE_CLIMB_WOJZZ)
    return fail(CDD_ERR_ARG, "bad dynamics");
  if (cfg->dynamics == CDD_GLIDE_CLIMB_WOJZZ && cfg->ndim != 2)
    return fail(CDD_ERR_UNSUPPORTED, "the woJzz dynamics is 2D only");
  if (cfg->load_kind < 0 || cfg->load_kind > 4) return fail(CDD_ERR_ARG, "bad load_kind");
  if (!(cfg->mu > 0) || !(cfg->nu < 0.5) || !(cfg->cfl > 0)) return fail(CDD_ERR_ARG, "bad mu/nu/cfl");
  if (cfg->load_mu < 0 || (cfg->load_mu > 0 && !(cfg->load_nu < 0.5)))
    return fail(CDD_ERR_ARG, "bad load_mu/load_nu");
  if (cfg->precision != CDD_FP64 && cfg->precision != CDD_FP32) return fail(CDD_ERR_ARG, "bad precision");
  HIPCHK(hipSetDevice(cfg->device));
  cdd_plan* p = new cdd_plan();
  memset(p, 0, sizeof(*p));
  p->cfg = *cfg;
  p->d = cfg->ndim;
  p->n = cfg->n;
  p->fp32 = cfg->precision == CDD_FP32;
  p->es = p->fp32 ? sizeof(float) : sizeof(double);
  const int n = cfg->n;
  p->cells = (long long)n * n * (p->d == 3 ? n : 1);
  const long long H = n / 2 + 1;
  p->spec_elems = 9LL * H * (p->d == 3 ? (long long)n * n : (long long)n);
  p->nblocks = flux_nblocks(p->d, n);
  p->npartial = p->nblocks > (int)((p->cells + 255) / 256) ? p->nblocks : (int)((p->cells + 255) / 256);
  const size_t fb = p->es * 9 * (size_t)p->cells;
#define ALLOC(ptr, bytes)                                 \
  do {                                                    \
    hipError_t e_ = hipMalloc((void**)&(ptr), (bytes));   \
    if (e_ != hipSuccess) {                               \
      cdd_plan_destroy(p);                                \
      return fail(CDD_ERR_HIP, "hipMalloc failed");       \
    }                                                     \
  } while (0)
  ALLOC(p->state[0], fb);
  if (cfg->rk_order == 2) ALLOC(p->state[1], fb);
  ALLOC(p->L0, fb);
  ALLOC(p->buf, fb);
  ALLOC(p->G, p->es * 6 * (size_t)p->cells);
  ALLOC(p->work9, fb);
  ALLOC(p->partial, sizeof(double) * (size_t)p->npartial);
  ALLOC(p->scalar, sizeof(double) * 8);
  ALLOC(p->spec, p->es * 2 * (size_t)p->spec_elems);
  ALLOC(p->tw, p->es * 2 * (size_t)n);
  ALLOC(p->clk, 2 * sizeof(Clock));  // [1]: saved clock of a segment replay
#undef ALLOC
  if (cfg->rk_order != 2) p->state[1] = p->buf;
  // twiddle table W_n^e = exp(-2 pi i e / n)
  {
    double* h = new double[2 * n];
    fill_twiddles(h, n);
    hipError_t e_;
    if (p->fp32) {
      float* hf = new float[2 * n];
      for (int e = 0; e < 2 * n; ++e) hf[e] = (float)h[e];
      e_ = hipMemcpy(p->tw, hf, sizeof(float) * 2 * n, hipMemcpyHostToDevice);
      delete[] hf;
    } else {
      e_ = hipMemcpy(p->tw, h, sizeof(double) * 2 * n, hipMemcpyHostToDevice);
    }
    delete[] h;
    if (e_ != hipSuccess) {
      cdd_plan_destroy(p);
      return fail(CDD_ERR_HIP, "twiddle upload failed");
    }
  }
  if (hipStreamCreateWithFlags(&p->own, hipStreamNonBlocking) != hipSuccess) {
    cdd_plan_destroy(p);
    return fail(CDD_ERR_HIP, "stream create failed");
  }
  p->stream = p->own;
  HIPCHK(hipMemsetAsync(p->state[0], 0, fb, p->stream));
  HIPCHK(hipMemsetAsync(p->scalar, 0, sizeof(double) * 8, p->stream));
  p->gp = green_params(*cfg);
  Clock c0;
  memset(&c0, 0, sizeof c0);
  c0.t = 0.0;
  c0.t_end = INFINITY;
  c0.cfl = cfg->cfl;
  c0.dt_rule = cfg->scheme == CDD_SCHEME_LLF ? 1 : 0;
  HIPCHK(hipMemcpy(p->clk, &c0, sizeof c0, hipMemcpyHostToDevice));
  HIPCHK(hipStreamSynchronize(p->stream));
  *out = p;
  return CDD_OK;
}

int cdd_plan_destroy(cdd_plan* p) {
  if (!p) return CDD_OK;
  (void)hipSetDevice(p->cfg.device);
  if (p->own) (void)hipStreamSynchronize(p->own);
  (void)hipFree(p->state[0]);
  if (p->cfg.rk_order == 2) (void)hipFree(p->state[1]);
  (void)hipFree(p->L0);
  (void)hipFree(p->buf);
  (void)hipFree(p->G);
  (void)hipFree(p->work9);
  (void)hipFree(p->partial);
  (void)hipFree(p->scalar);
  (void)hipFree(p->spec);
  (void)hipFree(p->tw);
  (void)hipFree(p->clk);
  (void)hipFree(p->adj);
  (void)hipFree(p->ubar);
  (void)hipFree(p->sbar);
  (void)hipFree(p->lpart);
  (void)hipFree(p->upart);
  (void)hipFree(p->codes);
  (void)hipFree(p->seg);
  if (p->host) (void)hipHostFree(p->host);
  (void)hipFree(p->epart);
  (void)hipFree(p->lseg);
  (void)hipFree(p->vac);
  (void)hipFree(p->dis_u);
  (void)hipFree(p->dis_L);
  (void)hipFree(p->dis_F1);
  (void)hipFree(p->dsig);
  (void)hipFree(p->nuc_src);
  (void)hipFree(p->valpha_f);
  if (p->own) (void)hipStreamDestroy(p->own);
  if (p->ev) {
    for (hipEvent_t e : *p->ev) (void)hipEventDestroy(e);
    delete p->ev;
    delete p->evkind;
  }
  delete p;
  return CDD_OK;
}

int cdd_set_stream(cdd_plan* p, void* stream) {
  if (!p) return fail(CDD_ERR_ARG, "null plan");
  // NULL selects the device's default (null) stream, as torch uses it
  p->stream = (hipStream_t)stream;
  return CDD_OK;
}

int cdd_set_state(cdd_plan* p, const void* beta, int on_device) {
  if (!p || !beta) return fail(CDD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  Clock c;
  if (sync_clock(p, &c)) return CDD_ERR_HIP;
  const size_t fb = p->es * 9 * (size_t)p->cells;
  HIPCHK(hipMemcpyAsync(p->state[cur_state(p)], beta, fb, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                        p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return CDD_OK;
}

int cdd_get_state(cdd_plan* p, void* beta, int on_device) {
  if (!p || !beta) return fail(CDD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  if (sync_clock(p, nullptr)) return CDD_ERR_HIP;
  const size_t fb = p->es * 9 * (size_t)p->cells;
  HIPCHK(hipMemcpyAsync(beta, p->state[cur_state(p)], fb, on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                        p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return CDD_OK;
}

int cdd_state_ptr(cdd_plan* p, void** beta_dev) {
  if (!p || !beta_dev) return fail(CDD_ERR_ARG, "null argument");
  if (sync_clock(p, nullptr)) return CDD_ERR_HIP;
  *beta_dev = (void*)p->state[cur_state(p)];
  return CDD_OK;
}

int cdd_set_time(cdd_plan* p, double t, double t_end, double fixed_dt) {
  if (!p) return fail(CDD_ERR_ARG, "null plan");
  HIPCHK(hipSetDevice(p->cfg.device));
  Clock c;
  if (sync_clock(p, &c)) return CDD_ERR_HIP;
  c.t = t;
  c.t_end = t_end;
  c.fixed_dt = fixed_dt;
  c.done = (t >= t_end) ? 1 : 0;
  c.status = 0;
  HIPCHK(hipMemcpy(p->clk, &c, sizeof c, hipMemcpyHostToDevice));
  return CDD_OK;
}

int cdd_get_clock(cdd_plan* p, double* t, double* last_dt, int* done, long long* nsteps) {
  if (!p) return fail(CDD_ERR_ARG, "null plan");
  HIPCHK(hipSetDevice(p->cfg.device));
  Clock c;
  if (sync_clock(p, &c)) return CDD_ERR_HIP;
  if (t) *t = c.t;
  if (last_dt) *last_dt = c.dt;
  if (done) *done = c.done;
  if (nsteps) *nsteps = c.nsteps;
  return CDD_OK;
}

static SpectralJob spectral_job(cdd_plan* p, int nc, const double* A, const double* B, const Clock* clk, double t,
                                double* out) {
  SpectralJob j;
  memset(&j, 0, sizeof j);
  j.d = p->d;
  j.n = p->n;
  j.nc = nc;
  j.cells = p->cells;
  j.src = SrcField<double>{A, B, clk, p->cells, nullptr};
  j.spec = p->spec;
  j.tw = p->tw;
  j.out = out;
  j.gp = p->gp;
  j.gp.clk = clk;
  j.gp.t = t;
  j.stream = p->stream;
  j.hook = p->timing ? timing_hook : nullptr;
  j.hook_ctx = p;
  j.fp32 = p->fp32;
  return j;
}

// The vacancies' chemical stress joins the stepper's (and the dynamics
// flux's) stress solves, not the state's CalculateSigma.
static void vac_stress(const cdd_plan* p, SpectralJob& j) {
  if (p->vacancy) {
    j.vac = p->vac;
    j.valpha = p->valpha;
    j.valpha_f = p->valpha_f;
  }
}

// The disorder stress of the state beta with disorder field dfield, into
// p->dsig (DisorderState.CalculateSigma's added term, DisorderState.py:
// 17-29): the spectral Nye tensor, rho' = d rho / (ME + 2 |rho|^2) in place,
// then sum_lt eps_lit irfftn(k_l rfftn(rho'_tj)) as a nine-component
// spectral pass.  clk: the stepper's clock (the passes skip once it is done).
static int disorder_stress(cdd_plan* p, const double* beta, const double* dfield, const Clock* clk) {
  HIPCHK(launch_spectral(spectral_job(p, 9, beta, nullptr, clk, 0.0, p->dsig)));
  HIPCHK(launch_disorder_weight(p->dsig, dfield, p->cells, p->stream));
  SpectralJob jc = spectral_job(p, 9, p->dsig, nullptr, clk, 0.0, p->dsig);
  jc.gp.curl = 1;
  HIPCHK(launch_spectral(jc));
  return CDD_OK;
}

static FluxJob flux_job(cdd_plan* p) {
  FluxJob f;
  memset(&f, 0, sizeof f);
  f.d = p->d;
  f.n = p->n;
  f.dyn = p->cfg.dynamics;
  f.cells = p->cells;
  f.G = p->G;
  f.Lambda = p->cfg.dynamics == CDD_GLIDE_ONLY ? 0.0 : p->cfg.Lambda;
  f.mu = p->gp.mu;
  f.lam = p->gp.lam;
  f.stream = p->stream;
  f.fp32 = p->fp32;
  f.llf = p->cfg.scheme == CDD_SCHEME_LLF;
  f.src = p->nuc_src;
  f.src_rate = p->nuc_rate;
  return f;
}

int cdd_sigma(cdd_plan* p, const void* beta_v, double t, void* sigma_v) {
  const double* beta = static_cast<const double*>(beta_v);  // float data for fp32 plans
  double* sigma = static_cast<double*>(sigma_v);
  if (!p || !beta || !sigma) return fail(CDD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  SpectralJob j = spectral_job(p, 6, beta, nullptr, nullptr, t, p->G);
  HIPCHK(launch_spectral(j));
  HIPCHK(launch_sigma_finalize(p->G, beta, sigma, p->cells, p->gp.mu, p->gp.lam, p->stream, p->fp32));
  return CDD_OK;
}

int cdd_rho(cdd_plan* p, const void* beta_v, void* rho_v) {
  const double* beta = static_cast<const double*>(beta_v);
  double* rho = static_cast<double*>(rho_v);
  if (!p || !beta || !rho) return fail(CDD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  SpectralJob j = spectral_job(p, 9, beta, nullptr, nullptr, 0.0, rho);
  HIPCHK(launch_spectral(j));
  return CDD_OK;
}

int cdd_sigma_variant(cdd_plan* p, const void* beta_v, int ktype, int source, void* sigma_v) {
  const double* beta = static_cast<const double*>(beta_v);
  double* sigma = static_cast<double*>(sigma_v);
  if (!p || !beta || !sigma) return fail(CDD_ERR_ARG, "null argument");
  if (ktype != CDD_KTYPE_K && ktype != CDD_KTYPE_SINK) return fail(CDD_ERR_ARG, "ktype must be CDD_KTYPE_K or CDD_KTYPE_SINK");
  if (source < CDD_SOURCE_BETAE || source > CDD_SOURCE_RHO) return fail(CDD_ERR_ARG, "bad stress source");
  if (p->fp32) return fail(CDD_ERR_ARG, "cdd_sigma_variant takes float64 plans");
  HIPCHK(hipSetDevice(p->cfg.device));
  if (source == CDD_SOURCE_RHO) {
    // the reference forms rho in real space (CalculateRhoFourier) and
    // transforms it again: two spectral chains
    SpectralJob jr = spectral_job(p, 9, beta, nullptr, nullptr, 0.0, p->work9);
    jr.gp.sink = ktype;
    HIPCHK(launch_spectral(jr));
    SpectralJob js = spectral_job(p, 9, p->work9, nullptr, nullptr, 0.0, sigma);
    js.gp.sink = ktype;
    js.gp.rho_sigma = 1;
    HIPCHK(launch_spectral(js));
    HIPCHK(launch_sym6_to_9(sigma, p->cells, p->stream));
    return CDD_OK;
  }
  if (source == CDD_SOURCE_BETAP) {
    // the M-tensor form, local terms and mean correction included
    SpectralJob jm = spectral_job(p, 6, beta, nullptr, nullptr, 0.0, sigma);
    memset(&jm.gp.load, 0, sizeof jm.gp.load);
    jm.gp.sink = ktype;
    jm.gp.mform = 1;
    HIPCHK(launch_spectral(jm));
    HIPCHK(launch_sym6_to_9(sigma, p->cells, p->stream));
    return CDD_OK;
  }
  // the betaE source ignores type (CalculateBetaElastic() is called without it)
  SpectralJob j = spectral_job(p, 6, beta, nullptr, nullptr, 0.0, p->G);
  memset(&j.gp.load, 0, sizeof j.gp.load);  // CalculateSigma carries no external load
  HIPCHK(launch_spectral(j));
  HIPCHK(launch_sigma_finalize(p->G, beta, sigma, p->cells, p->gp.mu, p->gp.lam, p->stream, 0));
  return CDD_OK;
}

int cdd_rho_variant(cdd_plan* p, const void* beta_v, int ktype, void* rho_v) {
  const double* beta = static_cast<const double*>(beta_v);
  double* rho = static_cast<double*>(rho_v);
  if (!p || !beta || !rho) return fail(CDD_ERR_ARG, "null argument");
  if (ktype != CDD_KTYPE_K && ktype != CDD_KTYPE_SINK) return fail(CDD_ERR_ARG, "ktype must be CDD_KTYPE_K or CDD_KTYPE_SINK");
  if (p->fp32) return fail(CDD_ERR_ARG, "cdd_rho_variant takes float64 plans");
  HIPCHK(hipSetDevice(p->cfg.device));
  SpectralJob j = spectral_job(p, 9, beta, nullptr, nullptr, 0.0, rho);
  j.gp.sink = ktype;
  HIPCHK(launch_spectral(j));
  return CDD_OK;
}

int cdd_rho_symmetric(cdd_plan* p, const void* beta_v, void* rho_v) {
  const double* beta = static_cast<const double*>(beta_v);
  double* rho = static_cast<double*>(rho_v);
  if (!p || !beta || !rho) return fail(CDD_ERR_ARG, "null argument");
  if (p->fp32) return fail(CDD_ERR_ARG, "cdd_rho_symmetric takes float64 plans");
  if (beta == rho) return fail(CDD_ERR_ARG, "cdd_rho_symmetric cannot work in place");
  HIPCHK(hipSetDevice(p->cfg.device));
  HIPCHK(launch_rho_symmetric(beta, rho, p->d, p->n, p->cells, p->stream));
  return CDD_OK;
}

int cdd_flux(cdd_plan* p, const void* beta_v, double t, void* rhs_v, void* speed_v) {
  const double* beta = static_cast<const double*>(beta_v);
  double* rhs = static_cast<double*>(rhs_v);
  double* speed = static_cast<double*>(speed_v);
  if (!p || !beta || !rhs) return fail(CDD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  SpectralJob j = spectral_job(p, 6, beta, nullptr, nullptr, t, p->G);
  vac_stress(p, j);
  HIPCHK(launch_spectral(j));
  FluxJob f = flux_job(p);
  f.A = beta;
  f.out = rhs;
  f.speed = speed;
  f.mode = 0;
  HIPCHK(launch_flux(f, nullptr));
  return CDD_OK;
}

// Buffers of one TVD-RK step: the stage states F0 (= u + dt L(u)) and F1
// (RK3: 3/4 u + 1/4 (F0 + dt L(F0))), the stress planes of each stage and the
// output.  Plain stepping shares one stress buffer and forms F0 in place
// over L0; the adjoint's segment recomputation keeps every stage.
struct StepBufs {
  const double* u;
  double* out;
  double* G[3];
  double* F0;
  double* F1;
};

static int launch_step(cdd_plan* p, const StepBufs& b) {
  const double nmax = (double)p->n;
  double* u = const_cast<double*>(b.u);
  // stage 1: L0 = L(u); dt from the stage-1 speeds
  if (p->timing) timing_hook(p, -1);
  {
    SpectralJob j1 = spectral_job(p, 6, u, nullptr, p->clk, 0.0, b.G[0]);
    j1.gp.log_stress = 1;
    vac_stress(p, j1);
    HIPCHK(launch_spectral(j1));
  }
  if (p->elog_on) {  // energy record of the step-start state (loads included)
    int nbe = 0;
    HIPCHK(launch_energy(b.G[0], u, p->cells, p->gp.mu, p->gp.el.nu, p->gp.lam, p->epart, &nbe, p->stream, p->fp32));
    HIPCHK(launch_energy_log(p->epart, nbe, p->clk, p->stream));
  }
  FluxJob f = flux_job(p);
  f.G = b.G[0];
  f.clk = p->clk;
  f.A = u;
  f.out = p->L0;
  f.mode = 0;
  f.partial = p->partial;
  if (p->disorder) {
    int rc = disorder_stress(p, u, p->dis_u, p->clk);
    if (rc) return rc;
    f.S9 = p->dsig;
    if (p->dis_adv) {
      f.Ad = p->dis_u;
      f.outd = p->dis_L;
    }
  }
  int nb = 0;
  HIPCHK(launch_flux(f, &nb));
  if (p->timing) timing_hook(p, KK_FLUX1);
  HIPCHK(launch_finalize_dt(p->partial, nb, p->clk, nmax, p->stream, 1, 0,
                            p->vacancy ? p->dt_bound : (p->nuc_src ? p->nuc_cap : 0.0)));
  if (p->timing) timing_hook(p, KK_FINALIZE);
  if (p->cfg.rk_order == 1) {
    // forward Euler (simpleTVD, plasticity_kernel.cu:1112-1137), in place
    HIPCHK(launch_euler(b.out, p->L0, 9 * p->cells, p->clk, p->stream));
    if (p->vacancy) {
      // the executable's VACANCIES: c += dt tr(rhs), then diffusion over the step
      HIPCHK(launch_vacancy_euler(p->vac, p->L0, p->cells, p->vbeta, p->clk, p->stream));
      SpectralJob jd = spectral_job(p, 1, p->vac, nullptr, p->clk, 0.0, p->vac);
      jd.gp.vgamma = p->vgamma;
      HIPCHK(launch_spectral(jd));
    }
    HIPCHK(launch_advance(p->clk, p->stream));
    if (p->timing) timing_hook(p, KK_ADVANCE);
    return CDD_OK;
  }
  // stage 2 on F0 = u + dt L0: the R2C pass forms F0 while reading u and
  // L0 for the stress solve and stores it in F0 (which may be L0 itself)
  {
    SpectralJob j2 = spectral_job(p, 6, u, p->L0, p->clk, 0.0, b.G[1]);
    j2.src.W = b.F0;
    vac_stress(p, j2);
    HIPCHK(launch_spectral(j2));
  }
  FluxJob f2 = flux_job(p);
  f2.G = b.G[1];
  f2.clk = p->clk;
  f2.A = b.F0;
  f2.U = u;
  f2.mode = 1;
  if (p->disorder) {
    // the advected field's F0 = d + dt L0_d, formed over L0_d
    const double* d2 = p->dis_u;
    if (p->dis_adv) {
      HIPCHK(launch_axpy_dt(p->dis_L, p->dis_u, p->dis_L, p->cells, p->clk, p->stream));
      d2 = p->dis_L;
    }
    int rc = disorder_stress(p, b.F0, d2, p->clk);
    if (rc) return rc;
    f2.S9 = p->dsig;
    if (p->dis_adv) {
      f2.Ad = p->dis_L;
      f2.Ud = p->dis_u;
      f2.outd = p->cfg.rk_order == 3 ? p->dis_F1 : p->dis_u;  // in place: Ud is read at the own cell only
    }
  }
  if (p->cfg.rk_order == 3) {
    f2.out = b.F1;
    f2.cu = 0.75;
    f2.cf = 0.25;
    HIPCHK(launch_flux(f2, nullptr));
    if (p->timing) timing_hook(p, KK_FLUX2);
    // stage 3 on F1
    SpectralJob j3 = spectral_job(p, 6, b.F1, nullptr, p->clk, 0.0, b.G[2]);
    vac_stress(p, j3);
    HIPCHK(launch_spectral(j3));
    FluxJob f3 = flux_job(p);
    f3.G = b.G[2];
    f3.clk = p->clk;
    f3.A = b.F1;
    f3.U = u;
    f3.out = b.out;  // may be u: u is only read at the thread's own cell
    f3.mode = 1;
    f3.cu = 1.0 / 3.0;
    f3.cf = 2.0 / 3.0;
    if (p->disorder) {
      int rc = disorder_stress(p, b.F1, p->dis_adv ? p->dis_F1 : p->dis_u, p->clk);
      if (rc) return rc;
      f3.S9 = p->dsig;
      if (p->dis_adv) {
        f3.Ad = p->dis_F1;
        f3.Ud = p->dis_u;
        f3.outd = p->dis_u;
      }
    }
    if (p->vacancy) {
      f3.vac = p->vac;
      f3.vbeta = p->vbeta;
    }
    HIPCHK(launch_flux(f3, nullptr));
    if (p->timing) timing_hook(p, KK_FLUX3);
  } else {
    f2.out = b.out;
    f2.cu = 0.5;
    f2.cf = 0.5;
    if (p->vacancy) {
      f2.vac = p->vac;
      f2.vbeta = p->vbeta;
    }
    HIPCHK(launch_flux(f2, nullptr));
    if (p->timing) timing_hook(p, KK_FLUX2);
  }
  if (p->vacancy) {
    // operator splitting: diffuse the vacancies over the step
    // (DiffusionEvolution, VacancyDynamics.py:40-70), in place
    SpectralJob jd = spectral_job(p, 1, p->vac, nullptr, p->clk, 0.0, p->vac);
    jd.gp.vgamma = p->vgamma;
    HIPCHK(launch_spectral(jd));
  }
  HIPCHK(launch_advance(p->clk, p->stream));
  if (p->timing) timing_hook(p, KK_ADVANCE);
  return CDD_OK;
}

int cdd_step(cdd_plan* p, int nsteps) {
  if (!p || nsteps < 0) return fail(CDD_ERR_ARG, "bad argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  for (int s = 0; s < nsteps; ++s) {
    const int ci = cur_state(p);
    StepBufs b;
    b.u = p->state[ci];
    b.out = p->cfg.rk_order != 2 ? p->state[ci] : p->state[ci ^ 1];
    b.G[0] = b.G[1] = b.G[2] = p->G;
    b.F0 = p->L0;
    b.F1 = p->buf;
    int rc = launch_step(p, b);
    if (rc) return rc;
    p->steps_launched += 1;
  }
  return CDD_OK;
}

int cdd_set_vacancy(cdd_plan* p, int enable, double alpha, double beta, double gamma, double dt_bound) {
  if (!p) return fail(CDD_ERR_ARG, "null plan");
  if (!enable) {
    p->vacancy = 0;
    return CDD_OK;
  }
  if (p->fp32) return fail(CDD_ERR_UNSUPPORTED, "the vacancy coupling is float64 only");
  if (p->cfg.scheme == CDD_SCHEME_LLF && p->cfg.rk_order != 1)
    return fail(CDD_ERR_UNSUPPORTED,
                "with the LLF scheme the vacancy coupling follows the executable's forward Euler step (rk_order 1)");
  if (p->disorder) return fail(CDD_ERR_UNSUPPORTED, "the vacancy and disorder couplings do not combine");
  if (p->nuc_src) return fail(CDD_ERR_UNSUPPORTED, "the vacancy coupling excludes the nucleation source");
  if (!(gamma >= 0.0) || !(dt_bound >= 0.0) || !isfinite(alpha) || !isfinite(beta))
    return fail(CDD_ERR_ARG, "bad vacancy parameters");
  HIPCHK(hipSetDevice(p->cfg.device));
  if (!p->vac) {
    HIPCHK(hipStreamSynchronize(p->stream));
    if (hipMalloc((void**)&p->vac, sizeof(double) * (size_t)p->cells) != hipSuccess)
      return fail(CDD_ERR_HIP, "hipMalloc failed (vacancy field)");
    HIPCHK(hipMemsetAsync(p->vac, 0, sizeof(double) * (size_t)p->cells, p->stream));
  }
  p->vacancy = 1;
  p->valpha = alpha;
  p->vbeta = beta;
  p->vgamma = gamma;
  p->dt_bound = dt_bound;
  return CDD_OK;
}

int cdd_set_disorder(cdd_plan* p, const void* dfield, int on_device, int advect) {
  if (!p) return fail(CDD_ERR_ARG, "null plan");
  if (!dfield) {
    p->disorder = 0;
    return CDD_OK;
  }
  if (p->fp32) return fail(CDD_ERR_UNSUPPORTED, "the disorder coupling is float64 only");
  if (p->cfg.scheme != CDD_SCHEME_CENTRAL_UPWIND)
    return fail(CDD_ERR_UNSUPPORTED, "the disorder coupling uses the central-upwind scheme");
  if (p->vacancy) return fail(CDD_ERR_UNSUPPORTED, "the vacancy and disorder couplings do not combine");
  HIPCHK(hipSetDevice(p->cfg.device));
  HIPCHK(hipStreamSynchronize(p->stream));
  const size_t plane = sizeof(double) * (size_t)p->cells;
  if (!p->dis_u && hipMalloc((void**)&p->dis_u, plane) != hipSuccess)
    return fail(CDD_ERR_HIP, "hipMalloc failed (disorder field)");
  if (!p->dsig && hipMalloc((void**)&p->dsig, 9 * plane) != hipSuccess)
    return fail(CDD_ERR_HIP, "hipMalloc failed (disorder stress)");
  if (advect) {
    if (!p->dis_L && hipMalloc((void**)&p->dis_L, plane) != hipSuccess)
      return fail(CDD_ERR_HIP, "hipMalloc failed (disorder stages)");
    if (!p->dis_F1 && hipMalloc((void**)&p->dis_F1, plane) != hipSuccess)
      return fail(CDD_ERR_HIP, "hipMalloc failed (disorder stages)");
  }
  HIPCHK(hipMemcpyAsync(p->dis_u, dfield, plane, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                        p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  p->disorder = 1;
  p->dis_adv = advect ? 1 : 0;
  return CDD_OK;
}

int cdd_disorder_get(cdd_plan* p, void* dfield, int on_device) {
  if (!p || !dfield) return fail(CDD_ERR_ARG, "null argument");
  if (!p->disorder) return fail(CDD_ERR_ARG, "the disorder coupling is not enabled (cdd_set_disorder)");
  HIPCHK(hipSetDevice(p->cfg.device));
  HIPCHK(hipMemcpyAsync(dfield, p->dis_u, sizeof(double) * (size_t)p->cells,
                        on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return CDD_OK;
}

static int dsig_alloc(cdd_plan* p) {
  if (p->fp32) return fail(CDD_ERR_UNSUPPORTED, "the disorder coupling is float64 only");
  if (!p->dsig && hipMalloc((void**)&p->dsig, 9 * sizeof(double) * (size_t)p->cells) != hipSuccess)
    return fail(CDD_ERR_HIP, "hipMalloc failed (disorder stress)");
  return CDD_OK;
}

int cdd_disorder_sigma(cdd_plan* p, const double* beta, const double* dfield, double t, double* sigma) {
  if (!p || !beta || !dfield || !sigma) return fail(CDD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  RCHK(dsig_alloc(p));
  HIPCHK(launch_spectral(spectral_job(p, 6, beta, nullptr, nullptr, t, p->G)));
  HIPCHK(launch_sigma_finalize(p->G, beta, sigma, p->cells, p->gp.mu, p->gp.lam, p->stream, 0));
  RCHK(disorder_stress(p, beta, dfield, nullptr));
  HIPCHK(launch_lincomb(9 * p->cells, 2, sigma, 1.0, p->dsig, 1.0, nullptr, 0.0, nullptr, 0.0, sigma, p->stream));
  return CDD_OK;
}

int cdd_disorder_flux(cdd_plan* p, const double* beta, const double* dfield, double t, double* rhs, double* rhs_d,
                      double* speed) {
  if (!p || !beta || !dfield || !rhs) return fail(CDD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  RCHK(dsig_alloc(p));
  if (p->cfg.scheme != CDD_SCHEME_CENTRAL_UPWIND)
    return fail(CDD_ERR_UNSUPPORTED, "the disorder coupling uses the central-upwind scheme");
  HIPCHK(launch_spectral(spectral_job(p, 6, beta, nullptr, nullptr, t, p->G)));
  RCHK(disorder_stress(p, beta, dfield, nullptr));
  FluxJob f = flux_job(p);
  f.A = beta;
  f.out = rhs;
  f.speed = speed;
  f.mode = 0;
  f.S9 = p->dsig;
  if (rhs_d) {
    f.Ad = dfield;
    f.outd = rhs_d;
  }
  HIPCHK(launch_flux(f, nullptr));
  return CDD_OK;
}

int cdd_set_source(cdd_plan* p, const void* src, int on_device, double rate, double dt_cap) {
  if (!p) return fail(CDD_ERR_ARG, "null plan");
  HIPCHK(hipSetDevice(p->cfg.device));
  if (!src) {
    HIPCHK(hipStreamSynchronize(p->stream));
    (void)hipFree(p->nuc_src);
    p->nuc_src = nullptr;
    p->nuc_rate = p->nuc_cap = 0.0;
    return CDD_OK;
  }
  if (p->fp32) return fail(CDD_ERR_UNSUPPORTED, "the nucleation source is float64 only");
  if (p->vacancy) return fail(CDD_ERR_UNSUPPORTED, "the nucleation source excludes the vacancy coupling");
  if (!isfinite(rate) || !(dt_cap >= 0.0)) return fail(CDD_ERR_ARG, "bad nucleation rate or time-step cap");
  const size_t fb = sizeof(double) * 9 * (size_t)p->cells;
  if (!p->nuc_src && hipMalloc((void**)&p->nuc_src, fb) != hipSuccess)
    return fail(CDD_ERR_HIP, "hipMalloc failed (nucleation source)");
  HIPCHK(hipMemcpyAsync(p->nuc_src, src, fb, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, p->stream));
  if (!on_device) HIPCHK(hipStreamSynchronize(p->stream));
  p->nuc_rate = rate;
  p->nuc_cap = dt_cap;
  return CDD_OK;
}

int cdd_vacancy_alpha(cdd_plan* p, const void* alpha, int on_device) {
  if (!p) return fail(CDD_ERR_ARG, "null plan");
  HIPCHK(hipSetDevice(p->cfg.device));
  if (!alpha) {
    HIPCHK(hipStreamSynchronize(p->stream));
    (void)hipFree(p->valpha_f);
    p->valpha_f = nullptr;
    return CDD_OK;
  }
  if (p->fp32) return fail(CDD_ERR_UNSUPPORTED, "the vacancy coupling is float64 only");
  const size_t b = sizeof(double) * (size_t)p->cells;
  if (!p->valpha_f && hipMalloc((void**)&p->valpha_f, b) != hipSuccess)
    return fail(CDD_ERR_HIP, "hipMalloc failed (alpha field)");
  HIPCHK(hipMemcpyAsync(p->valpha_f, alpha, b, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, p->stream));
  if (!on_device) HIPCHK(hipStreamSynchronize(p->stream));
  return CDD_OK;
}

int cdd_vacancy_set(cdd_plan* p, const void* c, int on_device) {
  if (!p || !c) return fail(CDD_ERR_ARG, "null argument");
  if (!p->vac) return fail(CDD_ERR_ARG, "the vacancy coupling is not enabled (cdd_set_vacancy)");
  HIPCHK(hipSetDevice(p->cfg.device));
  HIPCHK(hipMemcpyAsync(p->vac, c, sizeof(double) * (size_t)p->cells,
                        on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return CDD_OK;
}

int cdd_vacancy_get(cdd_plan* p, void* c, int on_device) {
  if (!p || !c) return fail(CDD_ERR_ARG, "null argument");
  if (!p->vac) return fail(CDD_ERR_ARG, "the vacancy coupling is not enabled (cdd_set_vacancy)");
  HIPCHK(hipSetDevice(p->cfg.device));
  HIPCHK(hipMemcpyAsync(c, p->vac, sizeof(double) * (size_t)p->cells,
                        on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return CDD_OK;
}

int cdd_energy(cdd_plan* p, const void* beta_v, double t, double* energy) {
  const double* beta = static_cast<const double*>(beta_v);
  if (!p || !energy) return fail(CDD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  if (!beta) {
    if (sync_clock(p, nullptr)) return CDD_ERR_HIP;
    beta = p->state[cur_state(p)];
  }
  HIPCHK(launch_spectral(spectral_job(p, 6, beta, nullptr, nullptr, t, p->G)));
  int nb = 0;
  HIPCHK(launch_energy(p->G, beta, p->cells, p->gp.mu, p->gp.el.nu, p->gp.lam, p->partial, &nb, p->stream, p->fp32));
  HIPCHK(launch_reduce_sum(p->partial, nb, p->scalar, p->stream, 0));
  HIPCHK(hipMemcpyAsync(energy, p->scalar, sizeof(double), hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return CDD_OK;
}

// ------------------------------------------------------------- adjoint
static int adj_alloc(cdd_plan* p) {
  if (p->fp32) return fail(CDD_ERR_UNSUPPORTED, "the adjoint is float64 only");
  if (p->vacancy) return fail(CDD_ERR_UNSUPPORTED, "the adjoint does not include the vacancy coupling");
  if (p->disorder) return fail(CDD_ERR_UNSUPPORTED, "the adjoint does not include the disorder coupling");
  if (p->cfg.scheme != CDD_SCHEME_CENTRAL_UPWIND)
    return fail(CDD_ERR_UNSUPPORTED, "the adjoint differentiates the central-upwind scheme");
  if (p->cfg.dynamics == CDD_GLIDE_CLIMB_WOJZZ)
    return fail(CDD_ERR_UNSUPPORTED, "the adjoint supports the glide+climb and glide-only dynamics");
  if (p->adj) return CDD_OK;
  const size_t fb = sizeof(double) * 9 * (size_t)p->cells;
  if (hipMalloc((void**)&p->adj, 6 * fb) != hipSuccess || hipMalloc((void**)&p->ubar, 2 * p->d * fb) != hipSuccess ||
      hipMalloc((void**)&p->sbar, fb) != hipSuccess ||
      hipMalloc((void**)&p->codes, sizeof(unsigned long long) * (size_t)p->cells) != hipSuccess ||
      hipMalloc((void**)&p->lpart, sizeof(double) * (size_t)std::max<long long>((p->cells + 255) / 256,
                                                                                flux_vjp_nblocks(p->d, p->n))) !=
          hipSuccess ||
      hipMalloc((void**)&p->upart, sizeof(double) * (size_t)std::max<long long>((p->cells + 255) / 256,
                                                                                flux_vjp_nblocks(p->d, p->n))) !=
          hipSuccess)
    return fail(CDD_ERR_HIP, "hipMalloc failed (adjoint workspace)");
  return CDD_OK;
}

// The uniaxial strain-rate control's stress depends on <beta_zz>
// (RunLoadingMPI.py:68-83): its transpose adds a uniform term to betabar_zz.
static bool uniaxial(const cdd_plan* p) { return p->cfg.load_kind == CDD_LOAD_UNIAXIAL; }

static double* adjx(cdd_plan* p, int k) { return p->adj + (size_t)k * 9 * p->cells; }

// out = J_L(beta)^T rbar; Lambda cotangent accumulated into scalar[1]
static int flux_vjp_impl(cdd_plan* p, const double* beta, double t, const double* rbar, double* out) {
  HIPCHK(launch_spectral(spectral_job(p, 6, beta, nullptr, nullptr, t, p->G)));
  VjpJob j;
  memset(&j, 0, sizeof j);
  j.d = p->d;
  j.n = p->n;
  j.dyn = p->cfg.dynamics;
  j.cells = p->cells;
  j.beta = beta;
  j.G = p->G;
  j.rbar = rbar;
  j.rscale = 1.0;
  j.dtp = nullptr;
  j.ubar = p->ubar;
  j.codes = p->codes;
  j.sbar = p->sbar;
  j.lpart = p->lpart;
  j.upart = uniaxial(p) ? p->upart : nullptr;
  j.Lambda = p->cfg.dynamics == CDD_GLIDE_ONLY ? 0.0 : p->cfg.Lambda;
  j.mu = p->gp.mu;
  j.lam = p->gp.lam;
  int nb = 0;
  HIPCHK(launch_flux_vjp_cell(j, &nb, p->stream));
  HIPCHK(launch_reduce_sum(p->lpart, nb, p->scalar + 1, p->stream, 1));
  if (j.upart) HIPCHK(launch_reduce_sum(p->upart, nb, p->scalar + 4, p->stream, 0));
  SpectralJob jt = spectral_job(p, 6, p->sbar, nullptr, nullptr, t, p->G);
  jt.trans = 1;
  jt.src_sym6 = 1;
  HIPCHK(launch_spectral(jt));
  GatherJob g;
  memset(&g, 0, sizeof g);
  g.d = p->d;
  g.n = p->n;
  g.cells = p->cells;
  g.codes = p->codes;
  g.ubar = p->ubar;
  g.sbar = p->sbar;
  g.ebar = p->G;
  g.mu = p->gp.mu;
  g.lam = p->gp.lam;
  if (j.upart) {
    g.uzz = p->scalar + 4;
    g.uzz_scale = -1.0 / (double)p->cells;
  }
  g.out = out;
  HIPCHK(launch_flux_vjp_gather(g, p->stream));
  return CDD_OK;
}

static int lincomb(cdd_plan* p, double* out, int nin, const double* a0, double c0, const double* a1 = nullptr,
                   double c1 = 0, const double* a2 = nullptr, double c2 = 0, const double* a3 = nullptr, double c3 = 0) {
  HIPCHK(launch_lincomb(9 * p->cells, nin, a0, c0, a1, c1, a2, c2, a3, c3, out, p->stream));
  return CDD_OK;
}

static int flux_into(cdd_plan* p, const double* beta, double t, double* out) {
  HIPCHK(launch_spectral(spectral_job(p, 6, beta, nullptr, nullptr, t, p->G)));
  FluxJob f = flux_job(p);
  f.A = beta;
  f.out = out;
  f.mode = 0;
  HIPCHK(launch_flux(f, nullptr));
  return CDD_OK;
}

int cdd_flux_vjp(cdd_plan* p, const double* beta, double t, const double* rbar, double* betabar, double* lambar) {
  if (!p || !beta || !rbar || !betabar) return fail(CDD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  RCHK(adj_alloc(p));
  HIPCHK(hipMemsetAsync(p->scalar + 1, 0, sizeof(double), p->stream));
  RCHK(flux_vjp_impl(p, beta, t, rbar, betabar));
  if (lambar) {
    HIPCHK(hipMemcpyAsync(lambar, p->scalar + 1, sizeof(double), hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
  }
  return CDD_OK;
}

int cdd_flux_frozen(cdd_plan* p, const double* beta, double t, unsigned long long* sel, int record, double* rhs) {
  if (!p || !beta || !sel || !rhs) return fail(CDD_ERR_ARG, "null argument");
  if (p->fp32 || p->cfg.scheme != CDD_SCHEME_CENTRAL_UPWIND || p->cfg.dynamics == CDD_GLIDE_CLIMB_WOJZZ ||
      p->vacancy || p->disorder || p->nuc_src)
    return fail(CDD_ERR_UNSUPPORTED,
                "the branch-frozen flux covers float64 central-upwind glide+climb and glide-only plans");
  HIPCHK(hipSetDevice(p->cfg.device));
  HIPCHK(launch_spectral(spectral_job(p, 6, beta, nullptr, nullptr, t, p->G)));
  FrozenJob f;
  memset(&f, 0, sizeof f);
  f.d = p->d;
  f.n = p->n;
  f.dyn = p->cfg.dynamics;
  f.cells = p->cells;
  f.beta = beta;
  f.G = p->G;
  f.sel = sel;
  f.record = record ? 1 : 0;
  f.rhs = rhs;
  f.Lambda = p->cfg.dynamics == CDD_GLIDE_ONLY ? 0.0 : p->cfg.Lambda;
  f.mu = p->gp.mu;
  f.lam = p->gp.lam;
  HIPCHK(launch_flux_frozen(f, p->stream));
  return CDD_OK;
}

int cdd_flux_tangent(cdd_plan* p, const double* beta, const double* dbeta, double t, double dLambda, double* rhs,
                     double* drhs) {
  if (!p || !beta || !dbeta || !rhs || !drhs) return fail(CDD_ERR_ARG, "null argument");
  if (p->fp32 || p->cfg.scheme != CDD_SCHEME_CENTRAL_UPWIND || p->cfg.dynamics == CDD_GLIDE_CLIMB_WOJZZ ||
      p->vacancy || p->disorder || p->nuc_src)
    return fail(CDD_ERR_UNSUPPORTED, "the tangent flux covers float64 central-upwind glide+climb and glide-only plans");
  HIPCHK(hipSetDevice(p->cfg.device));
  HIPCHK(launch_spectral(spectral_job(p, 6, beta, nullptr, nullptr, t, p->G)));
  // the stress of dbeta: the Green's part and the load's tangent, which is
  // nonzero only for the uniaxial control (its -(lam d + 2 mu d_zz d_zz)
  // <dbeta_zz>, the control with rate and initial strain 0)
  double* dG = p->work9;
  SpectralJob jd = spectral_job(p, 6, dbeta, nullptr, nullptr, t, dG);
  const bool uni = p->gp.load.kind == CDD_LOAD_UNIAXIAL;
  memset(&jd.gp.load, 0, sizeof jd.gp.load);
  if (uni) jd.gp.load.kind = CDD_LOAD_UNIAXIAL;
  jd.gp.log_stress = 0;
  HIPCHK(launch_spectral(jd));
  FrozenJob f;
  memset(&f, 0, sizeof f);
  f.d = p->d;
  f.n = p->n;
  f.dyn = p->cfg.dynamics;
  f.cells = p->cells;
  f.beta = beta;
  f.G = p->G;
  f.rhs = rhs;
  f.Lambda = p->cfg.dynamics == CDD_GLIDE_ONLY ? 0.0 : p->cfg.Lambda;
  f.mu = p->gp.mu;
  f.lam = p->gp.lam;
  f.dbeta = dbeta;
  f.dG = dG;
  f.dLambda = p->cfg.dynamics == CDD_GLIDE_ONLY ? 0.0 : dLambda;
  f.drhs = drhs;
  HIPCHK(launch_flux_frozen(f, p->stream));
  return CDD_OK;
}

int cdd_step_vjp(cdd_plan* p, const double* u, double t, double dt, const double* ubar_next, double* ubar,
                 double* lambar) {
  if (!p || !u || !ubar_next || !ubar) return fail(CDD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  RCHK(adj_alloc(p));
  HIPCHK(hipMemsetAsync(p->scalar + 1, 0, sizeof(double), p->stream));
  double *x0 = adjx(p, 0), *x1 = adjx(p, 1), *x2 = adjx(p, 2), *x3 = adjx(p, 3), *x4 = adjx(p, 4),
         *x5 = adjx(p, 5);
  if (p->cfg.rk_order == 2) {
    // forward: F0 = u + dt L(u)
    RCHK(flux_into(p, u, t, x0));
    RCHK(lincomb(p, x1, 2, u, 1.0, x0, dt));
    // reverse of u' = u/2 + (F0 + dt L(F0))/2
    RCHK(lincomb(p, x2, 1, ubar_next, 0.5 * dt));
    RCHK(flux_vjp_impl(p, x1, t, x2, x3));
    RCHK(lincomb(p, x4, 2, ubar_next, 0.5, x3, 1.0));  // F0bar
    RCHK(lincomb(p, x2, 1, x4, dt));
    RCHK(flux_vjp_impl(p, u, t, x2, x3));
    RCHK(lincomb(p, ubar, 3, ubar_next, 0.5, x4, 1.0, x3, 1.0));
  } else {
    // forward: F0 = u + dt L(u), F1 = 3/4 u + 1/4 (F0 + dt L(F0))
    RCHK(flux_into(p, u, t, x0));
    RCHK(lincomb(p, x1, 2, u, 1.0, x0, dt));
    RCHK(flux_into(p, x1, t, x0));
    RCHK(lincomb(p, x2, 3, u, 0.75, x1, 0.25, x0, 0.25 * dt));
    // reverse of u' = u/3 + 2/3 (F1 + dt L(F1))
    RCHK(lincomb(p, x3, 1, ubar_next, 2.0 / 3.0 * dt));
    RCHK(flux_vjp_impl(p, x2, t, x3, x0));
    RCHK(lincomb(p, x4, 2, ubar_next, 2.0 / 3.0, x0, 1.0));  // F1bar
    RCHK(lincomb(p, x3, 1, x4, 0.25 * dt));
    RCHK(flux_vjp_impl(p, x1, t, x3, x0));
    RCHK(lincomb(p, x5, 2, x4, 0.25, x0, 1.0));  // F0bar
    RCHK(lincomb(p, x3, 1, x5, dt));
    RCHK(flux_vjp_impl(p, u, t, x3, x0));
    RCHK(lincomb(p, ubar, 4, ubar_next, 1.0 / 3.0, x4, 0.75, x5, 1.0, x0, 1.0));
  }
  if (lambar) {
    HIPCHK(hipMemcpyAsync(lambar, p->scalar + 1, sizeof(double), hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
  }
  return CDD_OK;
}

int cdd_energy_grad(cdd_plan* p, const double* beta, double t, double* betabar, double* energy) {
  if (!p || !beta || !betabar) return fail(CDD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  RCHK(adj_alloc(p));
  HIPCHK(launch_spectral(spectral_job(p, 6, beta, nullptr, nullptr, t, p->G)));
  int nb = 0;
  HIPCHK(launch_energy_grad(p->G, beta, p->cells, p->gp.mu, p->gp.el.nu, p->gp.lam, p->sbar, p->lpart, &nb,
                            p->stream));
  HIPCHK(launch_reduce_sum(p->lpart, nb, p->scalar + 2, p->stream, 0));
  if (uniaxial(p)) {
    int nu = 0;
    HIPCHK(launch_uniaxial_partial(p->sbar, p->cells, p->gp.lam, p->gp.mu, p->upart, &nu, p->stream));
    HIPCHK(launch_reduce_sum(p->upart, nu, p->scalar + 5, p->stream, 0));
  }
  SpectralJob jt = spectral_job(p, 6, p->sbar, nullptr, nullptr, t, p->G);
  jt.trans = 1;
  HIPCHK(launch_spectral(jt));
  HIPCHK(launch_sigma_finalize(p->G, p->sbar, betabar, p->cells, p->gp.mu, p->gp.lam, p->stream));
  if (uniaxial(p))
    HIPCHK(launch_add_uniform_zz(betabar, p->cells, p->scalar + 5, -1.0 / (double)p->cells, p->stream));
  if (energy) {
    HIPCHK(hipMemcpyAsync(energy, p->scalar + 2, sizeof(double), hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
  }
  return CDD_OK;
}

// One RK stage of the segment reverse pass: out = sum_k coef_k term_k +
// J_L(beta)^T (rscale dt rbar), with beta's stress planes G kept from the
// recomputation and the per-block Lambda cotangents written to lpart.
static int stage_vjp(cdd_plan* p, const double* beta, const double* G, const double* rbar, double rscale,
                     const double* dtp, double* lpart, double* out, const double* csrc, int nterm, const double* t0,
                     double c0, const double* t1 = nullptr, double c1 = 0.0, const double* t2 = nullptr,
                     double c2 = 0.0) {
  VjpJob j;
  memset(&j, 0, sizeof j);
  j.d = p->d;
  j.n = p->n;
  j.dyn = p->cfg.dynamics;
  j.cells = p->cells;
  j.beta = beta;
  j.G = G;
  j.rbar = rbar;
  j.rscale = rscale;
  j.dtp = dtp;
  j.ubar = p->ubar;
  j.codes = p->codes;
  j.sbar = p->sbar;
  j.lpart = lpart;
  j.upart = uniaxial(p) ? p->upart : nullptr;
  j.Lambda = p->cfg.dynamics == CDD_GLIDE_ONLY ? 0.0 : p->cfg.Lambda;
  j.mu = p->gp.mu;
  j.lam = p->gp.lam;
  int nb = 0;
  HIPCHK(launch_flux_vjp_cell(j, &nb, p->stream));
  if (j.upart) HIPCHK(launch_reduce_sum(p->upart, nb, p->scalar + 4, p->stream, 0));
  SpectralJob jt = spectral_job(p, 6, p->sbar, nullptr, nullptr, 0.0, p->G);
  jt.trans = 1;
  jt.src_sym6 = 1;
  HIPCHK(launch_spectral(jt));
  GatherJob g;
  memset(&g, 0, sizeof g);
  g.d = p->d;
  g.n = p->n;
  g.cells = p->cells;
  g.codes = p->codes;
  g.ubar = p->ubar;
  g.sbar = p->sbar;
  g.ebar = p->G;
  g.mu = p->gp.mu;
  g.lam = p->gp.lam;
  g.nterm = nterm;
  g.term[0] = t0;
  g.coef[0] = c0;
  g.term[1] = t1;
  g.coef[1] = c1;
  g.term[2] = t2;
  g.coef[2] = c2;
  g.csrc = csrc;
  if (j.upart) {
    g.uzz = p->scalar + 4;
    g.uzz_scale = -1.0 / (double)p->cells;
  }
  g.out = out;
  HIPCHK(launch_flux_vjp_gather(g, p->stream));
  return CDD_OK;
}

// Layout of a segment's stored trajectory: the S stage states (9 planes)
// of every step, then the S stress solves (6 planes) of every step.  The
// recomputing reverse keeps one more 9-plane state (the segment's end) after
// them.
struct SegLayout {
  double* base;
  int S;
  long long cells;
  int nsteps;
  double* st(int j, int k) const { return base + ((size_t)j * S + k) * 9 * cells; }
  double* gs(int j, int k) const {
    return base + (size_t)nsteps * S * 9 * cells + ((size_t)j * S + k) * 6 * cells;
  }
  static size_t bytes(int nsteps, int S, long long cells) {
    return sizeof(double) * (size_t)nsteps * S * 15 * (size_t)cells;
  }
};

static int lseg_alloc(cdd_plan* p, int nsteps) {
  const size_t lneed = (size_t)p->cfg.rk_order * nsteps * flux_vjp_nblocks(p->d, p->n);
  if (lneed > p->lseg_len) {
    HIPCHK(hipStreamSynchronize(p->stream));
    (void)hipFree(p->lseg);
    p->lseg = nullptr;
    p->lseg_len = 0;
    if (hipMalloc((void**)&p->lseg, sizeof(double) * lneed) != hipSuccess)
      return fail(CDD_ERR_HIP, "hipMalloc failed (segment Lambda partials)");
    p->lseg_len = lneed;
  }
  return CDD_OK;
}

static int seg_alloc(cdd_plan* p, int nsteps) {
  const size_t plane = sizeof(double) * (size_t)p->cells;
  const size_t need = SegLayout::bytes(nsteps, p->cfg.rk_order, p->cells) + 9 * plane;
  if (need > p->seg_bytes) {
    HIPCHK(hipStreamSynchronize(p->stream));
    (void)hipFree(p->seg);
    p->seg = nullptr;
    p->seg_bytes = 0;
    if (hipMalloc((void**)&p->seg, need) != hipSuccess)
      return fail(CDD_ERR_HIP, ("hipMalloc failed (segment workspace of " + std::to_string(need) +
                                " bytes; use shorter checkpoint intervals)").c_str());
    p->seg_bytes = need;
  }
  return lseg_alloc(p, nsteps);
}

// Reverse sweep over a stored segment (last step first): the cotangent of
// the segment's end state ubar_end is carried to ubar0, the per-block Lambda
// cotangents accumulate into the plan's Lambda accumulator.
static int segment_reverse(cdd_plan* p, const SegLayout& L, const double* u0, const double* dts_dev, int nsteps,
                           const double* src_dev, const double* ubar_end, double* ubar0) {
  const int S = p->cfg.rk_order;
  const int nb = flux_vjp_nblocks(p->d, p->n);
  double *A = adjx(p, 0), *B = adjx(p, 1), *F0b = adjx(p, 2), *F1b = adjx(p, 3);
  const double* gin = ubar_end;
  for (int j = nsteps - 1; j >= 0; --j) {
    double* gout = j == 0 ? ubar0 : (gin == A ? B : A);
    const double* dtp = dts_dev + j;
    double* lp = p->lseg + (size_t)j * S * nb;
    const double* u = j == 0 ? u0 : L.st(j, 0);
    // the objective's own gradient at u_j (a uniform 9-vector) joins the last stage
    const double* sj = src_dev ? src_dev + 9 * (size_t)j : nullptr;
    if (S == 2) {
      // u' = u/2 + (F0 + dt L(F0))/2,  F0 = u + dt L(u)
      RCHK(stage_vjp(p, L.st(j, 1), L.gs(j, 1), gin, 0.5, dtp, lp + nb, F0b, nullptr, 1, gin, 0.5));
      RCHK(stage_vjp(p, u, L.gs(j, 0), F0b, 1.0, dtp, lp, gout, sj, 2, gin, 0.5, F0b, 1.0));
    } else {
      // u' = u/3 + 2/3 (F1 + dt L(F1)),  F1 = 3/4 u + 1/4 (F0 + dt L(F0))
      RCHK(stage_vjp(p, L.st(j, 2), L.gs(j, 2), gin, 2.0 / 3.0, dtp, lp + 2 * nb, F1b, nullptr, 1, gin, 2.0 / 3.0));
      RCHK(stage_vjp(p, L.st(j, 1), L.gs(j, 1), F1b, 0.25, dtp, lp + nb, F0b, nullptr, 1, F1b, 0.25));
      RCHK(stage_vjp(p, u, L.gs(j, 0), F0b, 1.0, dtp, lp, gout, sj, 3, gin, 1.0 / 3.0, F1b, 0.75, F0b, 1.0));
    }
    gin = gout;
  }
  HIPCHK(launch_reduce_sum(p->lseg, S * nsteps * nb, p->scalar + 3, p->stream, 1));
  return CDD_OK;
}

int cdd_segment_vjp(cdd_plan* p, const double* u0, double t0, const double* dts_dev, int nsteps,
                    const double* src_dev, const double* ubar_end, double* ubar0) {
  if (!p || !u0 || !dts_dev || !ubar_end || !ubar0 || nsteps < 1) return fail(CDD_ERR_ARG, "bad argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  RCHK(adj_alloc(p));
  RCHK(seg_alloc(p, nsteps));
  const int S = p->cfg.rk_order;
  const SegLayout L{p->seg, S, p->cells, nsteps};
  double* ufinal = L.gs(nsteps, 0);
  // recompute the segment from u0, replaying the recorded time steps and
  // keeping every stage state and its stress planes
  HIPCHK(launch_clock_segment(p->clk, 0, t0, dts_dev, nsteps, p->stream));
  for (int j = 0; j < nsteps; ++j) {
    StepBufs b;
    b.u = j == 0 ? u0 : L.st(j, 0);
    b.out = j + 1 < nsteps ? L.st(j + 1, 0) : ufinal;
    for (int k = 0; k < 3; ++k) b.G[k] = L.gs(j, k < S ? k : S - 1);
    b.F0 = L.st(j, 1);
    b.F1 = S == 3 ? L.st(j, 2) : nullptr;
    RCHK(launch_step(p, b));
  }
  HIPCHK(launch_clock_segment(p->clk, 1, 0.0, nullptr, 0, p->stream));
  return segment_reverse(p, L, u0, dts_dev, nsteps, src_dev, ubar_end, ubar0);
}

int cdd_segment_store_bytes(cdd_plan* p, int nsteps, unsigned long long* bytes) {
  if (!p || !bytes || nsteps < 1) return fail(CDD_ERR_ARG, "bad argument");
  *bytes = (unsigned long long)SegLayout::bytes(nsteps, p->cfg.rk_order, p->cells);
  return CDD_OK;
}

int cdd_segment_forward(cdd_plan* p, int nsteps, void* store, unsigned long long store_bytes) {
  if (!p || !store || nsteps < 1) return fail(CDD_ERR_ARG, "bad argument");
  if (p->fp32) return fail(CDD_ERR_UNSUPPORTED, "stored trajectories are float64 only");
  if (p->vacancy) return fail(CDD_ERR_UNSUPPORTED, "the adjoint does not include the vacancy coupling");
  if (p->disorder) return fail(CDD_ERR_UNSUPPORTED, "the adjoint does not include the disorder coupling");
  if (p->cfg.scheme != CDD_SCHEME_CENTRAL_UPWIND)
    return fail(CDD_ERR_UNSUPPORTED, "the adjoint differentiates the central-upwind scheme");
  if (store_bytes < SegLayout::bytes(nsteps, p->cfg.rk_order, p->cells))
    return fail(CDD_ERR_ARG, "segment store too small (cdd_segment_store_bytes)");
  HIPCHK(hipSetDevice(p->cfg.device));
  const int S = p->cfg.rk_order;
  const SegLayout L{static_cast<double*>(store), S, p->cells, nsteps};
  // the segment's start state is kept in the store; the last step writes the
  // buffer cdd_step would have ended on (RK2 ping-pongs by step parity)
  const int ci = cur_state(p);
  const int cf = S == 3 ? ci : ci ^ (nsteps & 1);
  HIPCHK(hipMemcpyAsync(L.st(0, 0), p->state[ci], sizeof(double) * 9 * (size_t)p->cells, hipMemcpyDeviceToDevice,
                        p->stream));
  for (int j = 0; j < nsteps; ++j) {
    StepBufs b;
    b.u = L.st(j, 0);
    b.out = j + 1 < nsteps ? L.st(j + 1, 0) : p->state[cf];
    for (int k = 0; k < 3; ++k) b.G[k] = L.gs(j, k < S ? k : S - 1);
    b.F0 = L.st(j, 1);
    b.F1 = S == 3 ? L.st(j, 2) : nullptr;
    RCHK(launch_step(p, b));
  }
  p->steps_launched += nsteps;
  return CDD_OK;
}

int cdd_segment_vjp_stored(cdd_plan* p, const void* store, unsigned long long store_bytes, const double* dts_dev,
                           int nsteps, const double* src_dev, const double* ubar_end, double* ubar0) {
  if (!p || !store || !dts_dev || !ubar_end || !ubar0 || nsteps < 1) return fail(CDD_ERR_ARG, "bad argument");
  if (store_bytes < SegLayout::bytes(nsteps, p->cfg.rk_order, p->cells))
    return fail(CDD_ERR_ARG, "segment store too small (cdd_segment_store_bytes)");
  HIPCHK(hipSetDevice(p->cfg.device));
  RCHK(adj_alloc(p));
  RCHK(lseg_alloc(p, nsteps));
  const SegLayout L{const_cast<double*>(static_cast<const double*>(store)), p->cfg.rk_order, p->cells, nsteps};
  return segment_reverse(p, L, L.st(0, 0), dts_dev, nsteps, src_dev, ubar_end, ubar0);
}

int cdd_lambda_bar(cdd_plan* p, double* value, int reset) {
  if (!p) return fail(CDD_ERR_ARG, "null plan");
  HIPCHK(hipSetDevice(p->cfg.device));
  if (value) {
    HIPCHK(hipMemcpyAsync(value, p->scalar + 3, sizeof(double), hipMemcpyDeviceToHost, p->stream));
  }
  if (reset) HIPCHK(hipMemsetAsync(p->scalar + 3, 0, sizeof(double), p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return CDD_OK;
}

// ------------------------------------------------------------- .plas files
// .plas records of a plan with the vacancy coupling on carry c (float64 plans)
static bool vac_record(const cdd_plan* p) { return p->vacancy && p->vac && !p->fp32; }

static int host_alloc(cdd_plan* p) {
  if (p->host) return CDD_OK;
  // one record: t and ten planes (the vacancy component of VACANCIES records)
  const size_t bytes = p->es * (1 + 10 * (size_t)p->cells);
  if (hipHostMalloc(&p->host, bytes) != hipSuccess) return fail(CDD_ERR_HIP, "hipHostMalloc failed (record staging)");
  return CDD_OK;
}

int cdd_plas_append(cdd_plan* p, const char* path) {
  if (!p || !path) return fail(CDD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  RCHK(host_alloc(p));
  Clock c;
  if (sync_clock(p, &c)) return CDD_ERR_HIP;
  const size_t fb = p->es * 9 * (size_t)p->cells;
  HIPCHK(launch_reverse_axes(p->state[cur_state(p)], p->work9, p->d, p->n, 9, p->fp32, p->stream));
  char* h = static_cast<char*>(p->host);
  HIPCHK(hipMemcpyAsync(h + p->es, p->work9, fb, hipMemcpyDeviceToHost, p->stream));
  // with the vacancy coupling on, the record carries c as its tenth
  // component (the executable's VACANCIES records, VACANCY_COMP = 9)
  const size_t cb = vac_record(p) ? p->es * (size_t)p->cells : 0;
  if (cb) {
    HIPCHK(launch_reverse_axes(p->vac, p->L0, p->d, p->n, 1, 0, p->stream));
    HIPCHK(hipMemcpyAsync(h + p->es + fb, p->L0, cb, hipMemcpyDeviceToHost, p->stream));
  }
  HIPCHK(hipStreamSynchronize(p->stream));
  if (p->fp32) {
    const float tf = (float)c.t;
    memcpy(h, &tf, sizeof tf);
  } else {
    memcpy(h, &c.t, sizeof c.t);
  }
  FILE* f = fopen(path, "ab");
  if (!f) return fail(CDD_ERR_ARG, (std::string("cannot open ") + path).c_str());
  const size_t want = p->es + fb + cb;
  const size_t got = fwrite(h, 1, want, f);
  const int fl = fflush(f);
  const int cl = fclose(f);
  if (got != want || fl != 0 || cl != 0) return fail(CDD_ERR_ARG, "short write to the .plas file");
  return CDD_OK;
}

int cdd_plas_resume(cdd_plan* p, const char* path, double* t, long long* records, int* torn) {
  if (!p || !path) return fail(CDD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  RCHK(host_alloc(p));
  FILE* f = fopen(path, "rb");
  if (!f) return fail(CDD_ERR_ARG, (std::string("cannot open ") + path).c_str());
  const size_t fb = p->es * 9 * (size_t)p->cells;
  const size_t cb = vac_record(p) ? p->es * (size_t)p->cells : 0;
  const long long rec = (long long)(p->es + fb + cb);
  if (fseeko(f, 0, SEEK_END) != 0) {
    fclose(f);
    return fail(CDD_ERR_ARG, "cannot seek in the .plas file");
  }
  const long long size = (long long)ftello(f);
  const long long nrec = size / rec;
  if (records) *records = nrec;
  if (torn) *torn = (size % rec) != 0;
  if (nrec < 1) {
    fclose(f);
    return fail(CDD_ERR_ARG, "no complete record in the .plas file (wrong grid size or precision?)");
  }
  // the last complete record; a torn tail (an interrupted append) is skipped
  if (fseeko(f, (off_t)((nrec - 1) * rec), SEEK_SET) != 0) {
    fclose(f);
    return fail(CDD_ERR_ARG, "cannot seek in the .plas file");
  }
  char* h = static_cast<char*>(p->host);
  const size_t got = fread(h, 1, (size_t)rec, f);
  fclose(f);
  if ((long long)got != rec) return fail(CDD_ERR_ARG, "short read from the .plas file");
  double tr;
  if (p->fp32) {
    float tf;
    memcpy(&tf, h, sizeof tf);
    tr = tf;
  } else {
    memcpy(&tr, h, sizeof tr);
  }
  Clock c;
  if (sync_clock(p, &c)) return CDD_ERR_HIP;
  HIPCHK(hipMemcpyAsync(p->work9, h + p->es, fb, hipMemcpyHostToDevice, p->stream));
  HIPCHK(launch_reverse_axes(p->work9, p->state[cur_state(p)], p->d, p->n, 9, p->fp32, p->stream));
  if (cb) {
    HIPCHK(hipMemcpyAsync(p->L0, h + p->es + fb, cb, hipMemcpyHostToDevice, p->stream));
    HIPCHK(launch_reverse_axes(p->L0, p->vac, p->d, p->n, 1, 0, p->stream));
  }
  HIPCHK(hipStreamSynchronize(p->stream));
  c.t = tr;
  c.done = (tr >= c.t_end) ? 1 : 0;
  HIPCHK(hipMemcpy(p->clk, &c, sizeof c, hipMemcpyHostToDevice));
  if (t) *t = tr;
  return CDD_OK;
}

int cdd_dt_record(cdd_plan* p, double* log_dev, long long cap) {
  if (!p || cap < 0) return fail(CDD_ERR_ARG, "bad argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  Clock c;
  if (sync_clock(p, &c)) return CDD_ERR_HIP;
  c.dt_log = cap > 0 ? log_dev : nullptr;
  c.log_base = c.nsteps;
  c.log_cap = cap;
  HIPCHK(hipMemcpy(p->clk, &c, sizeof c, hipMemcpyHostToDevice));
  return CDD_OK;
}

int cdd_speed_record(cdd_plan* p, double* log_dev, long long cap) {
  if (!p || cap < 0) return fail(CDD_ERR_ARG, "bad argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  Clock c;
  if (sync_clock(p, &c)) return CDD_ERR_HIP;
  c.vlog = cap > 0 ? log_dev : nullptr;
  c.vlog_base = c.nsteps;
  c.vlog_cap = cap;
  HIPCHK(hipMemcpy(p->clk, &c, sizeof c, hipMemcpyHostToDevice));
  return CDD_OK;
}

int cdd_stress_record(cdd_plan* p, double* log_dev, long long cap) {
  if (!p || cap < 0) return fail(CDD_ERR_ARG, "bad argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  Clock c;
  if (sync_clock(p, &c)) return CDD_ERR_HIP;
  c.slog = cap > 0 ? log_dev : nullptr;
  c.slog_base = c.nsteps;
  c.slog_cap = cap;
  HIPCHK(hipMemcpy(p->clk, &c, sizeof c, hipMemcpyHostToDevice));
  return CDD_OK;
}

int cdd_energy_record(cdd_plan* p, double* log_dev, long long cap) {
  if (!p || cap < 0) return fail(CDD_ERR_ARG, "bad argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  if (cap > 0 && !p->epart &&
      hipMalloc((void**)&p->epart, sizeof(double) * (size_t)((p->cells + 255) / 256)) != hipSuccess)
    return fail(CDD_ERR_HIP, "hipMalloc failed (energy partials)");
  Clock c;
  if (sync_clock(p, &c)) return CDD_ERR_HIP;
  c.elog = cap > 0 ? log_dev : nullptr;
  c.elog_base = c.nsteps;
  c.elog_cap = cap;
  HIPCHK(hipMemcpy(p->clk, &c, sizeof c, hipMemcpyHostToDevice));
  p->elog_on = cap > 0;
  return CDD_OK;
}

int cdd_status(cdd_plan* p, int* status) {
  if (!p || !status) return fail(CDD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  Clock c;
  if (sync_clock(p, &c)) return CDD_ERR_HIP;
  *status = c.status;
  return CDD_OK;
}

int cdd_instability(cdd_plan* p, const void* beta_v, double* value) {
  if (!p || !value) return fail(CDD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  const double* beta = static_cast<const double*>(beta_v);
  if (!beta) {
    if (sync_clock(p, nullptr)) return CDD_ERR_HIP;
    beta = p->state[cur_state(p)];
  }
  HIPCHK(launch_spectral(spectral_job(p, 9, beta, nullptr, nullptr, 0.0, p->work9)));
  int nb = 0;
  HIPCHK(launch_nyquist_mode(p->work9, p->d, p->n, p->cells, p->fp32, p->partial, &nb, p->stream));
  HIPCHK(launch_reduce_sum(p->partial, nb, p->scalar + 6, p->stream, 0));
  double v = 0.0;
  HIPCHK(hipMemcpyAsync(&v, p->scalar + 6, sizeof(double), hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  *value = fabs(v);
  return CDD_OK;
}

int cdd_dt_replay(cdd_plan* p, const double* sched_dev, long long len) {
  if (!p || len < 0) return fail(CDD_ERR_ARG, "bad argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  Clock c;
  if (sync_clock(p, &c)) return CDD_ERR_HIP;
  c.dt_sched = len > 0 ? sched_dev : nullptr;
  c.sched_base = c.nsteps;
  c.sched_len = len;
  HIPCHK(hipMemcpy(p->clk, &c, sizeof c, hipMemcpyHostToDevice));
  return CDD_OK;
}

int cdd_timing_enable(cdd_plan* p, int max_events) {
  if (!p || max_events < 0) return fail(CDD_ERR_ARG, "bad argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  HIPCHK(hipStreamSynchronize(p->stream));
  if (!p->ev) {
    p->ev = new std::vector<hipEvent_t>();
    p->evkind = new std::vector<int>();
  }
  for (hipEvent_t e : *p->ev) (void)hipEventDestroy(e);
  p->ev->clear();
  p->evkind->assign(max_events, -1);
  for (int i = 0; i < max_events; ++i) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    p->ev->push_back(e);
  }
  p->nev = 0;
  p->timing = max_events > 0;
  return CDD_OK;
}

int cdd_timing_read(cdd_plan* p, double* ms, int* counts, int nkinds) {
  if (!p || !ms || !counts) return fail(CDD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(p->cfg.device));
  HIPCHK(hipStreamSynchronize(p->stream));
  for (int k = 0; k < nkinds; ++k) {
    ms[k] = 0.0;
    counts[k] = 0;
  }
  for (int i = 1; i < p->nev; ++i) {
    const int k = (*p->evkind)[i];
    if (k < 0 || k >= nkinds) continue;
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, (*p->ev)[i - 1], (*p->ev)[i]));
    ms[k] += t;
    counts[k] += 1;
  }
  return CDD_OK;
}

}  // extern "C"
