// Batched stepping of many independent 2D realisations of uneven sizes
// (include/cdd.h, cdd_batch_*).  The reference runs one realisation per
// process per GPU (Plasticity/exec_batch.py:26-58); here all of them advance
// together.  Instances are grouped by grid side n; every group's buffers
// have uniform per-instance strides and each instance keeps its own device
// clock (own CFL dt and end time).  One step is captured once into a
// hipGraph and replayed for every step:
//   per stage: the spectral stress passes of every size group, each group on
//   its own forked capture stream (independent branches that run
//   concurrently; launch = all instances of the size, instance =
//   blockIdx.y), joined into ONE persistent multi-instance flux launch that
//   covers every instance of every size (k_flux_multi: a fixed grid walking
//   the flat (size, instance, tile) list); after the first stage the CFL
//   reductions of every instance, after the last the clock advance.
// Groups whose side the flux tile does not divide, and the per-size mode
// (cdd_batch_set_flux with per_cu = 0), launch the per-size k_flux instead.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "../../include/cdd.h"
#include "cdd_kernels.h"

using namespace cdd;

extern int cdd_internal_fail(int code, const char* msg);

#define BCHK(expr)                                                                \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      char b_[256];                                                               \
      snprintf(b_, sizeof b_, "%s failed: %s", #expr, hipGetErrorString(e_));     \
      return cdd_internal_fail(CDD_ERR_HIP, b_);                                  \
    }                                                                             \
  } while (0)

namespace {

struct Group {
  int n = 0, count = 0, tile = 0, stile = 0, nblocks = 0;
  long long cells = 0, fstride = 0, gstride = 0, sstride = 0;
  double* state[2] = {nullptr, nullptr};
  double* L0 = nullptr;
  double* buf = nullptr;
  double* G = nullptr;
  double* partial = nullptr;
  void* spec = nullptr;
  void* tw = nullptr;
  Clock* clk = nullptr;
  hipStream_t st = nullptr;
};

}  // namespace

struct cdd_batch {
  cdd_config cfg;
  GreenParams gp0;  // mu/nu/lam/A/C (norm per group)
  int flux_tile = 3;    // multi-instance flux variant (0-2: tile walk, 3: row march)
  int flux_per_cu = 2;  // its persistent blocks per CU; 0: per-size k_flux launches
  std::vector<Group> groups;
  std::vector<int> inst_group, inst_slot;
  std::vector<long long> set_global, set_nsteps;  // ping-pong bookkeeping (RK2)
  long long global_steps = 0;
  hipStream_t own = nullptr, stream = nullptr;
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  hipGraph_t graph[2] = {nullptr, nullptr};
  std::vector<hipEvent_t> ev;
};

static void free_graphs(cdd_batch* b) {
  for (int k = 0; k < 2; ++k) {
    if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
    if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
    b->exec[k] = nullptr;
    b->graph[k] = nullptr;
  }
}

// The buffers of stage s of one RK step of group g reading state[phase]
// (RK2 writes state[phase^1]): the spectral pass transforms SA (+ dt SB,
// written to SW when SB is set); the flux reads FA (and FU) and writes Fout.
struct StageIO {
  const double *SA, *SB;
  double* SW;
  const double *FA, *FU;
  double* Fout;
  int mode;
  double cu, cf;
  bool partial;
};

static int num_stages(const cdd_batch* b) { return b->cfg.rk_order == 3 ? 3 : 2; }

static StageIO stage_io(const cdd_batch* b, const Group& g, int phase, int s) {
  const int rk = b->cfg.rk_order;
  double* u = g.state[rk == 3 ? 0 : phase];
  StageIO io;
  memset(&io, 0, sizeof io);
  if (s == 0) {  // L0 = L(u), CFL partials
    io.SA = u;
    io.FA = u;
    io.Fout = g.L0;
    io.partial = true;
  } else if (s == 1) {  // F0 = u + dt L0 formed over L0
    io.SA = u;
    io.SB = g.L0;
    io.SW = g.L0;
    io.FA = g.L0;
    io.FU = u;
    io.mode = 1;
    if (rk == 3) {
      io.Fout = g.buf;
      io.cu = 0.75;
      io.cf = 0.25;
    } else {
      io.Fout = g.state[phase ^ 1];
      io.cu = 0.5;
      io.cf = 0.5;
    }
  } else {  // RK3: u' = u/3 + 2/3 (F1 + dt L(F1))
    io.SA = g.buf;
    io.FA = g.buf;
    io.FU = u;
    io.Fout = u;
    io.mode = 1;
    io.cu = 1.0 / 3.0;
    io.cf = 2.0 / 3.0;
  }
  return io;
}

static hipError_t group_spectral(cdd_batch* b, Group& g, const StageIO& io, hipStream_t st) {
  SpectralJob j;
  memset(&j, 0, sizeof j);
  j.d = 2;
  j.n = g.n;
  j.nc = 6;
  j.cells = g.cells;
  j.src = SrcField<double>{io.SA, io.SB, g.clk, g.cells, io.SW, g.fstride};
  j.spec = g.spec;
  j.tw = g.tw;
  j.out = g.G;
  j.gp = b->gp0;
  j.gp.norm = 1.0 / (double)g.cells;
  j.gp.clk = g.clk;
  j.stream = st;
  j.count = g.count;
  j.sstride = g.sstride;
  j.ostride = g.gstride;
  j.stile = g.stile;
  // the size groups' passes run concurrently: the policy follows all of them
  for (const Group& o : b->groups) j.ws_bytes += 6ll * (o.n / 2 + 1) * o.n * 16 * o.count;
  return launch_spectral(j);
}

// the per-size flux launch (tile variant g.tile)
static hipError_t group_flux(cdd_batch* b, Group& g, const StageIO& io, hipStream_t st) {
  FluxJob f;
  memset(&f, 0, sizeof f);
  f.d = 2;
  f.n = g.n;
  f.dyn = b->cfg.dynamics;
  f.cells = g.cells;
  f.A = io.FA;
  f.G = g.G;
  f.U = io.FU;
  f.out = io.Fout;
  f.partial = io.partial ? g.partial : nullptr;
  f.clk = g.clk;
  f.mode = io.mode;
  f.cu = io.cu;
  f.cf = io.cf;
  f.Lambda = b->cfg.dynamics == CDD_GLIDE_ONLY ? 0.0 : b->cfg.Lambda;
  f.mu = b->gp0.mu;
  f.lam = b->gp0.lam;
  f.stream = st;
  f.count = g.count;
  f.fstride = g.fstride;
  f.gstride = g.gstride;
  f.pstride = g.nblocks;
  f.tile = g.tile;
  return launch_flux(f, nullptr);
}

static bool in_multi(const cdd_batch* b, const Group& g) {
  if (b->flux_per_cu <= 0) return false;
  return b->flux_tile == 3 ? flux_rows_ok(g.n) : flux_multi_ok(g.n, b->flux_tile);
}

// The CFL reductions of the groups in the multi-instance launch, in one
// launch: the row-marching kernel leaves one atomically maximised slot per
// instance (read and reset), the tile walk one partial per tile.
static hipError_t multi_finalize(cdd_batch* b, hipStream_t st) {
  std::vector<double*> partial;
  std::vector<int> nb, count, clear;
  std::vector<long long> pstride;
  std::vector<Clock*> clk;
  std::vector<double> nmax;
  for (Group& g : b->groups) {
    if (!in_multi(b, g)) continue;
    partial.push_back(g.partial);
    nb.push_back(b->flux_tile == 3 ? 1 : g.nblocks);
    clear.push_back(b->flux_tile == 3 ? 1 : 0);
    pstride.push_back(g.nblocks);
    clk.push_back(g.clk);
    nmax.push_back((double)g.n);
    count.push_back(g.count);
  }
  if (partial.empty()) return hipSuccess;
  return launch_finalize_dt_groups((int)partial.size(), partial.data(), nb.data(), pstride.data(), clk.data(),
                                   nmax.data(), count.data(), clear.data(), st);
}

// One multi-instance flux launch of stage s over every group in_multi
// (only == -1), or over the groups other than group `only` (skip_only), or
// over group `only` alone.
static hipError_t multi_flux(cdd_batch* b, int phase, int s, hipStream_t st, int only = -1, bool skip_only = false) {
  FluxMultiJob m;
  memset(&m, 0, sizeof m);
  m.dyn = b->cfg.dynamics;
  m.tile = b->flux_tile;
  m.per_cu = b->flux_per_cu;
  m.Lambda = b->cfg.dynamics == CDD_GLIDE_ONLY ? 0.0 : b->cfg.Lambda;
  m.mu = b->gp0.mu;
  m.lam = b->gp0.lam;
  int k = 0;
  // largest sides first
  std::vector<int> order;
  for (size_t gi = 0; gi < b->groups.size(); ++gi) {
    if (!in_multi(b, b->groups[gi])) continue;
    if (only >= 0 && (((int)gi == only) == skip_only)) continue;
    order.push_back((int)gi);
  }
  if (order.empty()) return hipSuccess;
  std::sort(order.begin(), order.end(), [&](int x, int y) { return b->groups[x].n > b->groups[y].n; });
  for (int gi : order) {
    Group& g = b->groups[gi];
    const StageIO io = stage_io(b, g, phase, s);
    m.mode = io.mode;
    m.cu = io.cu;
    m.cf = io.cf;
    FluxGroup& G = m.g[k++];
    G.n = g.n;
    G.count = g.count;
    G.cells = g.cells;
    G.fstride = g.fstride;
    G.gstride = g.gstride;
    G.pstride = g.nblocks;
    G.A = io.FA;
    G.G = g.G;
    G.U = io.FU;
    G.out = io.Fout;
    G.partial = io.partial ? g.partial : nullptr;
    G.clk = g.clk;
  }
  return b->flux_tile == 3 ? launch_flux_rows(m, k, st) : launch_flux_multi(m, k, st);
}

// Capture one step of every group into graph[phase]: per stage, fork one
// capture stream per group from the batch stream for the spectral passes
// (and the per-size flux of groups outside the multi-instance launch), join
// them back, then the multi-instance flux on the batch stream.  The group of
// the largest side, whose stress solve ends last (its 2D contraction holds
// one block per CU), gets its own multi-instance flux launch on its stream
// right after that solve, and the other groups' launch runs as soon as
// theirs are done, beside the large group's last spectral passes (1.5-2%
// per batched step at sides 256-1024; one launch per group measured the same).
static int capture(cdd_batch* b, int phase) {
  // capture on the private stream (the caller's stream may be the null
  // stream, which cannot be captured); the graph is launched on the caller's
  hipStream_t main = b->own;
  BCHK(hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal));
  const int ng = (int)b->groups.size();
  int late = -1, nmulti = 0;
  for (int gi = 0; gi < ng; ++gi) {
    const Group& g = b->groups[gi];
    if (!in_multi(b, g)) continue;
    ++nmulti;
    if (late < 0 || g.n > b->groups[late].n ||
        (g.n == b->groups[late].n && g.count > b->groups[late].count))
      late = gi;
  }
  if (nmulti < 2) late = -1;
  hipError_t err = hipSuccess;
  for (int s = 0; s < num_stages(b) && err == hipSuccess; ++s) {
    hipEvent_t fork = b->ev[s * (1 + ng)];
    err = hipEventRecord(fork, main);
    for (int gi = 0; gi < ng && err == hipSuccess; ++gi) {
      Group& g = b->groups[gi];
      const StageIO io = stage_io(b, g, phase, s);
      err = hipStreamWaitEvent(g.st, fork, 0);
      if (err == hipSuccess) err = group_spectral(b, g, io, g.st);
      if (err == hipSuccess && !in_multi(b, g)) {
        err = group_flux(b, g, io, g.st);
        if (err == hipSuccess && s == 0)
          err = launch_finalize_dt(g.partial, g.nblocks, g.clk, (double)g.n, g.st, g.count, g.nblocks);
      }
      if (err == hipSuccess && gi == late) continue;  // joined after its own flux below
      hipEvent_t join = b->ev[s * (1 + ng) + 1 + gi];
      if (err == hipSuccess) err = hipEventRecord(join, g.st);
      if (err == hipSuccess) err = hipStreamWaitEvent(main, join, 0);
    }
    if (err == hipSuccess) err = multi_flux(b, phase, s, main, late, true);
    if (err == hipSuccess && late >= 0) {
      Group& g = b->groups[late];
      err = multi_flux(b, phase, s, g.st, late, false);
      hipEvent_t join = b->ev[s * (1 + ng) + 1 + late];
      if (err == hipSuccess) err = hipEventRecord(join, g.st);
      if (err == hipSuccess) err = hipStreamWaitEvent(main, join, 0);
    }
    if (s == 0 && err == hipSuccess) err = multi_finalize(b, main);
  }
  if (err == hipSuccess) {
    std::vector<Clock*> clk;
    std::vector<int> count;
    for (Group& g : b->groups) {
      clk.push_back(g.clk);
      count.push_back(g.count);
    }
    err = launch_advance_groups(ng, clk.data(), count.data(), main);
  }
  hipGraph_t graph = nullptr;
  hipError_t e2 = hipStreamEndCapture(main, &graph);
  if (err != hipSuccess || e2 != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    char m[160];
    snprintf(m, sizeof m, "graph capture failed: %s", hipGetErrorString(err != hipSuccess ? err : e2));
    return cdd_internal_fail(CDD_ERR_HIP, m);
  }
  b->graph[phase] = graph;
  BCHK(hipGraphInstantiate(&b->exec[phase], graph, nullptr, nullptr, 0));
  return CDD_OK;
}

// Zero every CFL partial (outside any capture): the row-marching flux
// maximises into slot 0 of each instance atomically, so a slot left by a
// timing launch or by another flux variant must not reach a later step.
static int zero_partials(cdd_batch* b) {
  BCHK(hipStreamSynchronize(b->stream));
  for (Group& g : b->groups) BCHK(hipMemset(g.partial, 0, sizeof(double) * (size_t)g.nblocks * g.count));
  BCHK(hipDeviceSynchronize());
  return CDD_OK;
}

static int sync_clocks(cdd_batch* b, std::vector<Clock>& out) {
  BCHK(hipStreamSynchronize(b->stream));
  out.resize(b->inst_group.size());
  for (size_t i = 0; i < b->inst_group.size(); ++i) {
    Group& g = b->groups[b->inst_group[i]];
    BCHK(hipMemcpy(&out[i], g.clk + b->inst_slot[i], sizeof(Clock), hipMemcpyDeviceToHost));
  }
  return CDD_OK;
}

// buffer holding instance i's state
static double* inst_state(cdd_batch* b, int i, const Clock& c) {
  Group& g = b->groups[b->inst_group[i]];
  int buf = 0;
  if (b->cfg.rk_order != 3) buf = (int)((b->set_global[i] + (c.nsteps - b->set_nsteps[i])) & 1);
  return g.state[buf] + (long long)b->inst_slot[i] * g.fstride;
}

extern "C" {

int cdd_batch_create(const cdd_config* cfg, const int* sizes, int count, cdd_batch** out) {
  if (!cfg || !sizes || !out || count < 1 || count > 4096) return cdd_internal_fail(CDD_ERR_ARG, "bad argument");
  if (cfg->ndim != 2) return cdd_internal_fail(CDD_ERR_UNSUPPORTED, "batched stepping is 2D");
  if (cfg->rk_order != 2 && cfg->rk_order != 3) return cdd_internal_fail(CDD_ERR_ARG, "rk_order must be 2 or 3");
  if (cfg->scheme != CDD_SCHEME_CENTRAL_UPWIND)
    return cdd_internal_fail(CDD_ERR_UNSUPPORTED, "batched stepping uses the central-upwind scheme");
  if (cfg->load_kind != CDD_LOAD_NONE) return cdd_internal_fail(CDD_ERR_UNSUPPORTED, "batched stepping has no load");
  if (cfg->precision != CDD_FP64) return cdd_internal_fail(CDD_ERR_UNSUPPORTED, "batched stepping is float64 only");
  if (cfg->dynamics != CDD_GLIDE_CLIMB && cfg->dynamics != CDD_GLIDE_ONLY)
    return cdd_internal_fail(CDD_ERR_UNSUPPORTED, "batched stepping supports glide+climb and glide-only");
  const int ok2[] = {16, 32, 64, 128, 256, 384, 512, 768, 1024};
  std::map<int, int> gidx;
  for (int i = 0; i < count; ++i) {
    bool ok = false;
    for (int v : ok2) ok |= (v == sizes[i]);
    if (!ok) return cdd_internal_fail(CDD_ERR_UNSUPPORTED, "unsupported instance size");
  }
  BCHK(hipSetDevice(cfg->device));
  cdd_batch* b = new cdd_batch();
  b->cfg = *cfg;
  const double mu = cfg->mu, nu = cfg->nu, lam = 2.0 * mu * nu / (1.0 - 2.0 * nu);
  memset(&b->gp0, 0, sizeof b->gp0);
  b->gp0.A = lam / (lam + 2.0 * mu);
  b->gp0.C = 2.0 * (lam + mu) / (lam + 2.0 * mu);
  b->gp0.mu = mu;
  b->gp0.lam = lam;
  b->gp0.el = Elastic{mu, nu, lam};
  for (int i = 0; i < count; ++i) {
    auto it = gidx.find(sizes[i]);
    if (it == gidx.end()) {
      gidx[sizes[i]] = (int)b->groups.size();
      Group g;
      g.n = sizes[i];
      b->groups.push_back(g);
      it = gidx.find(sizes[i]);
    }
    Group& g = b->groups[it->second];
    b->inst_group.push_back(it->second);
    b->inst_slot.push_back(g.count++);
  }
  b->set_global.assign(count, 0);
  b->set_nsteps.assign(count, 0);
  auto bail = [&](const char* m) {
    cdd_batch_destroy(b);
    return cdd_internal_fail(CDD_ERR_HIP, m);
  };
  for (Group& g : b->groups) {
    const int n = g.n;
    g.cells = (long long)n * n;
    g.fstride = 9 * g.cells;
    g.gstride = 6 * g.cells;
    g.sstride = 6LL * (n / 2 + 1) * n;
    g.nblocks = flux_nblocks_tile(2, n, 0);
    for (int t = 1; t <= 2; ++t) g.nblocks = flux_nblocks_tile(2, n, t) > g.nblocks ? flux_nblocks_tile(2, n, t) : g.nblocks;
    const size_t fb = sizeof(double) * (size_t)g.fstride * g.count;
    if (hipMalloc((void**)&g.state[0], fb) != hipSuccess) return bail("hipMalloc failed");
    if (cfg->rk_order == 2 && hipMalloc((void**)&g.state[1], fb) != hipSuccess) return bail("hipMalloc failed");
    if (hipMalloc((void**)&g.L0, fb) != hipSuccess) return bail("hipMalloc failed");
    if (cfg->rk_order == 3 && hipMalloc((void**)&g.buf, fb) != hipSuccess) return bail("hipMalloc failed");
    if (hipMalloc((void**)&g.G, sizeof(double) * (size_t)g.gstride * g.count) != hipSuccess) return bail("hipMalloc failed");
    if (hipMalloc(&g.spec, sizeof(double) * 2 * (size_t)g.sstride * g.count) != hipSuccess) return bail("hipMalloc failed");
    if (hipMalloc((void**)&g.partial, sizeof(double) * (size_t)g.nblocks * g.count) != hipSuccess)
      return bail("hipMalloc failed");
    // the row-marching flux maximises into slot 0 of each instance atomically
    if (hipMemset(g.partial, 0, sizeof(double) * (size_t)g.nblocks * g.count) != hipSuccess)
      return bail("memset failed");
    if (hipMalloc((void**)&g.clk, sizeof(Clock) * g.count) != hipSuccess) return bail("hipMalloc failed");
    if (hipMalloc(&g.tw, sizeof(double) * 2 * n) != hipSuccess) return bail("hipMalloc failed");
    std::vector<double> h(2 * n);
    for (int e = 0; e < n; ++e) {
      long double th = 2.0L * 3.14159265358979323846264338327950288L * (long double)e / (long double)n;
      h[2 * e] = (double)cosl(th);
      h[2 * e + 1] = (double)(-sinl(th));
    }
    if (hipMemcpy(g.tw, h.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice) != hipSuccess)
      return bail("twiddle upload failed");
    std::vector<Clock> c0(g.count);
    memset(c0.data(), 0, sizeof(Clock) * g.count);
    for (auto& c : c0) {
      c.t_end = INFINITY;
      c.cfl = cfg->cfl;
    }
    if (hipMemcpy(g.clk, c0.data(), sizeof(Clock) * g.count, hipMemcpyHostToDevice) != hipSuccess)
      return bail("clock upload failed");
    if (hipMemset(g.state[0], 0, fb) != hipSuccess) return bail("memset failed");
    if (g.state[1] && hipMemset(g.state[1], 0, fb) != hipSuccess) return bail("memset failed");
    if (hipStreamCreateWithFlags(&g.st, hipStreamNonBlocking) != hipSuccess) return bail("stream create failed");
  }
  if (hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking) != hipSuccess) return bail("stream create failed");
  b->stream = b->own;
  b->ev.resize(3 * (1 + b->groups.size()));
  for (auto& e : b->ev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return bail("event create failed");
  *out = b;
  return CDD_OK;
}

int cdd_batch_destroy(cdd_batch* b) {
  if (!b) return CDD_OK;
  (void)hipSetDevice(b->cfg.device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  free_graphs(b);
  for (Group& g : b->groups) {
    if (g.st) (void)hipStreamSynchronize(g.st);
    (void)hipFree(g.state[0]);
    (void)hipFree(g.state[1]);
    (void)hipFree(g.L0);
    (void)hipFree(g.buf);
    (void)hipFree(g.G);
    (void)hipFree(g.spec);
    (void)hipFree(g.partial);
    (void)hipFree(g.clk);
    (void)hipFree(g.tw);
    if (g.st) (void)hipStreamDestroy(g.st);
  }
  for (auto e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->own) (void)hipStreamDestroy(b->own);
  delete b;
  return CDD_OK;
}

int cdd_batch_set_stream(cdd_batch* b, void* stream) {
  if (!b) return cdd_internal_fail(CDD_ERR_ARG, "null batch");
  b->stream = (hipStream_t)stream;
  return CDD_OK;
}

int cdd_batch_size(cdd_batch* b, int i, int* n) {
  if (!b || !n || i < 0 || i >= (int)b->inst_group.size()) return cdd_internal_fail(CDD_ERR_ARG, "bad argument");
  *n = b->groups[b->inst_group[i]].n;
  return CDD_OK;
}

int cdd_batch_set_state(cdd_batch* b, int i, const double* beta, int on_device) {
  if (!b || !beta || i < 0 || i >= (int)b->inst_group.size()) return cdd_internal_fail(CDD_ERR_ARG, "bad argument");
  BCHK(hipSetDevice(b->cfg.device));
  std::vector<Clock> c;
  if (sync_clocks(b, c)) return CDD_ERR_HIP;
  Group& g = b->groups[b->inst_group[i]];
  b->set_global[i] = b->global_steps;
  b->set_nsteps[i] = c[i].nsteps;
  double* dst = inst_state(b, i, c[i]);
  BCHK(hipMemcpy(dst, beta, sizeof(double) * (size_t)g.fstride, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
  return CDD_OK;
}

int cdd_batch_get_state(cdd_batch* b, int i, double* beta, int on_device) {
  if (!b || !beta || i < 0 || i >= (int)b->inst_group.size()) return cdd_internal_fail(CDD_ERR_ARG, "bad argument");
  BCHK(hipSetDevice(b->cfg.device));
  std::vector<Clock> c;
  if (sync_clocks(b, c)) return CDD_ERR_HIP;
  Group& g = b->groups[b->inst_group[i]];
  BCHK(hipMemcpy(beta, inst_state(b, i, c[i]), sizeof(double) * (size_t)g.fstride,
                 on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost));
  return CDD_OK;
}

int cdd_batch_set_time(cdd_batch* b, int i, double t, double t_end, double fixed_dt) {
  if (!b || i < -1 || i >= (int)b->inst_group.size()) return cdd_internal_fail(CDD_ERR_ARG, "bad argument");
  BCHK(hipSetDevice(b->cfg.device));
  std::vector<Clock> c;
  if (sync_clocks(b, c)) return CDD_ERR_HIP;
  for (int k = 0; k < (int)b->inst_group.size(); ++k) {
    if (i >= 0 && k != i) continue;
    Clock cc = c[k];
    cc.t = t;
    cc.t_end = t_end;
    cc.fixed_dt = fixed_dt;
    cc.done = t >= t_end ? 1 : 0;
    Group& g = b->groups[b->inst_group[k]];
    BCHK(hipMemcpy(g.clk + b->inst_slot[k], &cc, sizeof(Clock), hipMemcpyHostToDevice));
  }
  return CDD_OK;
}

int cdd_batch_get_clock(cdd_batch* b, int i, double* t, double* last_dt, int* done, long long* nsteps) {
  if (!b || i < 0 || i >= (int)b->inst_group.size()) return cdd_internal_fail(CDD_ERR_ARG, "bad argument");
  BCHK(hipSetDevice(b->cfg.device));
  std::vector<Clock> c;
  if (sync_clocks(b, c)) return CDD_ERR_HIP;
  if (t) *t = c[i].t;
  if (last_dt) *last_dt = c[i].dt;
  if (done) *done = c[i].done;
  if (nsteps) *nsteps = c[i].nsteps;
  return CDD_OK;
}

int cdd_batch_set_tile(cdd_batch* b, int n, int variant) {
  if (!b || variant < 0 || variant > 2) return cdd_internal_fail(CDD_ERR_ARG, "bad argument");
  for (Group& g : b->groups)
    if (g.n == n) g.tile = variant;
  free_graphs(b);
  return CDD_OK;
}

int cdd_batch_set_flux(cdd_batch* b, int tile, int per_cu) {
  if (!b || tile < 0 || tile > 3 || per_cu < 0 || per_cu > (tile == 3 ? 4 : 2))
    return cdd_internal_fail(CDD_ERR_ARG, "bad argument");
  BCHK(hipSetDevice(b->cfg.device));
  if (zero_partials(b)) return CDD_ERR_HIP;
  b->flux_tile = tile;
  b->flux_per_cu = per_cu;
  free_graphs(b);
  return CDD_OK;
}

int cdd_batch_step(cdd_batch* b, int nsteps) {
  if (!b || nsteps < 0) return cdd_internal_fail(CDD_ERR_ARG, "bad argument");
  BCHK(hipSetDevice(b->cfg.device));
  for (int s = 0; s < nsteps; ++s) {
    const int phase = b->cfg.rk_order == 3 ? 0 : (int)(b->global_steps & 1);
    if (!b->exec[phase]) {
      int rc = capture(b, phase);
      if (rc) return rc;
    }
    BCHK(hipGraphLaunch(b->exec[phase], b->stream));
    b->global_steps += 1;
  }
  return CDD_OK;
}

// Autotuner hook: mean time (ms) of `reps` stage-1 flux launches of size
// group n with tile variant `variant` on the current states (no state change).
int cdd_batch_time_flux(cdd_batch* b, int n, int variant, int reps, double* ms) {
  if (!b || !ms || reps < 1 || variant < 0 || variant > 2) return cdd_internal_fail(CDD_ERR_ARG, "bad argument");
  BCHK(hipSetDevice(b->cfg.device));
  Group* gp = nullptr;
  for (Group& g : b->groups)
    if (g.n == n) gp = &g;
  if (!gp) return cdd_internal_fail(CDD_ERR_ARG, "no instances of this size");
  Group& g = *gp;
  FluxJob f;
  memset(&f, 0, sizeof f);
  f.d = 2;
  f.n = g.n;
  f.dyn = b->cfg.dynamics;
  f.cells = g.cells;
  f.A = g.state[0];
  f.G = g.G;
  f.out = g.L0;
  f.partial = g.partial;
  f.mode = 0;
  f.Lambda = b->cfg.dynamics == CDD_GLIDE_ONLY ? 0.0 : b->cfg.Lambda;
  f.mu = b->gp0.mu;
  f.lam = b->gp0.lam;
  f.stream = b->stream;
  f.count = g.count;
  f.fstride = g.fstride;
  f.gstride = g.gstride;
  f.pstride = g.nblocks;
  f.tile = variant;
  hipEvent_t e0, e1;
  BCHK(hipEventCreate(&e0));
  BCHK(hipEventCreate(&e1));
  BCHK(launch_flux(f, nullptr));  // warm-up
  BCHK(hipEventRecord(e0, b->stream));
  for (int r = 0; r < reps; ++r) BCHK(launch_flux(f, nullptr));
  BCHK(hipEventRecord(e1, b->stream));
  BCHK(hipEventSynchronize(e1));
  float t = 0.f;
  BCHK(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (zero_partials(b)) return CDD_ERR_HIP;
  *ms = t / reps;
  return CDD_OK;
}

int cdd_batch_set_spectral(cdd_batch* b, int n, int variant) {
  if (!b || variant < 0 || variant > 1) return cdd_internal_fail(CDD_ERR_ARG, "bad argument");
  for (Group& g : b->groups)
    if (g.n == n) g.stile = variant;
  free_graphs(b);
  return CDD_OK;
}

// Autotuner hook: mean time (ms) of `reps` first-stage stress solves of size
// group n (every instance) with R2C / C2R row tile `variant`, on the current
// states (writes only the group's stress planes and spectrum workspace).
int cdd_batch_time_spectral(cdd_batch* b, int n, int variant, int reps, double* ms) {
  if (!b || !ms || reps < 1 || variant < 0 || variant > 1) return cdd_internal_fail(CDD_ERR_ARG, "bad argument");
  BCHK(hipSetDevice(b->cfg.device));
  Group* gp = nullptr;
  for (Group& g : b->groups)
    if (g.n == n) gp = &g;
  if (!gp) return cdd_internal_fail(CDD_ERR_ARG, "no instances of this size");
  Group& g = *gp;
  const int s0 = g.stile;
  g.stile = variant;
  const StageIO io = stage_io(b, g, 0, 0);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = group_spectral(b, g, io, b->stream);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
  for (int r = 0; r < reps && e == hipSuccess; ++r) e = group_spectral(b, g, io, b->stream);
  if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  g.stile = s0;
  if (e != hipSuccess) return cdd_internal_fail(CDD_ERR_HIP, hipGetErrorString(e));
  *ms = t / reps;
  return CDD_OK;
}

// Autotuner hook: mean time (ms) of `reps` first-stage multi-instance flux
// launches over every eligible group with tile variant `tile` and `per_cu`
// persistent blocks per CU, on the current states (no state change).
int cdd_batch_time_flux_multi(cdd_batch* b, int tile, int per_cu, int reps, double* ms) {
  if (!b || !ms || reps < 1 || tile < 0 || tile > 3 || per_cu < 1 || per_cu > (tile == 3 ? 4 : 2))
    return cdd_internal_fail(CDD_ERR_ARG, "bad argument");
  BCHK(hipSetDevice(b->cfg.device));
  const int t0 = b->flux_tile, p0 = b->flux_per_cu;
  b->flux_tile = tile;
  b->flux_per_cu = per_cu;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = multi_flux(b, 0, 0, b->stream);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
  for (int r = 0; r < reps && e == hipSuccess; ++r) e = multi_flux(b, 0, 0, b->stream);
  if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  b->flux_tile = t0;
  b->flux_per_cu = p0;
  if (e != hipSuccess) return cdd_internal_fail(CDD_ERR_HIP, hipGetErrorString(e));
  if (zero_partials(b)) return CDD_ERR_HIP;
  *ms = t / reps;
  return CDD_OK;
}

}  // extern "C"
