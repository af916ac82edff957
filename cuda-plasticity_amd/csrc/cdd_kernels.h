// Internal launch interface between the C-ABI layer and the kernel files.
#pragma once
#include "cdd_common.h"
#include "cdd_fft.h"

struct cdd_config;

namespace cdd {

// A stencil/source field: value = A + dt * B (B optional, dt from clk).
template <typename T>
struct SrcField {
  const T* A;
  const T* B;
  const Clock* clk;
  long long cs;  // component stride (cells)
  T* W;          // optional: the formed A + dt B is also stored here (may alias B)
  long long istride;  // batched 2D: per-instance stride of A, B, W (elements)
};

struct GreenParams {
  double A, C;      // lam/(lam+2mu), 2(lam+mu)/(lam+2mu)
  double mu, lam;
  double norm;      // 1 / N^d
  double t;         // stage time when clk == nullptr
  const Clock* clk;
  LoadParams load;
  Elastic el;
  int log_stress;   // 1: the k = 0 thread appends the clock's stress-strain record
  double vgamma;    // nc = 1: diffusion filter exp(-vgamma dt k^2) (dt from clk)
  int curl;         // nc = 9: the disorder stress sum_lt eps_lit k_l F_tj (real k_l) in
                    // place of the Nye tensor's i k_l
  int rho_sigma;    // nc = 9: the free-boundary stress K_ijmn rho_mn of a Nye tensor (six
                    // symmetric outputs, the last three planes zero)
  int sink;         // 1: the sinktools' k vectors N sin(2 pi n / N) in place of 2 pi n_wrapped
  int mform;        // nc = 6: the whole stress by the M-tensor form (source 'betaP'), its
                    // k = 0 mode the mean correction; no load
};

// Called after each kernel launch with the kernel kind (HIP-event timing).
enum KernelKind {
  KK_R2C = 0, KK_COLS_FWD = 1, KK_CONTRACT = 2, KK_COLS_INV = 3, KK_C2R = 4,
  KK_FLUX1 = 5, KK_FINALIZE = 6, KK_FLUX2 = 7, KK_FLUX3 = 8, KK_ADVANCE = 9, KK_OTHER = 10, KK_COUNT = 11
};
typedef void (*LaunchHook)(void* ctx, int kind);

struct SpectralJob {
  LaunchHook hook;
  void* hook_ctx;
  int d, n, nc;          // nc = 6 (stress) or 9 (Nye tensor)
  long long cells;
  SrcField<double> src;
  void* spec;            // complex workspace, nc * (n/2+1) * n^(d-1)
  const void* tw;        // twiddle table for n
  double* out;           // nc real planes
  GreenParams gp;
  hipStream_t stream;
  // batched 2D: `count` instances of the same n with uniform strides; the
  // instance index is blockIdx.y of every 2D pass and clocks are gp.clk[i]
  int count;
  long long sstride;     // spectrum (complex elements)
  long long ostride;     // out (elements)
  int trans;             // 1: transposed Green's contraction (adjoint), no load mode
  int fp32;              // 1: src, spec, tw and out hold float data (src typed as a carrier)
  // vacancy coupling (fp64, nc = 6): the diagonal stress planes of out get
  // -valpha * vac (the vacancies' chemical stress)
  const double* vac;
  double valpha;
  const double* valpha_f;  // optional: alpha as a field (one plane), in place of valpha
  // R2C / C2R row tile: 0 the default rows per block, 1 half of them (2D
  // fp64 stress solves of sides >= 256; otherwise the default)
  int stile;
  // nc = 6: the source holds the six symmetric components (xx yy zz xy xz
  // yz) as planes, so the R2C reads them as they are (the flux adjoint's
  // stress cotangent); 0: it symmetrises a nine-component source
  int src_sym6;
  // working set that picks the passes' cache policy (streaming above
  // kStreamBytes); 0: this launch's spectrum.  Launches that run beside
  // others (the batch's size groups) give the spectrum of all of them.
  long long ws_bytes;
};
// Working sets above the 256 MB infinity cache stream (non-temporal loads
// and stores) through the spectral passes.
constexpr long long kStreamBytes = 256ll << 20;

hipError_t launch_spectral(const SpectralJob& job);
hipError_t launch_sym6_to_9(double* s, long long cells, hipStream_t st);
hipError_t launch_rho_symmetric(const double* b, double* rho, int d, int n, long long cells, hipStream_t st);
// the twiddle table W_n^e = exp(-2 pi i e / n), e < n, as 2 n doubles (host)
void fill_twiddles(double* h, int n);
hipError_t launch_sigma_finalize(const double* G, const double* beta, double* sig, long long cells, double mu,
                                 double lam, hipStream_t st, int fp32 = 0);

enum Dyn { DYN_GLIDE_CLIMB = 0, DYN_GLIDE_ONLY = 1, DYN_WOJZZ = 2 };

struct FluxJob {
  int d, n, dyn;
  long long cells;
  const double* A;     // stencil source
  const double* B;     // optional increment: f = A + dt * B
  const double* G;     // 6 planes of the non-local stress
  const double* U;     // optional state for the RK combination
  double* out;         // 9 planes
  double* speed;       // optional per-cell speed
  double* partial;     // optional per-block max of the speed
  Clock* clk;          // dt (for B and the combination)
  int mode;            // 0: out = rhs; 1: out = cu U + cf (f + dt rhs)
  double cu, cf;
  double Lambda, mu, lam;
  hipStream_t stream;
  // batched 2D (instance = blockIdx.y): strides of A/U/out, G and partial
  int count;
  long long fstride, gstride, pstride;
  int tile;            // flux tile variant (0: default, 1: 64x4, 2: 16x16)
  int fp32;            // 1: A, G, U, out, speed hold float data (partial stays double)
  // vacancy coupling (fp64, mode 1, the step's last stage): vac += vbeta
  // (tr out - tr U) at every cell
  double* vac;
  double vbeta;
  // 1: the CUDA executable's local Lax-Friedrichs flux (fp64; glide + climb
  // or glide only), plasticity_kernel.cu centralHJ :118-697
  int llf;
  // disorder (fp64, one instance, per-tile kernel): S9 (nine planes) is
  // added to the stress at every cell; Ad, the stencil source of an
  // advected scalar (the disorder field), gives outd = its rhs (mode 0) or
  // cu Ud + cf (Ad + dt rhs) (mode 1)
  const double* S9;
  const double* Ad;
  const double* Ud;
  double* outd;
  // fp64, per-tile kernel: a fixed source added to every cell's rhs,
  // rhs += src_rate * src (nine planes; the CUDA executable's dynamic
  // nucleation, plasticity_kernel.cu:1089-1091)
  const double* src;
  double src_rate;
};

hipError_t launch_flux(const FluxJob& job, int* nblocks_out);

// Persistent multi-instance 2D flux: one launch covers every instance of up
// to kFluxGroups size groups.  The work items are (group, instance, tile) in
// group order; a fixed grid of blocks walks them with a stride of the grid
// size.
// Per group: instance i's fields are A + i fstride (likewise U and out),
// G + i gstride, its clock clk + i and its per-tile CFL partials
// partial + i pstride (pstride >= the instance's tile count).
constexpr int kFluxGroups = 10;
struct FluxGroup {
  int n, count, ntiles, item0;  // side, instances, tiles per instance, first item
  long long cells, fstride, gstride, pstride;
  const double* A;
  const double* G;
  const double* U;
  double* out;
  double* partial;
  Clock* clk;
};
struct FluxMultiJob {
  int ngroups, nitems;  // filled by launch_flux_multi
  int dyn, mode;        // mode as FluxJob (0: out = rhs; 1: out = cu U + cf (f + dt rhs))
  int tile;             // tile variant (0: 32x8, 1: 64x4, 2: 16x16; 3: row marching)
  int per_cu;           // persistent blocks per CU (tiles: 1 or 2, 0: 2; rows: 1..4, 0: 2)
  double cu, cf, Lambda, mu, lam;
  FluxGroup g[kFluxGroups];
};
// whether a group of side n can join a multi-instance launch with this tile
bool flux_multi_ok(int n, int tile);
// fills ngroups' item ranges and nitems from g[0..ngroups) and launches
hipError_t launch_flux_multi(FluxMultiJob& job, int ngroups, hipStream_t st);
// The row-marching variant (sides that are multiples of 128): every block
// marches an equal share of the (group, instance, strip, row) space; each
// instance's CFL maximum is maximised atomically into partial[0] of the
// instance (pstride apart), which must be zero before the launch (batch
// creation zeroes it; launch_finalize_dt with clear resets it).  per_cu:
// blocks per CU (1..4, 0: 2).  ntiles/item0 become row-step counts.
bool flux_rows_ok(int n);
hipError_t launch_flux_rows(FluxMultiJob& job, int ngroups, hipStream_t st);
// the Green's-tensor, elastic and load constants of a configuration
GreenParams green_params(const cdd_config& cfg);
int flux_nblocks(int d, int n);
// clear (with nblocks = 1): the partial is one slot that k_flux_rows
// maximises atomically; it is reset to zero once read
hipError_t launch_finalize_dt(const double* partial, int nblocks, Clock* clk, double nmax, hipStream_t st,
                              int count = 1, long long pstride = 0, double dt_bound = 0.0, int clear = 0);
hipError_t launch_advance(Clock* clk, hipStream_t st, int count = 1);
// the same for ng instance groups in one launch per ten groups (group g:
// count[g] clocks at clk[g], speed partials at partial[g] + i pstride[g])
hipError_t launch_finalize_dt_groups(int ng, double* const* partial, const int* nb, const long long* pstride,
                                     Clock* const* clk, const double* nmax, const int* count, const int* clear,
                                     hipStream_t st);
hipError_t launch_advance_groups(int ng, Clock* const* clk, const int* count, hipStream_t st);
// forward Euler update u += dt L with the clock's dt (skipped once done)
hipError_t launch_euler(double* u, const double* L, long long n, const Clock* clk, hipStream_t st);
// the vacancy component's Euler step: c += vbeta dt tr(L)
hipError_t launch_vacancy_euler(double* c, const double* L, long long cells, double vbeta, const Clock* clk,
                                hipStream_t st);
// out = u + dt L with the clock's dt (out may alias L; skipped once done)
hipError_t launch_axpy_dt(double* out, const double* u, const double* L, long long n, const Clock* clk,
                          hipStream_t st);
// rho'_ij = d rho_ij / (ME + sum_mn 2 rho_mn^2), in place over the nine planes
hipError_t launch_disorder_weight(double* rho, const double* d, long long cells, hipStream_t st);
int flux_nblocks_tile(int d, int n, int tile);
hipError_t launch_energy(const double* G, const double* beta, long long cells, double mu, double nu, double lam,
                         double* partial, int* nblocks, hipStream_t st, int fp32 = 0);
hipError_t launch_reduce_sum(const double* partial, int n, double* out, hipStream_t st, int accumulate = 0);
// append the sum of the k_energy partials to the clock's energy log
hipError_t launch_energy_log(const double* partial, int nb, Clock* clk, hipStream_t st);
// per-block partials of sum over cells of (-1)^(i+j[+k]) |rho|
hipError_t launch_nyquist_mode(const void* rho, int d, int n, long long cells, int fp32, double* partial, int* nb,
                               hipStream_t st);
// reverse all grid axes of ncomp components (NumPy layout <-> the reference's
// transposed .ics/.plas layout); in and out must not alias
hipError_t launch_reverse_axes(const void* in, void* out, int d, int n, int ncomp, int fp32, hipStream_t st);

struct VjpJob {
  int d, n, dyn;
  long long cells;
  const double* beta;  // state the flux is linearised at
  const double* G;     // its non-local stress planes
  const double* rbar;  // cotangent of the flux (9 planes), used as rscale * dt * rbar
  double rscale;
  const double* dtp;   // optional device dt (the factor is 1 when null)
  double* ubar;        // out: 2 d 9 derivative-cotangent planes, scaled by n
  unsigned long long* codes;  // out: per-cell minmod selections (2 bits per axis and component)
  double* sbar;        // out: symmetrised stress cotangent (6 planes, xx yy zz xy xz yz)
  double* lpart;       // out: per-block Lambda cotangent
  double* upart;       // out (optional): per-block sum of lam tr(sbar) + 2 mu sbar_zz,
                       // the cotangent of the uniaxial control's betaE_zz
  double Lambda, mu, lam;
};
// out = J_L(beta)^T rbar assembled from the derivative cotangents, the stress
// cotangent and (optional) its spectral transpose ebar, plus sum_k coef_k term_k
struct GatherJob {
  int d, n;
  long long cells;
  const unsigned long long* codes;
  const double* ubar;
  const double* sbar;  // 6 planes, xx yy zz xy xz yz
  const double* ebar;
  double mu, lam;
  int nterm;
  const double* term[3];
  double coef[3];
  const double* uzz;    // optional device scalar: uzz_scale * (*uzz) is added to component zz
  double uzz_scale;
  const double* csrc;   // optional: 9 device constants added per component
  double* out;
};
// per-block partials of lam tr(sbar) + 2 mu sbar_zz (uniaxial control)
hipError_t launch_uniaxial_partial(const double* sbar, long long cells, double lam, double mu, double* partial,
                                   int* nblocks, hipStream_t st);
// out[zz] += scale * (*s) at every cell
hipError_t launch_add_uniform_zz(double* out, long long cells, const double* s, double scale, hipStream_t st);
int flux_vjp_nblocks(int d, int n);
hipError_t launch_clock_segment(Clock* clk, int mode, double t0, const double* dts, long long len, hipStream_t st);
hipError_t launch_flux_vjp_cell(const VjpJob& j, int* nblocks, hipStream_t st);
hipError_t launch_flux_vjp_gather(const GatherJob& g, hipStream_t st);
hipError_t launch_energy_grad(const double* G, const double* beta, long long cells, double mu, double nu, double lam,
                              double* sbar, double* partial, int* nblocks, hipStream_t st);
// The flux with every branch of the scheme fixed (the branch-frozen map the
// adjoint differentiates): two words per cell and stage in sel,
//   sel[idx]:         the minmod selection of the cell's own limiter site
//                     m_i per axis a and component c, 2 bits at 2 (9 a + c)
//                     (0, 1, 2: the first, middle, last argument; 3: zero)
//   sel[cells + idx]: per axis a, 8 bits at 8 a: the corner with the largest
//                     velocity component a (bits 0-2), the corner with the
//                     smallest (bits 3-5), whether the largest is > 0 (bit 6:
//                     a_plus is it, else 0) and whether the smallest is < 0
//                     (bit 7).
// record = 1 decides the branches at beta (the ordinary flux) and writes
// them; record = 0 takes them from sel, which makes the flux a smooth
// function of beta (and of Lambda).  With dbeta set (the tangent mode) the
// branches are decided at beta, sel is unused, and drhs receives the
// tangent-linear flux J_L(beta) dbeta + (dL/dLambda) dLambda, dG being the
// non-local stress planes of dbeta (its load's tangent included).
struct FrozenJob {
  int d, n, dyn;
  long long cells;
  const double* beta;
  const double* G;      // the non-local stress planes of beta
  unsigned long long* sel;
  int record;
  double* rhs;          // 9 planes
  double Lambda, mu, lam;
  const double* dbeta;  // tangent mode: the tangent of beta (9 planes)
  const double* dG;     // and of G (6 planes)
  double dLambda;
  double* drhs;         // out: the tangent of rhs (9 planes)
};
hipError_t launch_flux_frozen(const FrozenJob& j, hipStream_t st);
hipError_t launch_lincomb(long long len, int nin, const double* a0, double c0, const double* a1, double c1,
                          const double* a2, double c2, const double* a3, double c3, double* out, hipStream_t st);

}  // namespace cdd
