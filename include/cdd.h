/* C ABI of the MI355X-native continuum-dislocation-dynamics (CDD) stepper.
 *
 * The reference crosses its native boundary as a process + file exchange:
 * RunCUDA.simulation (Plasticity/RunCUDA.py:120-254) writes a transposed
 * .ics state (:9-22) and a #define header (:26-29, 54-95), runs
 * `make HEADER=...` (:214) and `./exe --device=N` (:225), and reads back
 * .plas records [t, 9 N^d] (assist.h:164-192).  Inside the process the hot
 * path is plasticity.cu runTest (:40-376) driving calculateFlux /
 * calculateSigma / centralHJ / simpleTVD of plasticity_kernel.cu
 * (:951-1206).  On the Python side the same path is the
 * PlasticitySystem -> FieldMover -> FieldDynamics -> PlasticityState call
 * chain.  This library replaces all of it with in-process calls on a plan
 * that owns its device buffers; the Python host layer
 * (cuda-plasticity_amd/Plasticity) binds these symbols through ctypes.
 *
 * Conventions
 *   - Fields are float64 arrays of shape (9, N0, N1[, N2]), C-contiguous,
 *     component c = 3 i + j in the sorted order xx xy xz yx yy yz zx zy zz
 *     (Plasticity/Fields/Fields.py:21, 247-249): NumPy layout, last axis
 *     contiguous.  Grids are square (side n) and periodic.
 *   - Pointers marked `dev` are device pointers on the plan's device
 *     (e.g. torch HIP tensors); they are borrowed for the call only and used
 *     in stream order on the plan's stream.
 *   - Every function returns 0 on success and a negative code on failure;
 *     cdd_last_error() returns the message of the calling thread's last
 *     failure.  The Python layer raises RuntimeError, as RunCUDA does on a
 *     non-zero exit status (Plasticity/RunCUDA.py:221-228).
 *   - Calls on different plans may run concurrently from different host
 *     threads (the exec_batch.py pattern); one plan is not thread-safe.
 *   - cdd_step never synchronises with the host: the CFL time step is
 *     reduced and kept on the device.  Functions that return host values
 *     (cdd_get_clock, cdd_energy) synchronise the plan's stream.
 */
#ifndef CDD_H
#define CDD_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CDD_OK = 0,
  CDD_ERR_ARG = -1,
  CDD_ERR_HIP = -2,
  CDD_ERR_UNSUPPORTED = -3
};

/* CDD_GLIDE_CLIMB_WOJZZ: the 2D variant of
 * Plasticity/FieldDynamics/CentralUpwindHJBetaPDynamics_woJzz.py (climb
 * suppression with tr(sigma)/2 along (ux_yy - uy_xy, uy_xx - ux_yx,
 * -(ux_zx + uy_zy)), its J term on H_xx and H_yy only, beta_zz held fixed);
 * the adjoint and batched entry points do not support it. */
enum { CDD_GLIDE_CLIMB = 0, CDD_GLIDE_ONLY = 1, CDD_GLIDE_CLIMB_WOJZZ = 2 };

/* Arithmetic and storage type of a plan's fields.  The reference computes in
 * float64 only (plasticity.h DOUBLE, numpy float64); CDD_FP32 is an added
 * variant whose error is stated against the float64 run.  Field buffers
 * (void* below) hold float64, or float32 for a CDD_FP32 plan; the adjoint
 * and batched entry points are float64 only. */
enum { CDD_FP64 = 0, CDD_FP32 = 1 };

enum {
  CDD_LOAD_NONE = 0,
  CDD_LOAD_STRAIN = 1,   /* RunPython.py:21-26,44-49 ExternalStrain, strain(t) = init + rate (t - t0) */
  CDD_LOAD_STRESS = 2,   /* RunPython.py:28-32 ExternalStress */
  CDD_LOAD_UNIAXIAL = 3, /* RunLoadingMPI.py:68-83 strain-rate control along z */
  CDD_LOAD_TENSOR = 4    /* sigma_ij += init_ij + rate_ij (t - t0); e.g. uniform shear */
};

/* Discretisation of the Hamilton-Jacobi flux.  CDD_SCHEME_CENTRAL_UPWIND is
 * the Python stepper's (CentralUpwindHJ3D.py:63-271: one-sided speeds per
 * axis, sqrt(|rho|^2 + ME)); CDD_SCHEME_LLF is the CUDA executable's
 * (plasticity_kernel.cu centralHJ :118-697 under plasticity.h's LLF, DOUBLE,
 * theta = 1: one local Lax-Friedrichs speed per cell, |rho| + sqrt(ME), its
 * own limiter form, dt = CFL / max speed / N and an end step of
 * t_end - t + ME, simpleTVD :1112-1137). */
enum { CDD_SCHEME_CENTRAL_UPWIND = 0, CDD_SCHEME_LLF = 1 };

typedef struct cdd_config {
  int ndim;          /* 2 or 3 */
  int n;             /* grid side: 8..512 powers of two in 3D; 8..1024 and 384, 768 in 2D */
  int dynamics;      /* CDD_GLIDE_CLIMB ('mdp' Lambda=1, 'gcd' Lambda=0, partial climb) or CDD_GLIDE_ONLY ('lvp') */
  int rk_order;      /* 3: Shu-Osher TVD-RK3 (FieldMover.py:61-141, plasticity_kernel.cu TVD3rd
                        :1140-1206); 2: SSP-RK2; 1: forward Euler (simpleTVD, CDD_SCHEME_LLF only) */
  double Lambda;     /* climb suppression (CentralUpwindHJBetaPDynamics3D.py:126-133) */
  double mu, nu;     /* isotropic elasticity (PlasticityState.py:23) */
  double cfl;        /* CFLsafeFactor (FieldMover.py:16,96) */
  int load_kind;     /* CDD_LOAD_* */
  int device;        /* HIP device ordinal */
  double load_init[9];
  double load_rate[9];
  double load_t0;
  double load_init_strain;
  double load_uniaxial_rate;
  int precision;     /* CDD_FP64 (default, 0) or CDD_FP32 */
  int scheme;        /* CDD_SCHEME_CENTRAL_UPWIND (default, 0) or CDD_SCHEME_LLF (float64; glide+climb
                        or glide-only; no batching, vacancies or adjoint) */
  double load_mu;    /* elastic constants of the CDD_LOAD_STRAIN stress (lam tr e + 2 mu e); 0: the plan's
                        mu, nu.  RunPython's ExternalStrain uses its module constants mu = 0.5, nu = 0.3
                        whatever the state's (Plasticity/RunPython.py:19-26), as does
                        PlottingTools.StrainStress (Plasticity/PlottingTools.py:73-74) */
  double load_nu;
} cdd_config;

typedef struct cdd_plan cdd_plan;

/* Create / destroy a stepper plan (replaces plasticity.cu:40-240 set-up and
 * plasticity_kernel.cu setupSystem :931-949). */
int cdd_plan_create(const cdd_config* cfg, cdd_plan** out);
int cdd_plan_destroy(cdd_plan* plan);

/* Run the plan's work on an existing hipStream_t, e.g. torch's current
 * stream; NULL is the default (null) stream.  A new plan uses a private
 * non-blocking stream. */
int cdd_set_stream(cdd_plan* plan, void* stream);

/* State transfer (replaces the .ics write / .plas read, RunCUDA.py:9-22 and
 * plasticity.cu:139-206,268-270).  on_device: 1 if the pointer is a device
 * pointer, 0 for host memory. */
int cdd_set_state(cdd_plan* plan, const void* beta, int on_device);
int cdd_get_state(cdd_plan* plan, void* beta, int on_device);

/* Clock: time at the start of the next step and Run's endTime
 * (FieldMover.py:20-36).  fixed_dt > 0 selects CalculateOneFixedTimeStep. */
int cdd_set_time(cdd_plan* plan, double t, double t_end, double fixed_dt);
int cdd_get_clock(cdd_plan* plan, double* t, double* last_dt, int* done, long long* nsteps);

/* sigma = PlasticityState.CalculateSigma(type='k', source='betaE') of beta
 * (PlasticityState.py:85-127,357-390) plus the plan's load at time t (the
 * GetSigma override, RunPython.py:44-49).  beta_dev, sigma_dev: (9, N^d). */
int cdd_sigma(cdd_plan* plan, const void* beta_dev, double t, void* sigma_dev);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the stress of cdd_sigma for the configuration cfg
 * (ndim, n, mu, nu, load, load_mu/load_nu, device; float64 only) without a
 * plan and without any allocation.  beta_dev, sigma_dev: (9, N^d) float64
 * device pointers on cfg->device; stream: a hipStream_t of that device
 * (NULL: the null stream).  The workspace is a static device array of 64 MiB
 * per GPU, held until the process ends; grids that do not fit are refused
 * (2D up to 768^2, 3D up to 64^3).  The first call on a device fills the
 * array's twiddle tables with a synchronous copy: make it before any graph
 * capture.  Calls that overlap on one GPU share the array and race. */
int cdd_sigma_noalloc(const cdd_config* cfg, const void* beta_dev, double t, void* sigma_dev, void* stream);

/* rho = PlasticityState.CalculateRhoFourier(type='k') (PlasticityState.py:264-304). */
int cdd_rho(cdd_plan* plan, const void* beta_dev, void* rho_dev);

/* The state's other stress and Nye-tensor forms, off the stepper's path
 * (float64 plans; no external load):
 *   cdd_sigma_variant: PlasticityState.CalculateSigma(type, source)
 *     (PlasticityState.py:85-230).  ktype CDD_KTYPE_K (the state's ktools)
 *     or CDD_KTYPE_SINK (its sinktools, k = N sin(2 pi n / N),
 *     FourierSpaceTools.py:13-61, PlasticityState.py:35-36); source
 *     CDD_SOURCE_BETAE (fixed volume; like the reference it ignores ktype),
 *     CDD_SOURCE_BETAP (the M-tensor form, same result for 'k') or
 *     CDD_SOURCE_RHO (free boundary: K_ijmn rho_mn of the Nye tensor of
 *     that ktype, zero mean).
 *   cdd_rho_variant: CalculateRhoFourier(type) (PlasticityState.py:264-304).
 *   cdd_rho_symmetric: CalculateRhoSymmetric (PlasticityState.py:232-262):
 *     the Nye tensor from central differences 0.5 N (f(i+1) - f(i-1))
 *     (NumericalMethods.py:6-13); rho_dev must not alias beta_dev.
 * beta_dev, sigma_dev, rho_dev: (9, N^d) float64 device arrays. */
#define CDD_KTYPE_K 0
#define CDD_KTYPE_SINK 1
#define CDD_SOURCE_BETAE 0
#define CDD_SOURCE_BETAP 1
#define CDD_SOURCE_RHO 2
int cdd_sigma_variant(cdd_plan* plan, const void* beta_dev, int ktype, int source, void* sigma_dev);
int cdd_rho_variant(cdd_plan* plan, const void* beta_dev, int ktype, void* rho_dev);
int cdd_rho_symmetric(cdd_plan* plan, const void* beta_dev, void* rho_dev);

/* rhs (9, N^d) and the CFL speed field a_tot + b_tot (+ c_tot) (N^d,
 * nullable) of BetaPDynamics.CalculateFlux(t, state, CFLCondition=True)
 * (CentralUpwindHJBetaPDynamics3D.py:214-221 / GlideOnly :255-262 ->
 * CentralUpwindHJ3D.py:63-271). */
int cdd_flux(cdd_plan* plan, const void* beta_dev, double t, void* rhs_dev, void* speed_dev);

/* Advance the plan's state by nsteps TVD-RK steps
 * (TVDRungeKutta_FieldMover.CalculateOneAdaptiveTimeStep, FieldMover.py:61-141;
 * plasticity.cu:242-271 loop).  Stops on the device once t >= t_end; no host
 * synchronisation. */
int cdd_step(cdd_plan* plan, int nsteps);

/* Vacancy coupling: switches the plan's stepping to the operator-splitting
 * mover of a VacancyState (OperatorSplittingTVDRK_FieldMover, FieldMover.py:
 * 318-380, with BetaP_VacancyDynamics, VacancyDynamics.py:16-70; the CUDA
 * executable's calculateKDiffusion, plasticity_kernel.cu:700-775).  Every
 * stress solve of a step sees sigma_ii - alpha c with the concentration c of
 * the step start (GetSigma); after the step c += beta (tr beta_new -
 * tr beta_old) and c is diffused by exp(-gamma dt k^2) in k space
 * (DiffusionEvolution); the adaptive dt is capped at dt_bound (> 0; the
 * mover's dtBound) before the end-time clamp.  cdd_flux includes the
 * -alpha c term (BetaP_VacancyDynamics.CalculateFlux); cdd_sigma does not
 * (PlasticityState.CalculateSigma).  float64 plans only; the adjoint refuses
 * such a plan.  c starts at zero; enable = 0 switches the coupling off and
 * keeps c.  The reference's fixed-step variant (:322-348) recouples the
 * step-start beta^P instead of the advanced one; here both keep the
 * advanced state (its adaptive variant, :349-380).
 *   cdd_vacancy_set / cdd_vacancy_get: the concentration field c (N^d
 *   doubles, the VacancyState's ('s','s') component), host or device. */
int cdd_set_vacancy(cdd_plan* plan, int enable, double alpha, double beta, double gamma, double dt_bound);
int cdd_vacancy_set(cdd_plan* plan, const void* c, int on_device);
/* alpha as a field (N^d float64 values, host or device) in place of the
 * scalar of cdd_set_vacancy: sigma_ii -= alpha(x) c(x) (the dynamics'
 * alpha * cfield, VacancyDynamics.py:24-28, with an array alpha).  NULL
 * returns to the scalar. */
int cdd_vacancy_alpha(cdd_plan* plan, const void* alpha, int on_device);
int cdd_vacancy_get(cdd_plan* plan, void* c, int on_device);

/* The CUDA executable's DYNAMIC_NUCLEATION (plasticity.cu:178-201,
 * plasticity_kernel.cu:31-33,1089-1091,1122-1125): every flux evaluation
 * gains rate * src (the executable's rhs += beta0dot / TIME_FINAL, every
 * stage) and the adaptive time step is capped at dt_cap (its
 * maxNucleationTimestep) before the end-time clamp.  src: 9 N^d float64
 * values, host (on_device = 0) or device; NULL switches the source off.
 * float64 plans; excludes the vacancy coupling.  The source does not
 * depend on the state, so the adjoint is unchanged. */
int cdd_set_source(cdd_plan* plan, const void* src, int on_device, double rate, double dt_cap);

/* Disorder (FieldDynamics/DisorderDynamics.py:17-44,
 * PlasticityStates/DisorderState.py:11-30, RunPython.py:196-214):
 * the stress gains sum_{l,t} eps_lit irfftn(k_l rfftn(rho'_tj)) with
 * rho' = d rho / (ME + 2 |rho|^2), rho the spectral Nye tensor, at every
 * stage.  advect = 1: d is the state's ('s', 's') component of a
 * DisorderState, carried by the same central-upwind flux (H_s = v . grad d)
 * and RK stages; advect = 0: d is the fixed field of BetaP_DisorderDynamics.
 * dfield: N^d float64 values; NULL disables.  float64, central-upwind
 * scheme; excludes the vacancy coupling, the adjoint and batching.
 * cdd_disorder_get reads the (advected) field back. */
int cdd_set_disorder(cdd_plan* plan, const void* dfield, int on_device, int advect);
int cdd_disorder_get(cdd_plan* plan, void* dfield, int on_device);
/* DisorderState.CalculateSigma for beta (9 planes) and d (device pointers,
 * float64): the state's stress plus the disorder term (9 planes, not
 * symmetric). */
int cdd_disorder_sigma(cdd_plan* plan, const double* beta, const double* dfield, double t, double* sigma);
/* The flux of the disorder dynamics at (beta, d): rhs of beta (9 planes),
 * rhs_d of the advected field (nullable: the fixed-field dynamics) and the
 * CFL speed (nullable); device pointers. */
int cdd_disorder_flux(cdd_plan* plan, const double* beta, const double* dfield, double t, double* rhs, double* rhs_d,
                      double* speed);

/* Total elastic energy sum_cells 1/2 S sigma sigma of beta (nullable: the
 * plan's state) at time t (PlasticityState.py:317-330, Observer.py:86). */
int cdd_energy(cdd_plan* plan, const void* beta_dev, double t, double* energy);

/* Device pointer of the plan's current state (9, N^d) after the last
 * synchronising call. */
int cdd_state_ptr(cdd_plan* plan, void** beta_dev);

/* HIP-event timing of the kernels cdd_step launches (the tracing the
 * reference does with cutil timers, plasticity.cu:169-170,227-228,320-350):
 * enable with room for max_events records (0 disables), run steps, then
 * read per-kind totals in ms and launch counts.  Kinds: 0 r2c, 1 column FFT
 * forward, 2 Green's contraction pass, 3 column FFT inverse, 4 c2r,
 * 5 flux stage 1, 6 CFL reduction, 7 flux stage 2, 8 flux stage 3,
 * 9 clock advance. */
/* Stress-strain log (replaces the per-state post-processing of
 * PlottingTools.StrainStress, Plasticity/PlottingTools.py:69-100, and the
 * sigma_zz of RunLoadingMPI's EnergyDownhillChecking,
 * Plasticity/MattTools/Run/RunLoadingMPI.py:20-58): from the next step on,
 * every step appends 8 doubles to log_dev (device, cap records): the step's
 * start time t, the mean stress including the external load (xx, yy, zz,
 * xy, xz, yz) of the state at t, and mean beta_zz.  Recorded by the stress
 * solve itself (its k = 0 mode): no extra pass, no host sync.  cap = 0 stops. */
int cdd_stress_record(cdd_plan* plan, double* log_dev, long long cap);

/* Device-side observers and failure detectors (no per-step host sync):
 *   cdd_energy_record: every step appends the total elastic energy of its
 *     start state (loads included) to log_dev (cap doubles; 0 stops): the
 *     series behind Observer.EnergyDownhillObserver (Observer.py:89-104) and
 *     RunLoadingMPI's EnergyDownhillChecking (RunLoadingMPI.py:20-58).
 *   cdd_status: 0, or 1 when a non-finite speed or right-hand side appeared
 *     in the step's first stage; the run then stops on the device with the
 *     state left at that step's start (a NaN guard the reference does not
 *     have).
 *   cdd_instability: |FFT(|rho|)| at the mode (N/2, ..., N/2) of beta
 *     (nullable: the plan's state), the measure of
 *     LargeKInstabilityChecking_FieldMover.CheckInstability
 *     (Plasticity/FieldMovers/LargeKInstabilityChecking_FieldMover.py:73-89),
 *     computed as sum over cells of (-1)^(i+j[+k]) |rho|. */
int cdd_energy_record(cdd_plan* plan, double* log_dev, long long cap);
/* Speed log: from the next step on, every step appends its CFL speed (the
 * largest one-sided corner speed of its first stage, the max that
 * FieldMover.CalculateOneAdaptiveTimeStep turns into dt, FieldMover.py:94-96)
 * to log_dev (cap doubles; 0 stops).  Logged also under a fixed time step
 * (CalculateOneFixedTimeStep), so a fixed-step run can check its CFL number
 * N max speed dt afterwards. */
int cdd_speed_record(cdd_plan* plan, double* log_dev, long long cap);
int cdd_status(cdd_plan* plan, int* status);
int cdd_instability(cdd_plan* plan, const void* beta_dev, double* value);

/* The reference's native output format (plasticity.cu:102-160,239,270;
 * assist.h:101-192): a .plas file is a sequence of records [t, 9 N^d], each
 * component stored with every grid axis reversed (the layout of
 * RunCUDA.convertStateToCUDA, Plasticity/RunCUDA.py:9-22), in the plan's
 * precision.  cdd_plas_append appends the current state and time (a
 * synchronising call).  cdd_plas_resume replaces the CONTINUE_RUN restart
 * (plasticity.cu:107-125): it loads the last complete record into the plan's
 * state and clock; unlike the reference's seek-from-end it checks that the
 * file holds whole records and skips a torn tail left by an interrupted
 * append (*torn = 1; *records = complete records; both nullable).  With the
 * vacancy coupling on, each record carries the concentration c as a tenth
 * component (the executable's VACANCIES records, [t, 10 N^d]). */
int cdd_plas_append(cdd_plan* plan, const char* path);
int cdd_plas_resume(cdd_plan* plan, const char* path, double* t, long long* records, int* torn);

int cdd_timing_enable(cdd_plan* plan, int max_events);
int cdd_timing_read(cdd_plan* plan, double* ms, int* counts, int nkinds);

/* ---- Adjoint (no reference counterpart: the reference has no gradients) --
 * Vector-Jacobian products with the time-step schedule frozen: the CFL
 * reduction (FieldMover.py:94-96) is treated as a constant.
 *   cdd_flux_vjp:    betabar = (dL/dbeta)^T rbar for the flux L of
 *                    BetaPDynamics.CalculateFlux at time t; *lambar =
 *                    (dL/dLambda) . rbar (nullable).
 *   cdd_step_vjp:    ubar = (du'/du)^T ubar_next for one TVD-RK step
 *                    u -> u' with step dt at time t (the plan's rk_order);
 *                    *lambar = (du'/dLambda) . ubar_next.
 *   cdd_energy_grad: total elastic energy of beta at t and its gradient.
 *   cdd_dt_record:   log every step's dt into a device array (cap entries).
 *   cdd_dt_replay:   take the next len steps' dt from a device array
 *                    (replaying a recorded schedule exactly).
 *   cdd_segment_vjp: reverse pass over one checkpoint segment: recompute
 *                    nsteps steps from the checkpoint u0 at time t0
 *                    replaying dts_dev (the recorded dt of those steps),
 *                    keeping every stage state and its stress planes, then
 *                    carry the cotangent of the segment's final state
 *                    (ubar_end) back to ubar0 (may alias ubar_end).
 *                    src_dev (nullable, nsteps x 9 device doubles): a
 *                    uniform per-component gradient of the objective at the
 *                    state starting step j, added to its cotangent (time-
 *                    distributed objectives such as a mean stress over the
 *                    last steps).  The plan's clock and state are left
 *                    unchanged; the Lambda cotangent is added to a device
 *                    accumulator.  No host synchronisation.  Workspace:
 *                    15 rk_order planes per step.
 * All loads are supported; the uniaxial strain-rate control's dependence on
 * <beta_zz> is differentiated (a uniform term in betabar_zz).
 *   cdd_lambda_bar:  read (nullable) and optionally reset that accumulator.
 * Stored trajectories (no counterpart in the reference, which has no
 * adjoint): when device memory allows, the forward run keeps each segment's
 * stage states so the reverse pass needs no recomputation.
 *   cdd_segment_store_bytes: device bytes of one nsteps segment's store
 *                    (15 rk_order planes per step).
 *   cdd_segment_forward: advance the plan's state by nsteps steps exactly as
 *                    cdd_step (CFL or fixed dt, dt record, loads, logs), and
 *                    keep the segment's start state, every stage state and
 *                    its stress planes in `store` (caller-owned device
 *                    memory).  The segment must not reach the end time.
 *   cdd_segment_vjp_stored: the reverse pass of cdd_segment_vjp over a
 *                    segment kept by cdd_segment_forward (same dts_dev,
 *                    src_dev, ubar_end, ubar0 and Lambda accumulator). */
int cdd_flux_vjp(cdd_plan* plan, const double* beta_dev, double t, const double* rbar_dev, double* betabar_dev,
                 double* lambar);
/* Branch-frozen flux (no reference counterpart; for checking the adjoint
 * against finite differences of a smooth map): the flux of cdd_flux (rhs_dev,
 * 9 planes) with every branch of the scheme -- the minmod selection of each
 * limiter site and, per axis, the corners of the largest and smallest
 * velocity and whether they pass zero (Max9/Min9 with 0,
 * CentralUpwindHJ3D.py:11-34,189-241) -- recorded into or replayed from
 * sel_dev (2 N^d words).  record = 1 decides them at beta (the flux is then
 * cdd_flux's to rounding) and writes them; record = 0 replays them, which
 * makes the flux a smooth function of beta and Lambda: the map the adjoint
 * differentiates.  float64 central-upwind glide+climb and glide-only plans. */
int cdd_flux_frozen(cdd_plan* plan, const double* beta_dev, double t, unsigned long long* sel_dev, int record,
                    double* rhs_dev);
/* Tangent-linear flux (no reference counterpart; forward-mode
 * differentiation of cdd_flux): rhs_dev = the flux at beta (cdd_flux's to
 * rounding) and drhs_dev = J_L(beta) dbeta + (dL/dLambda) dLambda, with the
 * branches decided at beta as the adjoint decides them.  The transpose of
 * this map is cdd_flux_vjp's.  float64 central-upwind glide+climb and
 * glide-only plans; all loads (the uniaxial control's tangent included). */
int cdd_flux_tangent(cdd_plan* plan, const double* beta_dev, const double* dbeta_dev, double t, double dLambda,
                     double* rhs_dev, double* drhs_dev);
int cdd_step_vjp(cdd_plan* plan, const double* u_dev, double t, double dt, const double* ubar_next_dev,
                 double* ubar_dev, double* lambar);
int cdd_energy_grad(cdd_plan* plan, const double* beta_dev, double t, double* betabar_dev, double* energy);
int cdd_dt_record(cdd_plan* plan, double* log_dev, long long cap);
int cdd_dt_replay(cdd_plan* plan, const double* sched_dev, long long len);
int cdd_segment_vjp(cdd_plan* plan, const double* u0_dev, double t0, const double* dts_dev, int nsteps,
                    const double* src_dev, const double* ubar_end_dev, double* ubar0_dev);
int cdd_lambda_bar(cdd_plan* plan, double* value, int reset);
int cdd_segment_store_bytes(cdd_plan* plan, int nsteps, unsigned long long* bytes);
int cdd_segment_forward(cdd_plan* plan, int nsteps, void* store_dev, unsigned long long store_bytes);
int cdd_segment_vjp_stored(cdd_plan* plan, const void* store_dev, unsigned long long store_bytes,
                           const double* dts_dev, int nsteps, const double* src_dev, const double* ubar_end_dev,
                           double* ubar0_dev);

/* ---- Batched independent realisations (2D) ------------------------------
 * The reference advances independent realisations one process per GPU
 * (Plasticity/exec_batch.py:26-58, RunCUDA.simulation per device/seed).
 * A batch advances `count` 2D realisations of possibly different sides
 * (sizes[i] in 16..1024 incl. 384, 768) together: one hipGraph-captured step
 * per launch advances every instance with its own CFL time step and end
 * time.  cfg gives the shared physics/stepping (ndim must be 2, no load). */
typedef struct cdd_batch cdd_batch;
int cdd_batch_create(const cdd_config* cfg, const int* sizes, int count, cdd_batch** out);
int cdd_batch_destroy(cdd_batch* batch);
int cdd_batch_set_stream(cdd_batch* batch, void* stream);
int cdd_batch_size(cdd_batch* batch, int i, int* n);
int cdd_batch_set_state(cdd_batch* batch, int i, const double* beta, int on_device);
int cdd_batch_get_state(cdd_batch* batch, int i, double* beta, int on_device);
/* i = -1 sets every instance */
int cdd_batch_set_time(cdd_batch* batch, int i, double t, double t_end, double fixed_dt);
int cdd_batch_get_clock(cdd_batch* batch, int i, double* t, double* last_dt, int* done, long long* nsteps);
/* Replays the captured step nsteps times (no host synchronisation). */
int cdd_batch_step(cdd_batch* batch, int nsteps);
/* The flux of every stage is one persistent multi-instance launch over all
 * instances of all sizes: tile = 0, 1, 2 (32x8, 64x4, 16x16 cells) walks the
 * (size, instance, tile) list with a fixed grid of per_cu (1 or 2) blocks per
 * CU; tile = 3 marches rows (sides that are multiples of 128): per_cu (1..4)
 * blocks per CU, each an equal share of the (size, instance, 128-column
 * strip, row) space.  per_cu = 0 launches one flux per size instead.  Sides
 * the chosen tile does not divide keep one launch per size.  Default: tile
 * 3, per_cu 2. */
int cdd_batch_set_flux(cdd_batch* batch, int tile, int per_cu);
/* Autotuner hooks: the per-size flux tile variant (sizes outside the
 * multi-instance launch, or per_cu = 0), the timing (ms per launch, on the
 * current states, no state change) of one per-size variant and of one
 * multi-instance configuration. */
int cdd_batch_set_tile(cdd_batch* batch, int n, int variant);
int cdd_batch_time_flux(cdd_batch* batch, int n, int variant, int reps, double* ms);
int cdd_batch_time_flux_multi(cdd_batch* batch, int tile, int per_cu, int reps, double* ms);
/* The stress solve's R2C / C2R row tile per size (0: the default rows per
 * block, 1: half of them; sides >= 256) and the timing of one stress solve
 * of a size group with a variant. */
int cdd_batch_set_spectral(cdd_batch* batch, int n, int variant);
int cdd_batch_time_spectral(cdd_batch* batch, int n, int variant, int reps, double* ms);

const char* cdd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
