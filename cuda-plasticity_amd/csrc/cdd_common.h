// Shared device types and helpers for the CDD stepper (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cdd {

// Machine epsilon of float64: the ME of Plasticity/NumericalMethods.py:4.
constexpr double kME = 2.220446049250313e-16;

template <typename T>
struct cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> operator*(cx<T> a, T s) { return {a.x * s, a.y * s}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cx<T> conj(cx<T> a) { return {a.x, -a.y}; }
// Loads and stores with a cache policy chosen at compile time: NT = the
// non-temporal (streaming) policy, for data a pass touches once and that
// would otherwise displace lines other accesses re-read.  Complex values go
// as one 2-T vector (2-T-aligned addresses).
template <bool NT, typename T>
__device__ __forceinline__ T ld_pol(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void st_pol(T* p, T v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}
template <bool NT, typename T>
__device__ __forceinline__ cx<T> ld_pol(const cx<T>* p) {
  if constexpr (NT) {
    typedef T V2 __attribute__((ext_vector_type(2)));
    const V2 v = __builtin_nontemporal_load(reinterpret_cast<const V2*>(p));
    return {v.x, v.y};
  } else {
    return *p;
  }
}
template <bool NT, typename T>
__device__ __forceinline__ void st_pol(cx<T>* p, cx<T> a) {
  if constexpr (NT) {
    typedef T V2 __attribute__((ext_vector_type(2)));
    V2 v;
    v.x = a.x;
    v.y = a.y;
    __builtin_nontemporal_store(v, reinterpret_cast<V2*>(p));
  } else {
    *p = a;
  }
}
// -i * a
template <typename T>
__device__ __forceinline__ cx<T> mul_ni(cx<T> a) { return {a.y, -a.x}; }
// +i * a
template <typename T>
__device__ __forceinline__ cx<T> mul_pi(cx<T> a) { return {-a.y, a.x}; }

// 1/sqrt(x) for x > 0: the hardware reciprocal square root refined by one
// Newton step y <- y (3/2 - x y^2 / 2), accurate to a few ulp.  The
// forward flux and the adjoint's recomputation share it, so both linearise
// at the same point.
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  return y * (1.5 - 0.5 * x * y * y);
}
__device__ __forceinline__ float rsqrt_nr(float x) {
  float y = __builtin_amdgcn_rsqf(x);
  return y * (1.5f - 0.5f * x * y * y);
}

// Run-time clock of one stepper instance, resident in device memory so the
// CFL time step never round-trips through the host.
struct Clock {
  double t;         // time at the start of the current step
  double t_end;     // integration end time (Run's endTime)
  double dt;        // time step of the current step
  double fixed_dt;  // > 0: fixed step (CalculateOneFixedTimeStep)
  double cfl;       // CFLsafeFactor
  double last_max;  // max speed seen by the last CFL reduction
  int done;         // 1 once t >= t_end (or on a failure): every step kernel returns at once
  int status;       // 0 ok; 1 a non-finite speed stopped the run (the state is the last step's start)
  long long nsteps;  // completed steps
  // frozen time-step schedule (adjoint replay): dt = dt_sched[nsteps - sched_base]
  const double* dt_sched;
  long long sched_base, sched_len;
  // optional log of the time steps taken: dt_log[nsteps - log_base] = dt
  double* dt_log;
  long long log_base, log_cap;
  // optional stress-strain log: 8 doubles per step (t, mean sigma xx yy zz
  // xy xz yz with the load, mean beta_zz) at slog[8 (nsteps - slog_base)]
  double* slog;
  long long slog_base, slog_cap;
  // optional energy log: total elastic energy of each step's start state
  double* elog;
  long long elog_base, elog_cap;
  // optional speed log: the step's CFL speed (last_max) at vlog[nsteps - vlog_base]
  double* vlog;
  long long vlog_base, vlog_cap;
  // 0: dt = CFL / (N max speed), the end step lands on t_end
  //    (FieldMover.py:94-111); 1: dt = CFL / max speed / N, the end step is
  //    t_end - t + ME (plasticity_kernel.cu simpleTVD :1121-1128)
  int dt_rule;
};

// Uniform load added to sigma (it enters the stress solve as the k = 0 mode).
struct LoadParams {
  int kind;  // 0 none, 1 strain, 2 stress, 3 uniaxial strain-rate, 4 full tensor
  int pad;
  double init[9];  // strain/stress: diag in [0..2]; tensor: 9 components
  double rate[9];
  double t0;
  double init_strain;    // uniaxial: initial <beta_zz>
  double uniaxial_rate;  // uniaxial: strain rate
  double smu, slam;      // strain load: its own mu and lam (RunPython's module constants)
};

struct Elastic {
  double mu, nu, lam;  // lam = 2 mu nu / (1 - 2 nu)
};

// Uniform stress shift of the load at time t; mean_bzz = <beta_zz> of the
// stage state (used by the uniaxial strain-rate control only).
__device__ __forceinline__ void load_shift(const LoadParams& L, const Elastic& E, double t,
                                           double mean_bzz, double s[9]) {
#pragma unroll
  for (int c = 0; c < 9; ++c) s[c] = 0.0;
  if (L.kind == 1) {  // RunPython.py:21-26, 44-49
    double e[3];
    for (int i = 0; i < 3; ++i) e[i] = L.init[i] + L.rate[i] * (t - L.t0);
    double tr = e[0] + e[1] + e[2];
    for (int i = 0; i < 3; ++i) s[4 * i] = L.slam * tr + 2.0 * L.smu * e[i];
  } else if (L.kind == 2) {  // RunPython.py:28-32
    for (int i = 0; i < 3; ++i) s[4 * i] = L.init[i] + L.rate[i] * (t - L.t0);
  } else if (L.kind == 3) {  // MattTools/Run/RunLoadingMPI.py:68-83
    double bzz = L.uniaxial_rate * t - (mean_bzz - L.init_strain);
    for (int i = 0; i < 3; ++i) s[4 * i] = E.lam * bzz;
    s[8] += 2.0 * E.mu * bzz;
  } else if (L.kind == 4) {
    for (int c = 0; c < 9; ++c) s[c] = L.init[c] + L.rate[c] * (t - L.t0);
  }
}

}  // namespace cdd
