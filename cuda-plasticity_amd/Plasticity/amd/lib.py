"""ctypes binding of libcdd.so (include/cdd.h).

The library is built in-tree by ``cuda-plasticity_amd/build_lib.py``; it is
required -- importing this module without it raises, there is no fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libcdd.so")

GLIDE_CLIMB, GLIDE_ONLY, GLIDE_CLIMB_WOJZZ = 0, 1, 2
FP64, FP32 = 0, 1
LOAD_NONE, LOAD_STRAIN, LOAD_STRESS, LOAD_UNIAXIAL, LOAD_TENSOR = 0, 1, 2, 3, 4
# flux discretisations: the Python stepper's central-upwind scheme, and the
# CUDA executable's local Lax-Friedrichs scheme (rk_order 1 = its simpleTVD)
SCHEME_CENTRAL_UPWIND, SCHEME_LLF = 0, 1
SCHEMES = {"central-upwind": SCHEME_CENTRAL_UPWIND, "llf": SCHEME_LLF}


class CDDError(RuntimeError):
    pass


class Config(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int),
        ("n", ctypes.c_int),
        ("dynamics", ctypes.c_int),
        ("rk_order", ctypes.c_int),
        ("Lambda", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("nu", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("load_kind", ctypes.c_int),
        ("device", ctypes.c_int),
        ("load_init", ctypes.c_double * 9),
        ("load_rate", ctypes.c_double * 9),
        ("load_t0", ctypes.c_double),
        ("load_init_strain", ctypes.c_double),
        ("load_uniaxial_rate", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("scheme", ctypes.c_int),
        ("load_mu", ctypes.c_double),
        ("load_nu", ctypes.c_double),
    ]


EXPORTS = [
    "cdd_plan_create", "cdd_plan_destroy", "cdd_set_stream", "cdd_set_state", "cdd_get_state",
    "cdd_set_time", "cdd_get_clock", "cdd_sigma", "cdd_rho", "cdd_flux", "cdd_step", "cdd_energy",
    "cdd_sigma_variant", "cdd_rho_variant", "cdd_rho_symmetric",
    "cdd_state_ptr", "cdd_last_error", "cdd_timing_enable", "cdd_timing_read",
    "cdd_batch_create", "cdd_batch_destroy", "cdd_batch_set_stream", "cdd_batch_size", "cdd_batch_set_state",
    "cdd_batch_get_state", "cdd_batch_set_time", "cdd_batch_get_clock", "cdd_batch_step", "cdd_batch_set_tile",
    "cdd_batch_time_flux", "cdd_batch_set_flux", "cdd_batch_time_flux_multi", "cdd_batch_set_spectral",
    "cdd_batch_time_spectral",
    "cdd_flux_vjp", "cdd_step_vjp", "cdd_energy_grad", "cdd_dt_record", "cdd_dt_replay",
    "cdd_segment_vjp", "cdd_lambda_bar", "cdd_stress_record", "cdd_plas_append", "cdd_plas_resume",
    "cdd_energy_record", "cdd_speed_record", "cdd_status", "cdd_instability", "cdd_segment_store_bytes", "cdd_segment_forward",
    "cdd_segment_vjp_stored", "cdd_set_vacancy", "cdd_vacancy_set", "cdd_vacancy_get",
    "cdd_set_disorder", "cdd_disorder_get", "cdd_disorder_sigma", "cdd_disorder_flux", "cdd_sigma_noalloc",
    "cdd_flux_frozen", "cdd_flux_tangent", "cdd_set_source",
    "cdd_vacancy_alpha",
]
KERNEL_KINDS = ["r2c", "cols_fwd", "contract", "cols_inv", "c2r", "flux_stage1", "cfl_reduce",
                "flux_stage2", "flux_stage3", "advance"]


def _hip_runtimes():
    """Paths of every HIP runtime library mapped into this process."""
    paths = set()
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                p = line.split()[-1]
                if "libamdhip64" in os.path.basename(p):
                    paths.add(os.path.realpath(p))
    except OSError:
        pass
    return paths


def _load():
    if not os.path.exists(LIBPATH):
        raise CDDError("libcdd.so is not built (%s); run __graft_entry__.build()" % LIBPATH)
    # torch ships its own HIP runtime under the same soname (libamdhip64.so.7).
    # Loading torch first lets libcdd.so bind to that copy; loading libcdd.so
    # first would map a second runtime from /opt/rocm, and the two then fail
    # to share the device ("no ROCm-capable device" in whichever starts last).
    import torch  # noqa: F401
    L = ctypes.CDLL(LIBPATH)
    rts = _hip_runtimes()
    if len(rts) > 1:
        raise CDDError("two HIP runtimes are mapped into this process (%s); import torch before "
                       "loading libcdd.so" % ", ".join(sorted(rts)))
    P = ctypes.c_void_p
    D = ctypes.c_double
    I = ctypes.c_int
    dp = ctypes.POINTER(ctypes.c_double)
    sig = {
        "cdd_plan_create": ([ctypes.POINTER(Config), ctypes.POINTER(P)], I),
        "cdd_plan_destroy": ([P], I),
        "cdd_set_stream": ([P, P], I),
        "cdd_set_state": ([P, P, I], I),
        "cdd_get_state": ([P, P, I], I),
        "cdd_set_time": ([P, D, D, D], I),
        "cdd_get_clock": ([P, dp, dp, ctypes.POINTER(I), ctypes.POINTER(ctypes.c_longlong)], I),
        "cdd_sigma": ([P, P, D, P], I),
        "cdd_rho": ([P, P, P], I),
        "cdd_sigma_variant": ([P, P, I, I, P], I),
        "cdd_rho_variant": ([P, P, I, P], I),
        "cdd_rho_symmetric": ([P, P, P], I),
        "cdd_sigma_noalloc": ([ctypes.POINTER(Config), P, D, P, P], I),
        "cdd_flux": ([P, P, D, P, P], I),
        "cdd_step": ([P, I], I),
        "cdd_energy": ([P, P, D, dp], I),
        "cdd_state_ptr": ([P, ctypes.POINTER(P)], I),
        "cdd_last_error": ([], ctypes.c_char_p),
        "cdd_timing_enable": ([P, I], I),
        "cdd_flux_vjp": ([P, P, D, P, P, dp], I),
        "cdd_step_vjp": ([P, P, D, D, P, P, dp], I),
        "cdd_flux_frozen": ([P, P, D, P, I, P], I),
        "cdd_flux_tangent": ([P, P, P, D, D, P, P], I),
        "cdd_energy_grad": ([P, P, D, P, dp], I),
        "cdd_dt_record": ([P, P, ctypes.c_longlong], I),
        "cdd_dt_replay": ([P, P, ctypes.c_longlong], I),
        "cdd_segment_vjp": ([P, P, D, P, I, P, P, P], I),
        "cdd_segment_store_bytes": ([P, I, ctypes.POINTER(ctypes.c_ulonglong)], I),
        "cdd_segment_forward": ([P, I, P, ctypes.c_ulonglong], I),
        "cdd_segment_vjp_stored": ([P, P, ctypes.c_ulonglong, P, I, P, P, P], I),
        "cdd_lambda_bar": ([P, dp, I], I),
        "cdd_set_vacancy": ([P, I, D, D, D, D], I),
        "cdd_vacancy_set": ([P, P, I], I),
        "cdd_vacancy_alpha": ([P, P, I], I),
        "cdd_vacancy_get": ([P, P, I], I),
        "cdd_set_disorder": ([P, P, I, I], I),
        "cdd_set_source": ([P, P, I, D, D], I),
        "cdd_disorder_get": ([P, P, I], I),
        "cdd_disorder_sigma": ([P, P, P, D, P], I),
        "cdd_disorder_flux": ([P, P, P, D, P, P, P], I),
        "cdd_stress_record": ([P, P, ctypes.c_longlong], I),
        "cdd_energy_record": ([P, P, ctypes.c_longlong], I),
        "cdd_speed_record": ([P, P, ctypes.c_longlong], I),
        "cdd_status": ([P, ctypes.POINTER(I)], I),
        "cdd_instability": ([P, P, dp], I),
        "cdd_plas_append": ([P, ctypes.c_char_p], I),
        "cdd_plas_resume": ([P, ctypes.c_char_p, dp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(I)], I),
        "cdd_batch_create": ([ctypes.POINTER(Config), ctypes.POINTER(I), I, ctypes.POINTER(P)], I),
        "cdd_batch_destroy": ([P], I),
        "cdd_batch_set_stream": ([P, P], I),
        "cdd_batch_size": ([P, I, ctypes.POINTER(I)], I),
        "cdd_batch_set_state": ([P, I, P, I], I),
        "cdd_batch_get_state": ([P, I, P, I], I),
        "cdd_batch_set_time": ([P, I, D, D, D], I),
        "cdd_batch_get_clock": ([P, I, dp, dp, ctypes.POINTER(I), ctypes.POINTER(ctypes.c_longlong)], I),
        "cdd_batch_step": ([P, I], I),
        "cdd_batch_set_tile": ([P, I, I], I),
        "cdd_batch_time_flux": ([P, I, I, I, dp], I),
        "cdd_batch_set_flux": ([P, I, I], I),
        "cdd_batch_time_flux_multi": ([P, I, I, I, dp], I),
        "cdd_batch_set_spectral": ([P, I, I], I),
        "cdd_batch_time_spectral": ([P, I, I, I, dp], I),
        "cdd_timing_read": ([P, dp, ctypes.POINTER(I), I], I),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    return L


lib = _load()


def check(rc):
    if rc != 0:
        raise CDDError(lib.cdd_last_error().decode(errors="replace"))


def _ptr(t):
    """Device pointer of a torch tensor (contiguous float64, or float32 for the
    fields of an fp32 plan, on the GPU)."""
    if not t.is_cuda:
        raise CDDError("expected a device tensor")
    if t.dtype.__repr__() not in ("torch.float64", "torch.float32"):
        raise CDDError("expected float64 or float32")
    if not t.is_contiguous():
        raise CDDError("expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())


class Plan:
    """One stepper plan: owns the device state, scratch and clock."""

    def __init__(self, ndim, n, dynamics=GLIDE_CLIMB, rk_order=3, Lambda=0.0, mu=1.0, nu=0.3, cfl=0.25,
                 load_kind=LOAD_NONE, load_init=None, load_rate=None, load_t0=0.0, load_init_strain=0.0,
                 load_uniaxial_rate=0.0, device=0, precision="fp64", scheme="central-upwind", load_mu=0.0,
                 load_nu=0.0):
        import torch
        self.torch = torch
        if precision not in ("fp64", "fp32"):
            raise CDDError("precision must be 'fp64' or 'fp32'")
        self.precision = precision
        self.dtype = torch.float32 if precision == "fp32" else torch.float64
        cfg = Config()
        cfg.precision = FP32 if precision == "fp32" else FP64
        if scheme not in SCHEMES:
            raise CDDError("scheme must be one of %s" % sorted(SCHEMES))
        cfg.scheme = SCHEMES[scheme]
        self.scheme = scheme
        cfg.ndim, cfg.n, cfg.dynamics, cfg.rk_order = ndim, n, dynamics, rk_order
        cfg.Lambda, cfg.mu, cfg.nu, cfg.cfl = Lambda, mu, nu, cfl
        cfg.load_kind, cfg.device = load_kind, device
        li = np.zeros(9) if load_init is None else np.asarray(load_init, float)
        lr = np.zeros(9) if load_rate is None else np.asarray(load_rate, float)
        if li.size == 3:
            li = np.concatenate([li, np.zeros(6)])
        if lr.size == 3:
            lr = np.concatenate([lr, np.zeros(6)])
        for c in range(9):
            cfg.load_init[c] = li[c]
            cfg.load_rate[c] = lr[c]
        cfg.load_t0 = load_t0
        cfg.load_init_strain = load_init_strain
        cfg.load_uniaxial_rate = load_uniaxial_rate
        # the strain load's own elastic constants (0: the plan's mu, nu)
        cfg.load_mu, cfg.load_nu = float(load_mu), float(load_nu)
        self.cfg = cfg
        self.ndim, self.n, self.device = ndim, n, device
        self.shape = (n,) * ndim
        h = ctypes.c_void_p()
        check(lib.cdd_plan_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.dev = torch.device("cuda", device)
        self.bind_stream()

    def bind_stream(self):
        """Run on torch's current stream so ordering with torch ops is implicit."""
        s = self.torch.cuda.current_stream(self.dev).cuda_stream
        check(lib.cdd_set_stream(self.h, ctypes.c_void_p(s)))

    def close(self):
        if getattr(self, "h", None):
            lib.cdd_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def empty(self, ncomp=9):
        return self.torch.empty((ncomp,) + self.shape, dtype=self.dtype, device=self.dev)

    def set_state(self, beta):
        t = self._as_dev(beta)
        check(lib.cdd_set_state(self.h, _ptr(t), 1))

    def get_state(self):
        out = self.empty()
        check(lib.cdd_get_state(self.h, _ptr(out), 1))
        return out

    def set_time(self, t, t_end=float("inf"), fixed_dt=0.0):
        check(lib.cdd_set_time(self.h, t, t_end, fixed_dt))

    def clock(self):
        t, dt = ctypes.c_double(), ctypes.c_double()
        done, ns = ctypes.c_int(), ctypes.c_longlong()
        check(lib.cdd_get_clock(self.h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(done), ctypes.byref(ns)))
        return t.value, dt.value, bool(done.value), ns.value

    def step(self, nsteps=1):
        check(lib.cdd_step(self.h, int(nsteps)))

    def sigma(self, beta, t=0.0):
        b = self._as_dev(beta)
        out = self.empty()
        check(lib.cdd_sigma(self.h, _ptr(b), t, _ptr(out)))
        return out

    def sigma_noalloc(self, beta, out, t=0.0, stream=None):
        """cdd_sigma_noalloc with this plan's configuration: the stress of beta
        into `out` (both (9,) + grid float64 device tensors) on `stream`
        (default: torch's current stream), using the library's static
        workspace instead of the plan's buffers."""
        b, o = self._as_dev(beta), out
        if o.dtype != self.torch.float64 or not o.is_contiguous() or tuple(o.shape) != (9,) + self.shape:
            raise CDDError("out must be a contiguous float64 (9,) + grid tensor")
        s = self.torch.cuda.current_stream(self.dev).cuda_stream if stream is None else stream
        check(lib.cdd_sigma_noalloc(ctypes.byref(self.cfg), _ptr(b), float(t), _ptr(o), ctypes.c_void_p(s)))
        return o

    def rho(self, beta):
        b = self._as_dev(beta)
        out = self.empty()
        check(lib.cdd_rho(self.h, _ptr(b), _ptr(out)))
        return out

    _KTYPES = {"k": 0, "sinK": 1}
    _SOURCES = {"betaE": 0, "betaP": 1, "rho": 2}

    def sigma_variant(self, beta, type="k", source="betaE"):
        """PlasticityState.CalculateSigma(type, source) of beta (no external
        load): type 'k' or 'sinK', source 'betaE', 'betaP' or 'rho'."""
        if type not in self._KTYPES or source not in self._SOURCES:
            raise CDDError("type must be 'k' or 'sinK', source 'betaE', 'betaP' or 'rho'")
        b = self._as_dev(beta)
        out = self.empty()
        check(lib.cdd_sigma_variant(self.h, _ptr(b), self._KTYPES[type], self._SOURCES[source], _ptr(out)))
        return out

    def rho_variant(self, beta, type="k"):
        """PlasticityState.CalculateRhoFourier(type) of beta."""
        if type not in self._KTYPES:
            raise CDDError("type must be 'k' or 'sinK'")
        b = self._as_dev(beta)
        out = self.empty()
        check(lib.cdd_rho_variant(self.h, _ptr(b), self._KTYPES[type], _ptr(out)))
        return out

    def rho_symmetric(self, beta):
        """PlasticityState.CalculateRhoSymmetric of beta (central differences)."""
        b = self._as_dev(beta)
        out = self.empty()
        check(lib.cdd_rho_symmetric(self.h, _ptr(b), _ptr(out)))
        return out

    def flux(self, beta, t=0.0):
        b = self._as_dev(beta)
        rhs = self.empty()
        speed = self.torch.empty(self.shape, dtype=self.dtype, device=self.dev)
        check(lib.cdd_flux(self.h, _ptr(b), t, _ptr(rhs), _ptr(speed)))
        return rhs, speed

    def energy(self, beta=None, t=0.0):
        e = ctypes.c_double()
        p = _ptr(self._as_dev(beta)) if beta is not None else None
        check(lib.cdd_energy(self.h, p, t, ctypes.byref(e)))
        return e.value

    # ---- adjoint (frozen time-step schedule)
    def flux_vjp(self, beta, t, rbar):
        b, r = self._as_dev(beta), self._as_dev(rbar)
        out = self.empty()
        lam = ctypes.c_double()
        check(lib.cdd_flux_vjp(self.h, _ptr(b), t, _ptr(r), _ptr(out), ctypes.byref(lam)))
        return out, lam.value

    def flux_frozen(self, beta, t, sel, record, out=None):
        """cdd_flux_frozen: the flux with its minmod and speed branches
        recorded into (record=True) or replayed from (False) `sel`, an int64
        device tensor of 2 * cells words."""
        b = self._as_dev(beta)
        if (sel.dtype != self.torch.int64 or not sel.is_contiguous() or sel.device != self.dev
                or sel.numel() != 2 * b[0].numel()):
            raise CDDError("sel must be a contiguous int64 tensor of 2 * cells words on the plan's device")
        if out is None:
            out = self.empty()
        check(lib.cdd_flux_frozen(self.h, _ptr(b), float(t), ctypes.c_void_p(sel.data_ptr()), 1 if record else 0,
                                  _ptr(out)))
        return out

    def flux_tangent(self, beta, dbeta, t=0.0, dLambda=0.0):
        """cdd_flux_tangent: (rhs, drhs), the flux at beta and its tangent
        along (dbeta, dLambda)."""
        b, db = self._as_dev(beta), self._as_dev(dbeta)
        rhs, drhs = self.empty(), self.empty()
        check(lib.cdd_flux_tangent(self.h, _ptr(b), _ptr(db), float(t), float(dLambda), _ptr(rhs), _ptr(drhs)))
        return rhs, drhs

    def step_vjp(self, u, t, dt, ubar_next):
        a, g = self._as_dev(u), self._as_dev(ubar_next)
        out = self.empty()
        lam = ctypes.c_double()
        check(lib.cdd_step_vjp(self.h, _ptr(a), t, dt, _ptr(g), _ptr(out), ctypes.byref(lam)))
        return out, lam.value

    def energy_grad(self, beta, t=0.0):
        b = self._as_dev(beta)
        out = self.empty()
        e = ctypes.c_double()
        check(lib.cdd_energy_grad(self.h, _ptr(b), t, _ptr(out), ctypes.byref(e)))
        return out, e.value

    def segment_vjp(self, u0, t0, dts, ubar_end, out=None, src=None):
        """Reverse pass over one checkpoint segment: recompute dts.numel()
        steps from the checkpoint u0 at time t0 replaying the recorded time
        steps dts (float64 device tensor), then return the cotangent of u0
        given the cotangent ubar_end of the segment's final state.  `out` may
        be ubar_end itself.  The Lambda cotangent accumulates on the device
        (lambda_bar).  src: optional (dts.numel(), 9) float64 device tensor,
        the uniform gradient of a time-distributed objective at the state
        that starts each step."""
        a, g = self._as_dev(u0), self._as_dev(ubar_end)
        if dts.dtype != self.torch.float64 or not dts.is_contiguous() or dts.device != self.dev:
            raise CDDError("dts must be a contiguous float64 tensor on the plan's device")
        if out is None:
            out = self.empty()
        elif (out.dtype != self.torch.float64 or not out.is_contiguous() or out.device != self.dev
              or tuple(out.shape) != (9,) + self.shape):
            raise CDDError("out must be a contiguous float64 (9,) + grid tensor on the plan's device")
        sp = None
        if src is not None:
            if (src.dtype != self.torch.float64 or not src.is_contiguous() or src.device != self.dev
                    or tuple(src.shape) != (dts.numel(), 9)):
                raise CDDError("src must be a contiguous float64 (nsteps, 9) tensor on the plan's device")
            sp = _ptr(src)
        check(lib.cdd_segment_vjp(self.h, _ptr(a), float(t0), _ptr(dts), int(dts.numel()), sp, _ptr(g), _ptr(out)))
        return out

    def set_vacancy(self, alpha=1.0, beta=1.0, gamma=1.0, dt_bound=None, enable=True):
        """Vacancy coupling (BetaP_VacancyDynamics with the operator-splitting
        mover): stress solves see sigma_ii - alpha c, and after every step
        c += beta (tr beta_new - tr beta_old) and c diffuses by
        exp(-gamma dt k^2); dt_bound caps the adaptive step (None: no cap).
        c starts at zero (vacancy_set)."""
        check(lib.cdd_set_vacancy(self.h, 1 if enable else 0, float(alpha), float(beta), float(gamma),
                                  float(dt_bound) if dt_bound else 0.0))

    def set_source(self, src, rate=1.0, dt_cap=0.0):
        """Dynamic nucleation: every flux evaluation gains rate * src (nine
        components, the grid shape) and the adaptive time step is capped at
        dt_cap (0: no cap).  None switches it off."""
        if src is None:
            check(lib.cdd_set_source(self.h, None, 0, 0.0, 0.0))
            return
        t = self._as_dev(src)
        check(lib.cdd_set_source(self.h, _ptr(t), 1, float(rate), float(dt_cap)))
        self.torch.cuda.current_stream(self.dev).synchronize()

    def vacancy_set(self, c):
        """Set the vacancy concentration field (gridShape, float64)."""
        t = self.torch.as_tensor(c, dtype=self.torch.float64)
        if tuple(t.shape) != self.shape:
            raise CDDError("c must have the grid shape %s" % (self.shape,))
        t = t.to(self.dev).contiguous()
        check(lib.cdd_vacancy_set(self.h, ctypes.c_void_p(t.data_ptr()), 1))

    def vacancy_alpha(self, alpha):
        """alpha as a field (the grid shape) in the vacancies' stress
        -alpha c; None returns to the scalar of set_vacancy."""
        if alpha is None:
            check(lib.cdd_vacancy_alpha(self.h, None, 0))
            return
        t = self.torch.as_tensor(alpha, dtype=self.torch.float64)
        if tuple(t.shape) != self.shape:
            raise CDDError("alpha must have the grid shape %s" % (self.shape,))
        t = t.to(self.dev).contiguous()
        check(lib.cdd_vacancy_alpha(self.h, ctypes.c_void_p(t.data_ptr()), 1))
        self.torch.cuda.current_stream(self.dev).synchronize()

    def vacancy_get(self):
        """The vacancy concentration field as a device tensor."""
        out = self.torch.empty(self.shape, dtype=self.torch.float64, device=self.dev)
        check(lib.cdd_vacancy_get(self.h, ctypes.c_void_p(out.data_ptr()), 1))
        return out

    def _scalar_field(self, d):
        t = self.torch.as_tensor(d, dtype=self.torch.float64)
        if tuple(t.shape) != self.shape:
            raise CDDError("the disorder field must have the grid shape %s" % (self.shape,))
        return t.to(self.dev).contiguous()

    def set_disorder(self, d, advect=True):
        """Disorder coupling: every stress solve gains the disorder term of d
        (DisorderState.CalculateSigma); advect=True carries d with the flux
        and the RK stages as the ('s', 's') component of a DisorderState,
        advect=False keeps it fixed (BetaP_DisorderDynamics).  None disables."""
        if d is None:
            check(lib.cdd_set_disorder(self.h, None, 0, 0))
            return
        t = self._scalar_field(d)
        check(lib.cdd_set_disorder(self.h, _ptr(t), 1, 1 if advect else 0))

    def disorder_get(self):
        """The (advected) disorder field as a device tensor."""
        out = self.torch.empty(self.shape, dtype=self.torch.float64, device=self.dev)
        check(lib.cdd_disorder_get(self.h, _ptr(out), 1))
        return out

    def disorder_sigma(self, beta, d, t=0.0):
        """The stress of a DisorderState (nine planes, not symmetric)."""
        b, dd = self._as_dev(beta), self._scalar_field(d)
        out = self.empty()
        check(lib.cdd_disorder_sigma(self.h, _ptr(b), _ptr(dd), float(t), _ptr(out)))
        return out

    def disorder_flux(self, beta, d, t=0.0, advect=True):
        """(rhs of beta, rhs of d or None, speed) of the disorder dynamics."""
        b, dd = self._as_dev(beta), self._scalar_field(d)
        rhs = self.empty()
        rd = self.torch.empty(self.shape, dtype=self.torch.float64, device=self.dev) if advect else None
        speed = self.torch.empty(self.shape, dtype=self.torch.float64, device=self.dev)
        check(lib.cdd_disorder_flux(self.h, _ptr(b), _ptr(dd), float(t), _ptr(rhs), _ptr(rd) if advect else None,
                                    _ptr(speed)))
        return rhs, rd, speed

    def segment_store_bytes(self, nsteps):
        """Device bytes of one nsteps segment's stored trajectory."""
        v = ctypes.c_ulonglong()
        check(lib.cdd_segment_store_bytes(self.h, int(nsteps), ctypes.byref(v)))
        return v.value

    def _store_ptr(self, store, nsteps):
        t = self.torch
        if not store.is_cuda or store.device != self.dev or not store.is_contiguous():
            raise CDDError("store must be a contiguous tensor on the plan's device")
        nbytes = store.numel() * store.element_size()
        if nbytes < self.segment_store_bytes(nsteps):
            raise CDDError("store too small for %d steps (segment_store_bytes)" % nsteps)
        return ctypes.c_void_p(store.data_ptr()), nbytes

    def segment_forward(self, nsteps, store):
        """Advance the state by nsteps steps like step(), keeping the
        segment's start state, every stage state and its stress planes in
        `store` (a device tensor of at least segment_store_bytes(nsteps)
        bytes) for segment_vjp_stored."""
        ptr, nbytes = self._store_ptr(store, nsteps)
        check(lib.cdd_segment_forward(self.h, int(nsteps), ptr, nbytes))

    def segment_vjp_stored(self, store, dts, ubar_end, out=None, src=None):
        """segment_vjp over a segment kept by segment_forward: no
        recomputation (same dts, ubar_end, out, src and Lambda cotangent)."""
        g = self._as_dev(ubar_end)
        if dts.dtype != self.torch.float64 or not dts.is_contiguous() or dts.device != self.dev:
            raise CDDError("dts must be a contiguous float64 tensor on the plan's device")
        nsteps = int(dts.numel())
        ptr, nbytes = self._store_ptr(store, nsteps)
        if out is None:
            out = self.empty()
        elif (out.dtype != self.torch.float64 or not out.is_contiguous() or out.device != self.dev
              or tuple(out.shape) != (9,) + self.shape):
            raise CDDError("out must be a contiguous float64 (9,) + grid tensor on the plan's device")
        sp = None
        if src is not None:
            if (src.dtype != self.torch.float64 or not src.is_contiguous() or src.device != self.dev
                    or tuple(src.shape) != (nsteps, 9)):
                raise CDDError("src must be a contiguous float64 (nsteps, 9) tensor on the plan's device")
            sp = _ptr(src)
        check(lib.cdd_segment_vjp_stored(self.h, ptr, nbytes, _ptr(dts), nsteps, sp, _ptr(g), _ptr(out)))
        return out

    def lambda_bar(self, reset=True):
        """Lambda cotangent accumulated by segment_vjp since the last reset."""
        v = ctypes.c_double()
        check(lib.cdd_lambda_bar(self.h, ctypes.byref(v), 1 if reset else 0))
        return v.value

    def _f64_log(self, t, what, shape=None):
        """Device pointer of a float64 log/schedule tensor the library reads or
        writes as double[numel]; anything else is refused before a launch."""
        if t.dtype != self.torch.float64:
            raise CDDError("%s must be float64, got %s" % (what, t.dtype))
        if not t.is_contiguous():
            raise CDDError("%s must be contiguous" % what)
        if shape is not None and (t.dim() != len(shape) or any(a is not None and a != b
                                                             for a, b in zip(shape, t.shape))):
            raise CDDError("%s must have shape %s, got %s" % (what, shape, tuple(t.shape)))
        if not t.is_cuda or t.device != self.dev:
            raise CDDError("%s must be on the plan's device %s, got %s" % (what, self.dev, t.device))
        return ctypes.c_void_p(t.data_ptr())

    def dt_record(self, log):
        """Log the next log.numel() time steps into the float64 device tensor `log`
        (None stops logging)."""
        if log is None:
            check(lib.cdd_dt_record(self.h, None, 0))
        else:
            check(lib.cdd_dt_record(self.h, self._f64_log(log, "dt log"), log.numel()))

    def plas_append(self, path):
        """Append [t, state] to a .plas file in the reference's layout."""
        check(lib.cdd_plas_append(self.h, os.fsencode(path)))

    def plas_resume(self, path):
        """Load the last complete record of a .plas file into the state and
        clock; returns (t, complete records, torn tail)."""
        t, n, torn = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_int()
        check(lib.cdd_plas_resume(self.h, os.fsencode(path), ctypes.byref(t), ctypes.byref(n), ctypes.byref(torn)))
        return t.value, n.value, bool(torn.value)

    def energy_record(self, log):
        """Append the total elastic energy of every step's start state to the
        float64 device tensor `log` (None stops)."""
        if log is None:
            check(lib.cdd_energy_record(self.h, None, 0))
        else:
            check(lib.cdd_energy_record(self.h, self._f64_log(log, "energy log"), log.numel()))

    def speed_record(self, log):
        """Append every step's CFL speed (the largest one-sided corner speed
        of its first stage) to the float64 device tensor `log` (None stops);
        logged under fixed time steps too."""
        if log is None:
            check(lib.cdd_speed_record(self.h, None, 0))
        else:
            check(lib.cdd_speed_record(self.h, self._f64_log(log, "speed log"), log.numel()))

    def status(self):
        """0, or 1 if a non-finite speed stopped the run on the device."""
        v = ctypes.c_int()
        check(lib.cdd_status(self.h, ctypes.byref(v)))
        return v.value

    def instability(self, beta=None):
        """|FFT(|rho|)| at the (N/2, ..., N/2) mode of beta (default: the state)."""
        v = ctypes.c_double()
        b = _ptr(self._as_dev(beta)) if beta is not None else None
        check(lib.cdd_instability(self.h, b, ctypes.byref(v)))
        return v.value

    def stress_record(self, log):
        """Append one stress-strain record per step to the float64 device tensor
        `log` of shape (cap, 8): t, mean sigma (xx, yy, zz, xy, xz, yz, load
        included) at the step start, mean beta_zz.  None stops recording."""
        if log is None:
            check(lib.cdd_stress_record(self.h, None, 0))
        else:
            ptr = self._f64_log(log, "stress log", shape=(None, 8))
            check(lib.cdd_stress_record(self.h, ptr, log.shape[0]))

    def dt_replay(self, sched):
        """Take the next sched.numel() time steps from the float64 device tensor
        `sched` (None returns to the CFL rule)."""
        if sched is None:
            check(lib.cdd_dt_replay(self.h, None, 0))
        else:
            check(lib.cdd_dt_replay(self.h, self._f64_log(sched, "dt schedule"), sched.numel()))

    def timing_enable(self, max_events):
        check(lib.cdd_timing_enable(self.h, int(max_events)))

    def timing_read(self):
        k = len(KERNEL_KINDS)
        ms = (ctypes.c_double * k)()
        cnt = (ctypes.c_int * k)()
        check(lib.cdd_timing_read(self.h, ms, cnt, k))
        return {KERNEL_KINDS[i]: (ms[i], cnt[i]) for i in range(k)}

    def _as_dev(self, beta):
        torch = self.torch
        if isinstance(beta, np.ndarray):
            beta = torch.from_numpy(np.ascontiguousarray(beta, dtype=np.float64))
        t = beta.to(device=self.dev, dtype=self.dtype).contiguous()
        if tuple(t.shape) != (9,) + self.shape:
            raise CDDError("expected shape %s, got %s" % ((9,) + self.shape, tuple(t.shape)))
        return t


class Batch:
    """Many independent 2D realisations of uneven sides advanced together
    (one hipGraph-captured step per launch, per-instance CFL clocks)."""

    def __init__(self, sizes, dynamics=GLIDE_ONLY, rk_order=2, Lambda=0.0, mu=1.0, nu=0.3, cfl=0.25, device=0):
        import torch
        self.torch = torch
        cfg = Config()
        cfg.ndim, cfg.n, cfg.dynamics, cfg.rk_order = 2, 0, dynamics, rk_order
        cfg.Lambda, cfg.mu, cfg.nu, cfg.cfl, cfg.device = Lambda, mu, nu, cfl, device
        self.sizes = [int(s) for s in sizes]
        arr = (ctypes.c_int * len(self.sizes))(*self.sizes)
        h = ctypes.c_void_p()
        check(lib.cdd_batch_create(ctypes.byref(cfg), arr, len(self.sizes), ctypes.byref(h)))
        self.h = h
        self.dev = torch.device("cuda", device)
        s = torch.cuda.current_stream(self.dev).cuda_stream
        check(lib.cdd_batch_set_stream(self.h, ctypes.c_void_p(s)))

    def close(self):
        if getattr(self, "h", None):
            lib.cdd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.sizes)

    def set_state(self, i, beta):
        torch = self.torch
        n = self.sizes[i]
        if isinstance(beta, np.ndarray):
            beta = torch.from_numpy(np.ascontiguousarray(beta, dtype=np.float64))
        t = beta.to(device=self.dev, dtype=torch.float64).contiguous()
        if tuple(t.shape) != (9, n, n):
            raise CDDError("instance %d expects shape %s" % (i, (9, n, n)))
        check(lib.cdd_batch_set_state(self.h, i, _ptr(t), 1))

    def get_state(self, i):
        n = self.sizes[i]
        out = self.torch.empty((9, n, n), dtype=self.torch.float64, device=self.dev)
        check(lib.cdd_batch_get_state(self.h, i, _ptr(out), 1))
        return out

    def set_time(self, i, t, t_end=float("inf"), fixed_dt=0.0):
        check(lib.cdd_batch_set_time(self.h, i, t, t_end, fixed_dt))

    def clock(self, i):
        t, dt = ctypes.c_double(), ctypes.c_double()
        done, ns = ctypes.c_int(), ctypes.c_longlong()
        check(lib.cdd_batch_get_clock(self.h, i, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(done),
                                      ctypes.byref(ns)))
        return t.value, dt.value, bool(done.value), ns.value

    def step(self, nsteps=1):
        check(lib.cdd_batch_step(self.h, int(nsteps)))

    def set_tile(self, n, variant):
        check(lib.cdd_batch_set_tile(self.h, int(n), int(variant)))

    def time_flux(self, n, variant, reps=5):
        ms = ctypes.c_double()
        check(lib.cdd_batch_time_flux(self.h, int(n), int(variant), int(reps), ctypes.byref(ms)))
        return ms.value

    def set_flux(self, tile=0, per_cu=2):
        """The multi-instance flux launch's tile variant and persistent blocks
        per CU (per_cu = 0: one flux launch per size instead)."""
        check(lib.cdd_batch_set_flux(self.h, int(tile), int(per_cu)))

    def set_spectral(self, n, variant):
        """Size n's stress-solve row tile (0: default rows per block, 1: half)."""
        check(lib.cdd_batch_set_spectral(self.h, int(n), int(variant)))

    def time_spectral(self, n, variant, reps=5):
        ms = ctypes.c_double()
        check(lib.cdd_batch_time_spectral(self.h, int(n), int(variant), int(reps), ctypes.byref(ms)))
        return ms.value

    def time_flux_multi(self, tile, per_cu, reps=5):
        ms = ctypes.c_double()
        check(lib.cdd_batch_time_flux_multi(self.h, int(tile), int(per_cu), int(reps), ctypes.byref(ms)))
        return ms.value
