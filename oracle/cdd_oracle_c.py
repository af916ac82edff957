"""ctypes binding of the C/OpenMP restatement of the oracle (oracle/cdd_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/ (checked against the NumPy oracle,
oracle/cdd_oracle.py, on small grids) to check the GPU path at sizes the
NumPy oracle cannot reach, and by bench.py's cpu_baseline leg.  The product
path never loads it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cdd_oracle.c")
LIB = os.path.join(HERE, "libcdd_oracle.so")


def build():
    """gcc -O3 -fopenmp; no FP contraction, so each operation rounds as in NumPy."""
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O3", "-std=gnu11", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", SRC, "-o", tmp, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle C build failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    return LIB


class _Cfg(ctypes.Structure):
    _fields_ = [("d", ctypes.c_int), ("n", ctypes.c_int), ("dyn", ctypes.c_int), ("order", ctypes.c_int),
                ("Lambda", ctypes.c_double), ("mu", ctypes.c_double), ("nu", ctypes.c_double),
                ("cfl", ctypes.c_double), ("load_kind", ctypes.c_int), ("init", ctypes.c_double * 9),
                ("rate", ctypes.c_double * 9), ("t0", ctypes.c_double), ("init_strain", ctypes.c_double),
                ("uniaxial_rate", ctypes.c_double), ("load_mu", ctypes.c_double), ("load_nu", ctypes.c_double),
                ("fixed_dt", ctypes.c_double)]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("libcdd_oracle.so is not built; run __graft_entry__.build()")
        _lib = ctypes.CDLL(LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        _lib.oc_run.argtypes = [ctypes.POINTER(_Cfg), dp, ctypes.c_double, ctypes.c_int, ctypes.c_double, dp, dp]
        _lib.oc_run.restype = ctypes.c_int
        _lib.oc_sigma.argtypes = [ctypes.POINTER(_Cfg), dp, ctypes.c_double, dp]
        _lib.oc_flux.argtypes = [ctypes.POINTER(_Cfg), dp, dp, dp, dp]
    return _lib


def _cfg(beta, dynamics, Lambda, mu, nu, cfl, order, load):
    from oracle import cdd_oracle as O
    d = beta.ndim - 1
    n = beta.shape[1]
    if any(s != n for s in beta.shape[1:]) or n & (n - 1):
        raise ValueError("square power-of-two grids only")
    c = _Cfg(d=d, n=n, dyn=1 if dynamics == O.GLIDE_ONLY else 0, order=order, Lambda=Lambda, mu=mu, nu=nu, cfl=cfl)
    if load is not None and load.kind != "none":
        if load.kind == "strain":
            c.load_kind = 1
            c.init[:3], c.rate[:3], c.t0 = list(load.initial), list(load.rate), load.t0
            if load.elastic is not None:
                c.load_mu, c.load_nu = load.elastic
        elif load.kind == "stress":
            c.load_kind = 2
            c.init[:3], c.rate[:3], c.t0 = list(load.initial), list(load.rate), load.t0
        elif load.kind == "uniaxial":
            c.load_kind = 3
            c.init_strain, c.uniaxial_rate = load.init_strain, load.uniaxial_rate
        elif load.kind == "tensor":
            c.load_kind = 4
            c.init[:], c.rate[:], c.t0 = list(load.tensor0), list(load.tensor_rate), load.t0
        else:
            raise ValueError(load.kind)
    return c


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def run(dyn, beta, t0, nsteps, cfl, order=3, t_end=np.inf, fixed_dt=None):
    """Same contract as oracle.cdd_oracle.run for a cdd_oracle.Dynamics:
    returns (u, t, dts)."""
    L = _load()
    u = np.ascontiguousarray(beta, dtype=np.float64).copy()
    c = _cfg(u, dyn.dynamics, dyn.Lambda, dyn.mu, dyn.nu, cfl, order, dyn.load)
    if fixed_dt is not None:
        c.fixed_dt = float(fixed_dt)
    dts = np.zeros(nsteps)
    t = ctypes.c_double()
    k = L.oc_run(ctypes.byref(c), _dp(u), float(t0), int(nsteps), float(t_end), ctypes.byref(t), _dp(dts))
    return u, t.value, dts[:k]


def sigma(dyn, t, beta):
    """dyn.get_sigma(beta, t): the stress with the dynamics' load."""
    L = _load()
    u = np.ascontiguousarray(beta, dtype=np.float64)
    c = _cfg(u, dyn.dynamics, dyn.Lambda, dyn.mu, dyn.nu, 0.25, 3, dyn.load)
    sig = np.empty_like(u)
    L.oc_sigma(ctypes.byref(c), _dp(u), float(t), _dp(sig))
    return sig


def flux(dyn, t, beta):
    """(rhs, speed) like cdd_oracle.Dynamics.flux."""
    L = _load()
    u = np.ascontiguousarray(beta, dtype=np.float64)
    c = _cfg(u, dyn.dynamics, dyn.Lambda, dyn.mu, dyn.nu, 0.25, 3, dyn.load)
    sig = np.empty_like(u)
    rhs = np.empty_like(u)
    speed = np.empty(u.shape[1:])
    L.oc_sigma(ctypes.byref(c), _dp(u), float(t), _dp(sig))
    L.oc_flux(ctypes.byref(c), _dp(u), _dp(sig), _dp(rhs), _dp(speed))
    return rhs, speed
