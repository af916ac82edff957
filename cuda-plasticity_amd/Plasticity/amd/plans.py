"""Plan cache: one C-ABI plan per (shape, physics, stepping) configuration."""
from . import lib as L

_cache = {}


def get_plan(ndim, n, device, **kw):
    key = (ndim, n, str(device)) + tuple(sorted((k, tuple(v) if isinstance(v, (list, tuple)) else v)
                                                for k, v in kw.items()))
    p = _cache.get(key)
    if p is None:
        idx = device.index if hasattr(device, "index") and device.index is not None else 0
        p = L.Plan(ndim, n, device=idx, **kw)
        _cache[key] = p
    p.bind_stream()
    return p


def analysis_plan(gridShape, mu, nu, device):
    n = gridShape[0]
    if any(s != n for s in gridShape):
        raise ValueError("the GPU stepper supports square grids only")
    return get_plan(len(gridShape), n, device, mu=mu, nu=nu)


def clear():
    for p in _cache.values():
        p.close()
    _cache.clear()
