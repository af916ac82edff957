"""cdd_sigma_noalloc, the stress solve that allocates nothing (its workspace
is the library's static device array), against cdd_sigma of a plan: two
calls in a row, the second on a smaller grid, bit for bit; then a 2D grid at
the array's size limit and a refused oversize grid."""
import ctypes
import sys

import numpy as np
import pytest

from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from Plasticity.amd import lib
    return lib


def _case(L, d, n, seed, **load):
    import torch
    beta = torch.from_numpy(SpectralGaussianField((n,) * d, 2.0, 0.01, seed=seed)).cuda()
    p = L.Plan(d, n, dynamics=L.GLIDE_CLIMB, rk_order=3, Lambda=0.9, mu=1.0, nu=0.3, **load)
    want = p.sigma(beta, 0.35)
    out = torch.full_like(beta, float("nan"))
    stream = torch.cuda.current_stream().cuda_stream
    L.check(L.lib.cdd_sigma_noalloc(ctypes.byref(p.cfg), ctypes.c_void_p(beta.data_ptr()), 0.35,
                                    ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream)))
    torch.cuda.synchronize()
    return out, want


def test_noalloc_sigma_matches_plan_sigma(L):
    import torch
    uniaxial = dict(load_kind=L.LOAD_UNIAXIAL, load_init_strain=0.001, load_uniaxial_rate=0.05)
    for d, n, seed, load in ((3, 64, 1, uniaxial), (3, 32, 2, uniaxial), (2, 768, 3, {}),
                             (2, 128, 4, dict(load_kind=L.LOAD_STRAIN, load_init=(0.001, 0, -0.002) + (0,) * 6,
                                              load_rate=(0.01, 0, 0) + (0,) * 6, load_mu=0.5, load_nu=0.3))):
        say("noalloc sigma", d, n)
        out, want = _case(L, d, n, seed, **load)
        assert torch.equal(out, want), (d, n, (out - want).abs().max().item())


def test_noalloc_sigma_refuses_grids_beyond_its_scratch(L):
    import torch
    say("noalloc sigma oversize")
    cfg = L.Config()
    cfg.ndim, cfg.n, cfg.mu, cfg.nu, cfg.device = 3, 128, 1.0, 0.3, 0
    x = torch.zeros((9, 128, 128, 128), dtype=torch.float64, device="cuda")
    rc = L.lib.cdd_sigma_noalloc(ctypes.byref(cfg), ctypes.c_void_p(x.data_ptr()), 0.0,
                                 ctypes.c_void_p(x.data_ptr()), None)
    assert rc == -3
    assert b"scratch holds" in L.lib.cdd_last_error()
    cfg.precision = L.FP32
    cfg.n = 16
    assert L.lib.cdd_sigma_noalloc(ctypes.byref(cfg), ctypes.c_void_p(x.data_ptr()), 0.0,
                                   ctypes.c_void_p(x.data_ptr()), None) == -3
