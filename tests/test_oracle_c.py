"""The C/OpenMP restatement of the oracle (oracle/cdd_oracle.c) against the
NumPy oracle on small grids: flux, stress and whole RK runs agree to
rounding, so the C oracle can stand in for it where NumPy is too slow."""
import numpy as np
import pytest

from oracle import cdd_oracle as O
from oracle import cdd_oracle_c as OC
from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


@pytest.fixture(scope="module", autouse=True)
def built():
    import os
    if not os.path.exists(OC.LIB):
        OC.build()


@pytest.mark.parametrize("shape,dyn,Lam", [((16, 16), O.GLIDE_ONLY, 0.0), ((16, 16), O.GLIDE_CLIMB, 0.9),
                                           ((8, 8, 8), O.GLIDE_CLIMB, 0.9), ((8, 8, 8), O.GLIDE_ONLY, 0.0),
                                           ((8, 8, 8), O.GLIDE_CLIMB, 0.0)])
def test_flux_matches_numpy_oracle(shape, dyn, Lam):
    beta = SpectralGaussianField(shape, 2.0, 0.01, seed=30)
    dyn_ = O.Dynamics(dyn, Lam)
    r, s = OC.flux(dyn_, 0.0, beta)
    rr, ss = dyn_.flux(0.0, beta)
    assert rel(r, rr) < 1e-13 and rel(s, ss) < 1e-14


LOADS = [O.Load(), O.Load("strain", initial=(0.001, 0.0, -0.001), rate=(-0.005, -0.005, 0.01), t0=0.1),
         O.Load("strain", initial=(0.001, 0.0, -0.001), rate=(-0.005, -0.005, 0.01), t0=0.1, elastic=(0.5, 0.3)),
         O.Load("stress", initial=(0.002, 0.0, 0.0), rate=(0.0, 0.0, 0.01)),
         O.Load("uniaxial", init_strain=0.001, uniaxial_rate=0.05),
         O.Load("tensor", tensor0=(0, 0.005, 0, 0.005, 0, 0, 0, 0, 0))]


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("li", range(len(LOADS)))
def test_runs_match_numpy_oracle(order, li):
    beta = SpectralGaussianField((8, 8, 8), 2.0, 0.01, seed=31)
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.9, load=LOADS[li])
    u, t, dts = OC.run(dyn, beta, 0.2, 4, 0.25, order=order)
    ur, tr, dtr = O.run(dyn, beta, 0.2, 4, 0.25, order=order)
    assert len(dts) == len(dtr) and np.allclose(dts, dtr, rtol=1e-13, atol=0)
    assert abs(t - tr) <= 1e-14 * tr
    assert rel(u, ur) < 1e-13


def test_2d_run_matches_numpy_oracle():
    beta = SpectralGaussianField((32, 32), 3.0, 0.01, seed=32)
    dyn = O.Dynamics(O.GLIDE_ONLY, 0.0)
    u, t, dts = OC.run(dyn, beta, 0.0, 10, 0.25, order=2)
    ur, tr, _ = O.run(dyn, beta, 0.0, 10, 0.25, order=2)
    assert rel(u, ur) < 1e-13 and abs(t - tr) <= 1e-14 * tr


def test_end_time_clamp():
    beta = SpectralGaussianField((16, 16), 2.0, 0.01, seed=33)
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.5)
    _, t1, d1 = O.run(dyn, beta, 0.0, 1, 0.25)
    t_end = 2.5 * d1[0]
    u, t, dts = OC.run(dyn, beta, 0.0, 10, 0.25, t_end=t_end)
    ur, tr, dtr = O.run(dyn, beta, 0.0, 10, 0.25, t_end=t_end)
    assert len(dts) == len(dtr) == 3 and abs(t - t_end) < 1e-15 and rel(u, ur) < 1e-13


def test_fixed_step_and_sigma_match_numpy_oracle():
    beta = SpectralGaussianField((8, 8, 8), 2.0, 0.01, seed=34)
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.9, load=O.Load("uniaxial", init_strain=0.001, uniaxial_rate=0.05))
    u, t, dts = OC.run(dyn, beta, 0.1, 3, 0.25, order=3, fixed_dt=0.004)
    ur, tr, dtr = O.run(dyn, beta, 0.1, 3, 0.25, order=3, fixed_dt=0.004)
    assert list(dts) == [0.004] * 3 and list(dtr) == [0.004] * 3
    assert abs(t - tr) <= 1e-15 and rel(u, ur) < 1e-13
    assert rel(OC.sigma(dyn, 0.3, beta), dyn.get_sigma(beta, 0.3)) < 1e-13
