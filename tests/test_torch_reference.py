"""The float64 torch restatement the adjoint tests differentiate
(tests/torch_reference.py) against the oracle on the same fields: stress,
flux (both dynamics), RK2/RK3 steps, loads and energy.  CPU only."""
import numpy as np
import pytest
import torch

from oracle import cdd_oracle as O
from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField

import torch_reference as TR

MU, NU = 1.0, 0.3


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a))


@pytest.mark.parametrize("shape", [(16, 16), (8, 8, 8)])
def test_sigma_and_energy(shape):
    beta = SpectralGaussianField(shape, 2.0, 0.01, seed=41)
    s = TR.sigma(t(beta), MU, NU).numpy()
    assert rel(s, O.sigma(beta, MU, NU)) < 1e-14
    e = float(TR.energy(t(beta), MU, NU))
    assert abs(e - O.elastic_energy(beta, MU, NU)) <= 1e-13 * abs(e)


@pytest.mark.parametrize("shape", [(16, 16), (8, 8, 8)])
@pytest.mark.parametrize("dyn,Lambda", [(O.GLIDE_CLIMB, 0.9), (O.GLIDE_CLIMB, 0.0), (O.GLIDE_ONLY, 0.0)])
def test_flux(shape, dyn, Lambda):
    beta = SpectralGaussianField(shape, 2.0, 0.01, seed=42)
    sig = O.sigma(beta, MU, NU)
    lvp = dyn == O.GLIDE_ONLY
    got = TR.flux(t(beta), t(sig), None if lvp else torch.tensor(Lambda, dtype=torch.float64), lvp).numpy()
    want, _ = O.central_upwind_flux(beta, sig, dyn, Lambda)
    assert rel(got, want) < 1e-13


@pytest.mark.parametrize("order", [2, 3])
def test_rk_step_with_uniaxial_control(order):
    shape = (8, 8, 8)
    beta = SpectralGaussianField(shape, 2.0, 0.01, seed=43)
    init, rate, t0 = float(beta[8].mean()), 0.05, 0.3
    load = O.Load("uniaxial", init_strain=init, uniaxial_rate=rate)
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.9, MU, NU, load=load)
    dt = 1e-3
    step = O.tvd_rk3_step if order == 3 else O.tvd_rk2_step
    want, dtw = step(dyn.flux, t0, beta, np.inf, 0.25, fixed_dt=dt)
    got = TR.rk_step(t(beta), dt, torch.tensor(0.9, dtype=torch.float64), False, MU, NU, order,
                     TR.uniaxial_shift(t0, rate, init, MU, NU)).numpy()
    assert dtw == dt
    assert rel(got, want) < 1e-13
    ms = TR.mean_stress(t(beta), MU, NU, TR.uniaxial_shift(t0, rate, init, MU, NU)).numpy()
    np.testing.assert_allclose(ms, O.mean_stress(beta, t0, MU, NU, load), rtol=1e-12, atol=1e-17)
