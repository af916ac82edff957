"""Pins of the oracle's vacancy coupling (oracle/cdd_oracle.py: Vacancy,
vacancy_diffuse, vacancy_sigma, run_vacancy), following
FieldDynamics/VacancyDynamics.py:16-70 and the operator-splitting mover
FieldMovers/FieldMover.py:318-380 -- closed forms and invariants, no
reference code is run."""
import numpy as np
import pytest

from oracle import cdd_oracle as O
from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField


@pytest.mark.parametrize("shape", [(16, 16), (8, 8, 8)])
def test_diffusion_of_single_modes_is_the_heat_kernel(shape):
    """c = c0 + a cos(2 pi m x_i / N): each mode decays by exp(-gamma dt k^2)
    with k = 2 pi m (radians per unit length of the unit cell); the mean is
    untouched."""
    n = shape[0]
    gamma, dt = 0.37, 0.0123
    grids = np.meshgrid(*[np.arange(n)] * len(shape), indexing="ij")
    for axis, m in [(0, 3), (len(shape) - 1, 2)]:
        c = 0.3 + 0.2 * np.cos(2 * np.pi * m * grids[axis] / n)
        got = O.vacancy_diffuse(c, dt, gamma)
        want = 0.3 + 0.2 * np.exp(-gamma * dt * (2 * np.pi * m) ** 2) * np.cos(2 * np.pi * m * grids[axis] / n)
        assert np.max(np.abs(got - want)) < 1e-14


def test_vacancy_stress_is_a_uniform_pressure_shift():
    """sigma_ii -= alpha c, nothing else changes."""
    beta = SpectralGaussianField((16, 16), 2.0, 0.01, seed=3)
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.5, 1.0, 0.3)
    c = np.random.default_rng(1).standard_normal((16, 16)) * 0.01
    vac = O.Vacancy(alpha=0.7)
    s0 = dyn.get_sigma(beta, 0.0)
    s1 = O.vacancy_sigma(dyn, vac, beta, c, 0.0)
    for k in range(9):
        want = s0[k] - 0.7 * c if k in (0, 4, 8) else s0[k]
        assert np.max(np.abs(s1[k] - want)) < 1e-15


@pytest.mark.parametrize("order", [2, 3])
def test_vacancy_mass_balance(order):
    """Diffusion keeps sum(c); so sum(c') - sum(c) = beta (sum tr beta' -
    sum tr beta) over a step -- the sign and the factor of the coupling."""
    beta = SpectralGaussianField((16, 16), 2.0, 0.01, seed=4)
    c = 0.01 * SpectralGaussianField((16, 16), 2.0, 1.0, seed=5)[0]
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.3, 1.0, 0.3)
    vac = O.Vacancy(alpha=1.0, beta=0.6, gamma=0.05)
    u, c1, t, dts = O.run_vacancy(dyn, vac, beta, c, 0.0, 1, 0.25, order=order)
    dtr = (u[0] + u[4] + u[8]).sum() - (beta[0] + beta[4] + beta[8]).sum()
    assert abs(dtr) > 1e-8  # climb changes the trace here
    assert abs((c1.sum() - c.sum()) - 0.6 * dtr) < 1e-12 * max(1.0, abs(c.sum()))


def test_without_coupling_c_only_diffuses_over_the_step():
    """beta = 0: the concentration changes only by diffusion, over the
    step's own dt, after the beta^P update."""
    beta = SpectralGaussianField((16, 16), 2.0, 0.01, seed=6)
    c = 0.01 * SpectralGaussianField((16, 16), 2.0, 1.0, seed=7)[0]
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.4, 1.0, 0.3)
    u, c1, t, dts = O.run_vacancy(dyn, O.Vacancy(alpha=1.0, beta=0.0, gamma=0.2), beta, c, 0.0, 2, 0.25)
    assert np.max(np.abs(c1 - O.vacancy_diffuse(O.vacancy_diffuse(c, dts[0], 0.2), dts[1], 0.2))) < 1e-17


def test_uniform_concentration_acts_as_a_uniform_compressive_load():
    """A uniform c0 (beta = gamma = 0 keep it uniform) is a stress load of
    -alpha c0 on the diagonal: the run equals the plain run with that
    'stress' load, bit for bit."""
    beta = SpectralGaussianField((16, 16), 2.0, 0.01, seed=11)
    c0 = 0.02
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.4, 1.0, 0.3)
    u, c1, t, dts = O.run_vacancy(dyn, O.Vacancy(alpha=0.9, beta=0.0, gamma=0.0), beta,
                                  np.full((16, 16), c0), 0.0, 3, 0.25)
    loaded = O.Dynamics(O.GLIDE_CLIMB, 0.4, 1.0, 0.3, O.Load("stress", initial=(-0.9 * c0,) * 3))
    ur, tr, dtr = O.run(loaded, beta, 0.0, 3, 0.25)
    assert np.array_equal(u, ur) and list(dts) == list(dtr)
    plain, _, _ = O.run(dyn, beta, 0.0, 3, 0.25)
    assert np.max(np.abs(u - plain)) > 1e-9  # the load does act


def test_dt_bound_caps_the_adaptive_step():
    beta = SpectralGaussianField((16, 16), 2.0, 0.01, seed=8)
    c = np.zeros((16, 16))
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.5, 1.0, 0.3)
    _, _, _, dts = O.run_vacancy(dyn, O.Vacancy(), beta, c, 0.0, 2, 0.25)
    cap = 0.5 * dts[0]
    _, _, t, dts2 = O.run_vacancy(dyn, O.Vacancy(dt_bound=cap), beta, c, 0.0, 2, 0.25)
    assert list(dts2) == [cap, cap] and t == 2 * cap


def test_zero_coupling_is_the_plain_run():
    """alpha = beta = gamma = 0: beta^P follows the plain mover bit for bit and
    c stays put."""
    beta = SpectralGaussianField((16, 16), 2.0, 0.01, seed=9)
    c = 0.01 * SpectralGaussianField((16, 16), 2.0, 1.0, seed=10)[0]
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.5, 1.0, 0.3)
    u, c1, t, dts = O.run_vacancy(dyn, O.Vacancy(0.0, 0.0, 0.0), beta, c, 0.0, 3, 0.25)
    ur, tr, dtr = O.run(dyn, beta, 0.0, 3, 0.25)
    assert np.array_equal(u, ur) and t == tr
    assert np.max(np.abs(c1 - c)) < 1e-16


def test_constant_alpha_field_equals_scalar_alpha():
    """An alpha field that is constant is the scalar alpha, bit for bit (the
    reference's alpha * cfield broadcasts either)."""
    beta = SpectralGaussianField((16, 16), 2.0, 0.01, seed=81)
    c = 0.01 * SpectralGaussianField((16, 16), 2.0, 1.0, seed=82)[0]
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.5, 1.0, 0.3)
    a = O.run_vacancy(dyn, O.Vacancy(0.7, 1.0, 0.05), beta, c, 0.0, 3, 0.25)
    b = O.run_vacancy(dyn, O.Vacancy(np.full((16, 16), 0.7), 1.0, 0.05), beta, c, 0.0, 3, 0.25)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[3], b[3])
