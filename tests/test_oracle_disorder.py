"""Pins of the oracle's disorder restatement (oracle/cdd_oracle.py
disorder_weight, disorder_curl, disorder_sigma, DisorderDynamics;
DisorderState.py:15-30, DisorderDynamics.py:22-39) by closed forms of the
curl term on single Fourier modes (the real k_l factor and irfftn's real
kz = 0 / Nyquist planes), exact linearity, the zero-disorder limit, and the
advected scalar's identity with a tensor row that has no subtraction term."""
import numpy as np
import pytest

from oracle import cdd_oracle as O
from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField

MU, NU = 0.5, 0.3


def relmax(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _mode(shape, axis, m, kind="cos"):
    n = shape[axis]
    idx = np.arange(n)
    f = np.cos(2 * np.pi * m * idx / n) if kind == "cos" else np.sin(2 * np.pi * m * idx / n)
    sh = [1] * len(shape)
    sh[axis] = n
    return np.broadcast_to(f.reshape(sh), shape).copy()


@pytest.mark.parametrize("kind", ["cos", "sin"])
@pytest.mark.parametrize("t,j", [(1, 0), (2, 1), (0, 2)])
def test_curl_of_a_mode_along_the_last_axis_3d(kind, t, j):
    """Along the rfft axis the factor k_z acts as |d/dz|: a mode m gives
    2 pi m times the same mode, in component (i, j) with sign perm[z, i, t]."""
    shape, m, a = (8, 8, 8), 3, 0.7
    rp = np.zeros((9,) + shape)
    rp[O.cidx(t, j)] = a * _mode(shape, 2, m, kind)
    out = O.disorder_curl(rp)
    want = np.zeros_like(out)
    for i in range(3):
        want[O.cidx(i, j)] = O.PERM[2, i, t] * 2 * np.pi * m * a * _mode(shape, 2, m, kind)
    assert np.max(np.abs(out - want)) < 1e-12


@pytest.mark.parametrize("axis", [0, 1])
def test_curl_of_a_mode_along_another_axis_vanishes_below_nyquist_3d(axis):
    """k_l times a mode along a non-rfft axis is purely imaginary on the
    kz = 0 plane, which irfftn reads as real: no contribution."""
    shape = (8, 8, 8)
    rp = np.zeros((9,) + shape)
    rp[O.cidx(1, 0)] = _mode(shape, axis, 2)
    rp[O.cidx(2, 2)] = _mode(shape, axis, 1, "sin")
    assert np.max(np.abs(O.disorder_curl(rp))) < 1e-12


@pytest.mark.parametrize("axis", [0, 1])
def test_curl_of_the_nyquist_mode_along_another_axis_3d(axis):
    """At the Nyquist index the wavenumber is +pi N on its own (no -k
    partner), so (-1)^x_l survives with factor pi N and sign perm[l, i, t]."""
    shape, a = (8, 8, 8), 0.3
    n = shape[axis]
    rp = np.zeros((9,) + shape)
    t, j = 1, 0
    rp[O.cidx(t, j)] = a * _mode(shape, axis, n // 2)
    out = O.disorder_curl(rp)
    want = np.zeros_like(out)
    for i in range(3):
        want[O.cidx(i, j)] = O.PERM[axis, i, t] * np.pi * n * a * _mode(shape, axis, n // 2)
    assert np.max(np.abs(out - want)) < 1e-11


def test_curl_2d_last_axis_and_no_z_terms():
    """2D: the rfft axis is y; l = z is skipped (and kz = 0 anyway)."""
    shape, m, a = (16, 16), 5, 1.1
    rp = np.zeros((9,) + shape)
    rp[O.cidx(2, 1)] = a * _mode(shape, 1, m)  # t = z, j = y
    rp[O.cidx(0, 0)] = a * _mode(shape, 1, m, "sin")  # t = x, j = x
    out = O.disorder_curl(rp)
    want = np.zeros_like(out)
    for i in range(3):
        want[O.cidx(i, 1)] += O.PERM[1, i, 2] * 2 * np.pi * m * a * _mode(shape, 1, m)
        want[O.cidx(i, 0)] += O.PERM[1, i, 0] * 2 * np.pi * m * a * _mode(shape, 1, m, "sin")
    assert np.max(np.abs(out - want)) < 1e-11
    assert np.any(want != 0)


def test_weight_matches_its_modulus():
    """rho' = d rho / (ME + 2 |rho|_F^2), summed here as a Frobenius norm."""
    b = SpectralGaussianField((16, 16), 2.0, 0.01, seed=1)
    rho = O.rho_fourier(b)
    d = SpectralGaussianField((16, 16), 3.0, 0.05, seed=2)[0]
    w = O.disorder_weight(rho, d)
    fro = np.einsum("c...,c...->...", rho, rho)
    assert relmax(w, d * rho / (O.ME + 2 * fro)) < 1e-14
    assert np.max(np.abs(O.disorder_weight(rho, np.zeros_like(d)))) == 0.0


@pytest.mark.parametrize("shape", [(16, 16), (8, 8, 8)])
def test_disorder_sigma_limits(shape):
    b = SpectralGaussianField(shape, 2.0, 0.01, seed=3)
    d = SpectralGaussianField(shape, 2.0, 0.05, seed=4)[0]
    s0 = O.sigma(b, MU, NU)
    assert np.array_equal(O.disorder_sigma(b, np.zeros(shape), MU, NU), s0)
    ds1 = O.disorder_sigma(b, d, MU, NU) - s0
    ds2 = O.disorder_sigma(b, 2 * d, MU, NU) - s0
    assert relmax(ds2, 2 * ds1) < 1e-12
    assert np.max(np.abs(ds1)) > 1e-3 * np.max(np.abs(s0))  # a real contribution
    # not symmetric in general
    assert np.max(np.abs(ds1[O.cidx(0, 1)] - ds1[O.cidx(1, 0)])) > 1e-6 * np.max(np.abs(ds1))


@pytest.mark.parametrize("dyn,Lambda", [(O.GLIDE_CLIMB, 0.0), (O.GLIDE_CLIMB, 0.7), (O.GLIDE_ONLY, 0.0)])
def test_advected_scalar_is_the_zx_row_in_2d(dyn, Lambda):
    """In 2D the zx row of H has no subtraction and no climb term, so with
    d = beta_zx the scalar's rhs is the zx rhs bit for bit."""
    b = SpectralGaussianField((16, 16), 2.0, 0.01, seed=5)
    u = np.concatenate([b, b[O.cidx(2, 0)][None]])
    dd = O.DisorderDynamics(dyn, Lambda, MU, NU)
    rhs, speed = dd.flux(0.0, u)
    assert rhs.shape == (10, 16, 16)
    assert np.array_equal(rhs[9], rhs[O.cidx(2, 0)])
    # the beta part is the plain flux with the disorder stress
    ref, sp = O.central_upwind_flux(b, dd.get_sigma(u, 0.0), dyn, Lambda)
    assert np.array_equal(rhs[:9], ref) and np.array_equal(speed, sp)


@pytest.mark.parametrize("dyn,Lambda", [(O.GLIDE_CLIMB, 0.7), (O.GLIDE_ONLY, 0.0)])
def test_advected_scalar_3d_is_the_central_upwind_combination(dyn, Lambda):
    """3D: rhs_d = -sum_o w_o sum_a V_o[a] d_a^o / (a b c)_tot
    + sum_a (d_a^+ - d_a^-) a_m a_p / a_tot, built here from the oracle's
    velocities; a uniform d has rhs 0."""
    n = 8
    b = SpectralGaussianField((n,) * 3, 2.0, 0.01, seed=6)
    d = SpectralGaussianField((n,) * 3, 2.0, 0.05, seed=7)[0]
    u = np.concatenate([b, d[None]])
    dd = O.DisorderDynamics(dyn, Lambda, MU, NU)
    sig = dd.get_sigma(u, 0.0)
    rhs, _ = dd.flux(0.0, u)
    lam = 0.0 if dyn == O.GLIDE_ONLY else Lambda
    du = [O.find_derivatives(b, a) for a in range(3)]
    dd_ = [O.find_derivatives(d[None], a) for a in range(3)]
    opts = ['---', '--+', '-+-', '-++', '+--', '+-+', '++-', '+++']
    sel = {'+': 0, '-': 1}
    V = {o: O._velocity(dyn, [du[a][sel[o[a]]] for a in range(3)], sig, lam) for o in opts}
    ap = [O.max9(*[V[o][a] for o in opts], 0.) for a in range(3)]
    am = [np.fabs(O.min9(*[V[o][a] for o in opts], 0.)) for a in range(3)]
    at = [ap[a] + am[a] + O.ME for a in range(3)]
    w = {'+': am, '-': ap}
    acc = 0.0
    for o in opts:
        hs = sum(V[o][a] * dd_[a][sel[o[a]]][0] for a in range(3))
        acc = acc - hs * w[o[0]][0] * w[o[1]][1] * w[o[2]][2]
    want = acc / at[0] / at[1] / at[2]
    for a in range(3):
        want = want + (dd_[a][0][0] - dd_[a][1][0]) * am[a] * ap[a] / at[a]
    assert relmax(rhs[9], want) < 1e-12
    u[9] = 0.25
    rhs, _ = dd.flux(0.0, u)
    assert np.max(np.abs(rhs[9])) == 0.0


def test_runs_without_disorder_are_the_plain_runs():
    b = SpectralGaussianField((16, 16), 2.0, 0.01, seed=8)
    plain = O.Dynamics(O.GLIDE_CLIMB, 0.5, MU, NU)
    ref, tref, dref = O.run(plain, b, 0.0, 4, 0.25, order=3)
    adv = O.DisorderDynamics(O.GLIDE_CLIMB, 0.5, MU, NU)
    u, t, dts = O.run(adv, np.concatenate([b, np.zeros((1, 16, 16))]), 0.0, 4, 0.25, order=3)
    assert np.array_equal(u[:9], ref) and np.array_equal(dts, dref) and not np.any(u[9])
    fixed = O.DisorderDynamics(O.GLIDE_CLIMB, 0.5, MU, NU, disorder=np.zeros((16, 16)))
    u2, _, _ = O.run(fixed, b, 0.0, 4, 0.25, order=3)
    assert np.array_equal(u2, ref)
