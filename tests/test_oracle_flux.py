"""Pins of the oracle's limiter, velocities, Hamiltonians, central-upwind
combination and TVD Runge-Kutta steppers (closed forms and identities)."""
import numpy as np
import pytest

from oracle import cdd_oracle as O
from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField


def textbook_minmod(a, b, c):
    out = np.zeros(np.broadcast(a, b, c).shape)
    pos = (a > 0) & (b > 0) & (c > 0)
    neg = (a <= 0) & (b <= 0) & (c <= 0)
    out[pos] = np.minimum(np.minimum(a, b), c)[pos]
    out[neg] = np.maximum(np.maximum(a, b), c)[neg]
    return out


def test_minmod_table_and_selections():
    rng = np.random.default_rng(0)
    vals = np.array([-2.0, -1.0, -0.5, 0.0, 0.5, 1.0, 2.0])
    a, b, c = np.meshgrid(vals, vals, vals, indexing="ij")
    np.testing.assert_array_equal(O.minmod3(a, b, c), textbook_minmod(a, b, c))
    r = rng.standard_normal((3, 1000))
    np.testing.assert_array_equal(O.minmod3(*r), textbook_minmod(*r))
    r = rng.standard_normal((9, 500))
    np.testing.assert_array_equal(O.max9(*r), r.max(axis=0))
    np.testing.assert_array_equal(O.min9(*r), r.min(axis=0))
    np.testing.assert_array_equal(O.max5(*r[:4], 0.), np.maximum(r[:4].max(axis=0), 0))
    np.testing.assert_array_equal(O.min5(*r[:4], 0.), np.minimum(r[:4].min(axis=0), 0))


def test_limited_derivatives_exact_on_quadratic_and_linear():
    n = 32
    Dx = 1.0 / n
    i = np.arange(n, dtype=float)
    for f, fp in [(0.7 * i * i, 1.4 * i / Dx), (-0.3 * i * i, -0.6 * i / Dx), (2.5 * i, 2.5 / Dx * np.ones(n))]:
        field = f[None, :, None] * np.ones((1, n, n))  # Dx = 1/N_last
        p, m = O.find_derivatives(field, 0)
        sl = slice(2, n - 2)
        np.testing.assert_allclose(p[0, sl, 0], fp[sl], rtol=1e-13)
        np.testing.assert_allclose(m[0, sl, 0], fp[sl], rtol=1e-13)
    # at a local extremum the limiter slope is zero: u+ = D_i/Dx, u- = D_{i-1}/Dx
    f = np.zeros(n)
    f[10] = 1.0
    field = f[None, :, None] * np.ones((1, n, n))  # Dx = 1/N_last
    p, m = O.find_derivatives(field, 0)
    assert p[0, 10, 0] == -1.0 / Dx and m[0, 10, 0] == 1.0 / Dx


def _random_corner(d, shape=(40,), seed=0):
    rng = np.random.default_rng(seed)
    us = [rng.standard_normal((9,) + shape) for _ in range(d)]
    sig = rng.standard_normal((9,) + shape)
    for i in range(3):
        for j in range(i):
            sig[O.cidx(i, j)] = sig[O.cidx(j, i)]
    return us, sig


def _levi_civita_rho(us):
    """rho_ij = -eps_ilm u_l[m, j] with u_l = 0 for l >= d."""
    d = len(us)
    U = np.zeros((3, 3, 3) + us[0].shape[1:])  # U[l, m, j]
    for l in range(d):
        U[l] = us[l].reshape((3, 3) + us[0].shape[1:])
    return -np.einsum("ilm,lmj...->ij...", O.PERM, U), U


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("Lambda", [0.0, 0.9, 1.0])
def test_velocity_levi_civita_form(d, Lambda):
    us, sig = _random_corner(d, seed=d)
    v = O.velocity_glide_climb(us, sig, Lambda)
    rho, U = _levi_civita_rho(us)
    rhomod = np.sqrt(O.ME + np.sum(rho * rho, axis=(0, 1)))
    S = sig.reshape((3, 3) + sig.shape[1:]).copy()
    tr = S[0, 0] + S[1, 1] + S[2, 2]
    for i in range(3):
        S[i, i] = S[i, i] - Lambda * tr / 3.0
    # v_l = S_st (d_s beta_lt - d_l beta_st) / |rho|
    expect = (np.einsum("st...,slt...->l...", S, U) - np.einsum("st...,lst...->l...", S, U)) / rhomod
    np.testing.assert_allclose(np.array(v), expect, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("Lambda", [0.0, 0.9])
def test_hamiltonian_levi_civita_and_trace(d, Lambda):
    us, sig = _random_corner(d, seed=10 + d)
    V = O.velocity_glide_climb(us, sig, Lambda)
    H = O.hamiltonian(us, V, Lambda)
    rho, U = _levi_civita_rho(us)
    Va = np.array(V)
    H0 = np.einsum("ilm,l...,mj...->ij...", O.PERM, Va, rho)
    trH0 = H0[0, 0] + H0[1, 1] + H0[2, 2]
    expect = H0.copy()
    for i in range(3):
        expect[i, i] = expect[i, i] - Lambda / 3.0 * trH0
    np.testing.assert_allclose(H.reshape((3, 3) + H.shape[1:]), expect, rtol=1e-10, atol=1e-11)
    trH = H[0] + H[4] + H[8]
    np.testing.assert_allclose(trH, (1.0 - Lambda) * trH0, rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("d", [2, 3])
def test_lvp_velocity_identities(d):
    us, sig = _random_corner(d, seed=20 + d)
    v = np.array(O.velocity_lvp(us, sig))
    v0 = np.array(O.velocity_glide_climb(us, sig, 0.0))  # already / |rho|
    rho, U = _levi_civita_rho(us)
    rhomod = np.sqrt(O.ME + np.sum(rho * rho, axis=(0, 1)))
    # q_a = sum_{b != a} u_a[b, b] - sum_{b != a} u_b[a, b]  ( = -rho_bb-type trace vector)
    q = np.zeros((3,) + sig.shape[1:])
    for a in range(3):
        for b in range(3):
            if b != a:
                q[a] += U[a, b, b] - U[b, a, b]
    qq = np.sum(q * q, axis=0)
    w0 = v0 * rhomod  # unnormalised climb velocity
    rhov = O.ME + np.sum(w0 * q, axis=0)
    expect = (w0 + q * rhov / (O.ME + qq)) / rhomod
    np.testing.assert_allclose(v, expect, rtol=1e-9, atol=1e-11)
    # v'.q = v0.q + |q|^2 (ME + v0.q)/(ME + |q|^2)  (code sign, not the docstring's)
    lhs = np.sum(v * rhomod * q, axis=0)
    w0q = np.sum(w0 * q, axis=0)
    np.testing.assert_allclose(lhs, w0q + qq * (O.ME + w0q) / (O.ME + qq), rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("dyn,Lambda", [(O.GLIDE_CLIMB, 0.0), (O.GLIDE_CLIMB, 1.0), (O.GLIDE_ONLY, 0.0)])
@pytest.mark.parametrize("shape", [(8, 8), (6, 6, 6)])
def test_flux_constant_field(dyn, Lambda, shape):
    beta = np.ones((9,) + shape) * 0.01 * np.arange(1, 10)[(slice(None),) + (None,) * len(shape)]
    sig = O.sigma(beta, 1.0, 0.3)
    rhs, speed = O.central_upwind_flux(beta, sig, dyn, Lambda)
    assert np.all(rhs == 0.0)
    np.testing.assert_allclose(speed, len(shape) * O.ME, rtol=1e-15)


@pytest.mark.parametrize("dyn,Lambda", [(O.GLIDE_CLIMB, 0.9), (O.GLIDE_ONLY, 0.0)])
@pytest.mark.parametrize("shape", [(16, 16), (8, 8, 8)])
def test_flux_translation_equivariance(dyn, Lambda, shape):
    beta = SpectralGaussianField(shape, 2.0, 0.01, seed=1)
    sig = O.sigma(beta, 1.0, 0.3)
    rhs, speed = O.central_upwind_flux(beta, sig, dyn, Lambda)
    sh = (3, -2, 1)[:len(shape)]
    ax = tuple(range(1, len(shape) + 1))
    rhs2, speed2 = O.central_upwind_flux(np.roll(beta, sh, ax), np.roll(sig, sh, ax), dyn, Lambda)
    np.testing.assert_array_equal(rhs2, np.roll(rhs, sh, ax))
    np.testing.assert_array_equal(speed2, np.roll(speed, sh, tuple(a - 1 for a in ax)))


def _rotate90_2d(f):
    """Rotate a (9, N, N) tensor field by +90 degrees about z (about grid
    point 0): f'(r) = Q f(Q^T r) Q^T with Q = [[0,-1,0],[1,0,0],[0,0,1]]."""
    Q = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], dtype=float)
    n = f.shape[1]
    i = np.arange(n)
    # Q^T r = (y, -x): new[x, y] = old[y, -x]
    X, Yy = np.meshgrid(i, i, indexing="ij")
    g = f[:, Yy, (-X) % n]
    T = g.reshape((3, 3, n, n))
    return np.einsum("ia,abxy,jb->ijxy", Q, T, Q).reshape((9, n, n))


@pytest.mark.parametrize("dyn,Lambda", [(O.GLIDE_CLIMB, 0.0), (O.GLIDE_CLIMB, 0.9), (O.GLIDE_ONLY, 0.0)])
def test_flux_rotation_equivariance_2d(dyn, Lambda):
    shape = (16, 16)
    # L = 3 keeps the Nyquist modes (where k = +pi N breaks the rotation
    # symmetry of the half-spectrum Green's tensor) at the 1e-20 level
    beta = SpectralGaussianField(shape, 3.0, 0.01, seed=2)
    dyn_ = O.Dynamics(dyn, Lambda)
    rhs, speed = dyn_.flux(0.0, beta)
    rhs_r, speed_r = dyn_.flux(0.0, _rotate90_2d(beta))
    scale = np.max(np.abs(rhs))
    np.testing.assert_allclose(rhs_r, _rotate90_2d(rhs), atol=1e-12 * scale)
    # speed: a_tot and b_tot swap roles; their sum is invariant
    n = shape[0]
    i = np.arange(n)
    X, Yy = np.meshgrid(i, i, indexing="ij")
    np.testing.assert_allclose(speed_r, speed[Yy, (-X) % n], rtol=1e-12)


@pytest.mark.parametrize("dyn,Lambda", [(O.GLIDE_CLIMB, 0.0), (O.GLIDE_CLIMB, 0.9), (O.GLIDE_ONLY, 0.0)])
def test_flux_3d_z_invariant_equals_2d(dyn, Lambda):
    n = 16
    b2 = SpectralGaussianField((n, n), 2.0, 0.01, seed=3)
    b3 = np.repeat(b2[:, :, :, None], n, axis=3)
    d2 = O.Dynamics(dyn, Lambda)
    rhs2, sp2 = d2.flux(0.0, b2)
    rhs3, sp3 = d2.flux(0.0, b3)
    for k in range(n):
        # differs only through c_tot = c_p + c_m + ME in 3D
        np.testing.assert_allclose(rhs3[..., k], rhs2, rtol=0, atol=1e-6 * np.max(np.abs(rhs2)))


class _Linear:
    def __init__(self, lam, speed):
        self.lam, self.speed = lam, speed

    def __call__(self, t, u):
        return self.lam * u, np.full(u.shape[1:], self.speed)


@pytest.mark.parametrize("order", [2, 3])
def test_rk_amplification(order):
    u = np.ones((9, 4, 4))
    lam, speed, cfl = -3.0, 2.0, 0.25
    step = O.tvd_rk3_step if order == 3 else O.tvd_rk2_step
    u1, dt = step(_Linear(lam, speed), 0.0, u, np.inf, cfl)
    assert dt == cfl / (4 * speed)
    z = lam * dt
    amp = 1 + z + z * z / 2 + (z ** 3 / 6 if order == 3 else 0.0)
    np.testing.assert_allclose(u1, amp, rtol=1e-15)
    # clamp to the end time
    u1, dt = step(_Linear(lam, speed), 1.0, u, 1.0 + 1e-3, cfl)
    assert dt == pytest.approx(1e-3, rel=1e-12)


def test_stage_time_is_step_start():
    seen = []

    def flux(t, u):
        seen.append(t)
        return 0 * u, np.ones(u.shape[1:])

    O.tvd_rk3_step(flux, 0.7, np.ones((9, 2, 2)), np.inf, 0.1)
    assert seen == [0.7, 0.7, 0.7]


def test_energy_goes_downhill_small_cfl():
    beta = SpectralGaussianField((16, 16), 2.0, 0.01, seed=4)
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.0)
    e = [O.elastic_energy(beta)]
    u = beta
    t = 0.0
    for _ in range(5):
        u, dt = O.tvd_rk3_step(dyn.flux, t, u, np.inf, 0.05)
        t += dt
        e.append(O.elastic_energy(u))
    assert all(e[i + 1] <= e[i] * (1 + 1e-12) for i in range(5))


def test_wojzz_reduces_to_glide_climb_at_lambda_zero_and_freezes_zz():
    from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField
    beta = SpectralGaussianField((16, 16), 2.5, 0.01, seed=40)
    sig = O.sigma(beta)
    r0, s0 = O.central_upwind_flux(beta, sig, O.GLIDE_CLIMB, 0.0)
    rw, sw = O.central_upwind_flux(beta, sig, O.GLIDE_CLIMB_WOJZZ, 0.0)
    assert np.array_equal(rw[:8], r0[:8]) and np.array_equal(sw, s0)
    assert not rw[8].any()
    rw, _ = O.central_upwind_flux(beta, sig, O.GLIDE_CLIMB_WOJZZ, 0.7)
    assert not rw[8].any()
    # the Lambda term changes the flux
    assert np.abs(rw[:8] - r0[:8]).max() > 1e-6 * np.abs(r0).max()


def test_wojzz_rotation_equivariance():
    """x and y enter the woJzz climb term symmetrically: the flux commutes
    with the 90-degree rotation of FieldInitializer.RotateState_90degree."""
    from Plasticity.FieldInitializers.FieldInitializer import RotateState_90degree, SpectralGaussianField
    from Plasticity.PlasticityStates.PlasticityState import PlasticityState
    beta = SpectralGaussianField((16, 16), 2.5, 0.01, seed=41)
    rot = lambda a: RotateState_90degree(PlasticityState((16, 16), field=a, device="cpu"), device="cpu").GetOrderParameterField().numpy()
    dyn = O.Dynamics(O.GLIDE_CLIMB_WOJZZ, 0.6)
    r = dyn.flux(0.0, beta)[0]
    rr = dyn.flux(0.0, rot(beta))[0]
    np.testing.assert_allclose(rr, rot(r), rtol=0, atol=1e-12 * np.abs(r).max())


@pytest.mark.parametrize("d,n", [(2, 32), (3, 12)])
@pytest.mark.parametrize("dyn,lam", [(O.GLIDE_CLIMB, 0.9), (O.GLIDE_CLIMB, 0.0), (O.GLIDE_ONLY, 0.0)])
def test_one_signed_speeds_give_the_upwind_hamiltonian(d, n, dyn, lam):
    """Where every corner velocity has one sign along each axis, the
    central-upwind combination is the pure upwind scheme: rhs = -H at the
    corner of the upwind one-sided derivatives (u^- where the speed is
    positive, u^+ where it is negative), up to the ME regularisers of the
    speed totals, and the numerical-diffusion terms vanish
    (CentralUpwindHJ3D.py:99-188 with a_m = 0 or a_p = 0).  Exchanging the
    corner weights of an axis (H_+ with a_p instead of a_m) gives -H at the
    downwind corner instead, which this fails."""
    beta = SpectralGaussianField((n,) * d, 1.5, 0.01, seed=41)
    sig = O.sigma(beta, 1.0, 0.3)
    sig[O.cidx(0, 1)] += 0.004
    sig[O.cidx(1, 0)] += 0.004
    sig[O.cidx(d - 1, d - 1)] -= 0.003
    rhs, _ = O.central_upwind_flux(beta, sig, dyn, lam)
    ders = [O.find_derivatives(beta, a) for a in range(d)]
    corners = [tuple(c) for c in np.ndindex(*(2,) * d)]  # 0: '+', 1: '-'
    V = {c: O._velocity(dyn, [ders[a][c[a]] for a in range(d)], sig, lam) for c in corners}
    H = {c: O._hamiltonian(dyn, [ders[a][c[a]] for a in range(d)], V[c], lam) for c in corners}
    vmax = [np.max([V[c][a] for c in corners], axis=0) for a in range(d)]
    vmin = [np.min([V[c][a] for c in corners], axis=0) for a in range(d)]
    # cells with one-signed speeds on every axis, well above ME
    floor = 1e-6 * max(np.abs(v).max() for v in vmax + vmin)
    ok = np.ones(beta.shape[1:], bool)
    up = []
    for a in range(d):
        pos, neg = vmin[a] >= floor, vmax[a] <= -floor
        ok &= pos | neg
        up.append(np.where(pos, 1, 0))  # positive speed: u^- corner
    assert ok.sum() >= 0.2 * ok.size, ok.sum()
    want = np.zeros_like(rhs)
    for c in corners:
        sel = ok.copy()
        for a in range(d):
            sel &= up[a] == c[a]
        want[:, sel] = -H[c][:, sel]
    scale = np.abs(want[:, ok]).max()
    err = np.abs(rhs[:, ok] - want[:, ok]).max()
    assert err <= 1e-8 * scale, (err, scale)
    # and the downwind corner (a weight exchange on any axis) is far off
    for a in range(d):
        wrong = np.zeros_like(rhs)
        for c in corners:
            sel = ok.copy()
            for b in range(d):
                sel &= (up[b] == c[b]) if b != a else (up[b] != c[b])
            wrong[:, sel] = -H[c][:, sel]
        assert np.abs(rhs[:, ok] - wrong[:, ok]).max() > 1e-3 * scale
