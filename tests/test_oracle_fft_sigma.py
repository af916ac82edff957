"""Pins of the oracle's FFT, k-grid, stress and Nye-tensor functions.

Each check uses a closed form, an identity or brute force, chosen so that a
dropped term, a wrong sign or a transposed index fails.
"""
import numpy as np
import pytest

from oracle import cdd_oracle as O


def brute_rfftn(a):
    """O(N^2) DFT over every axis, keeping the half spectrum of the last."""
    out = a.astype(complex)
    for ax in range(a.ndim):
        n = a.shape[ax]
        idx = np.arange(n)
        W = np.exp(-2j * np.pi * np.outer(idx, idx) / n)
        out = np.moveaxis(np.tensordot(W, np.moveaxis(out, ax, 0), axes=(1, 0)), 0, ax)
    sl = [slice(None)] * a.ndim
    sl[-1] = slice(0, a.shape[-1] // 2 + 1)
    return out[tuple(sl)]


@pytest.mark.parametrize("shape", [(8, 8), (4, 4, 4), (6, 8), (4, 6, 8)])
def test_numpy_rfftn_matches_brute_force_dft(shape):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(shape)
    ref = brute_rfftn(a)
    got = O.rfftn(a)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    back = O.irfftn(got, shape)
    assert np.max(np.abs(back - a)) <= 1e-15 * 10
    # Parseval over the full spectrum
    full = np.fft.fftn(a)
    assert abs(np.sum(np.abs(full) ** 2) / a.size - np.sum(a * a)) < 1e-12 * np.sum(a * a)


@pytest.mark.parametrize("shape", [(8, 8), (16, 16), (8, 8, 8)])
def test_kgrid_closed_form(shape):
    K = O.ktools(shape)
    d = len(shape)
    n = shape[0]
    for a in range(d):
        m = K.kmask.shape[a]
        idx = np.arange(m)
        wrapped = np.where(idx > n // 2, idx - n, idx)  # Nyquist stays positive
        expect = 2 * np.pi * wrapped
        sl = [None] * d
        sl[a] = slice(None)
        np.testing.assert_allclose(np.broadcast_to(expect[tuple(sl)], K.kmask.shape), K.k[a], rtol=1e-15, atol=1e-12)
    assert K.kmask.flat[0] == 0.0 and K.kmask.sum() == K.kmask.size - 1
    if d == 2:
        assert np.all(K.kz == 0)


def _grid(shape):
    return np.meshgrid(*[np.arange(s) / s for s in shape], indexing="ij")


@pytest.mark.parametrize("shape", [(16, 16), (8, 8, 8)])
def test_sigma_constant_betaP(shape):
    mu, nu = 1.0, 0.3
    lam = O.lame_lambda(mu, nu)
    B = np.array([[0.3, -0.1, 0.2], [0.05, -0.4, 0.7], [0.11, 0.13, 0.5]])
    beta = np.zeros((9,) + shape)
    for c, (i, j) in enumerate(O.COMPS):
        beta[c] = B[i, j]
    s = O.sigma(beta, mu, nu)
    expect = -(lam * np.trace(B) * np.eye(3) + mu * (B + B.T))
    for c, (i, j) in enumerate(O.COMPS):
        np.testing.assert_allclose(s[c], expect[i, j], atol=1e-14)


@pytest.mark.parametrize("shape", [(16, 16), (8, 8, 8)])
def test_sigma_compatible_is_zero(shape):
    """beta_mj = d_m u_j for a smooth periodic u gives sigma = 0."""
    d = len(shape)
    from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField
    u = list(SpectralGaussianField(shape, 3.0, 1.0, seed=3)[:3])  # band-limited
    K = O.ktools(shape)
    beta = np.zeros((9,) + shape)
    for m in range(d):
        for j in range(3):
            beta[O.cidx(m, j)] = O.irfftn(1j * K.k[m] * O.rfftn(u[j]), shape)
    s = O.sigma(beta, 1.0, 0.3)
    assert np.max(np.abs(s)) < 1e-9 * np.max(np.abs(beta))


@pytest.mark.parametrize("shape", [(16, 16), (8, 8, 8)])
def test_sigma_antisymmetric_is_zero(shape):
    rng = np.random.default_rng(4)
    beta = np.zeros((9,) + shape)
    for i, j in [(0, 1), (0, 2), (1, 2)]:
        w = rng.standard_normal(shape)
        beta[O.cidx(i, j)] = w
        beta[O.cidx(j, i)] = -w
    s = O.sigma(beta, 1.0, 0.3)
    assert np.max(np.abs(s)) < 1e-13


@pytest.mark.parametrize("shape", [(16, 16), (8, 8, 8)])
def test_sigma_single_mode(shape):
    """beta_zy = a cos 2 pi x: the Green's tensor vanishes (k along x has no
    y or z part), so betaE = -betaP and sigma_yz = sigma_zy = -mu a cos 2 pi x."""
    mu, a = 1.0, 0.02
    g = _grid(shape)
    beta = np.zeros((9,) + shape)
    beta[O.cidx(2, 1)] = a * np.cos(2 * np.pi * g[0])
    s = O.sigma(beta, mu, 0.3)
    for c, (i, j) in enumerate(O.COMPS):
        expect = -mu * a * np.cos(2 * np.pi * g[0]) if (i, j) in [(1, 2), (2, 1)] else 0.0
        np.testing.assert_allclose(s[c], expect, atol=1e-15)


@pytest.mark.parametrize("shape", [(16, 16), (8, 8, 8)])
def test_sigma_symmetric_divergence_free_and_betaP_source(shape):
    from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField
    beta = SpectralGaussianField(shape, 3.0, 0.01, seed=5)
    mu, nu = 1.0, 0.3
    s = O.sigma(beta, mu, nu)
    for i in range(3):
        for j in range(3):
            np.testing.assert_allclose(s[O.cidx(i, j)], s[O.cidx(j, i)], atol=1e-17)
    K = O.ktools(shape)
    d = len(shape)
    scale = np.max(np.abs(s))
    for i in range(3):
        div = sum(1j * K.k[j] * O.rfftn(s[O.cidx(i, j)]) for j in range(d))
        div = O.irfftn(div, shape)
        assert np.max(np.abs(div)) < 1e-10 * scale * shape[0]
    s2 = O.sigma_betaP_source(beta, mu, nu)
    assert np.max(np.abs(s - s2)) < 1e-13 * scale


@pytest.mark.parametrize("shape", [(16, 16), (8, 8, 8)])
def test_rho_closed_forms(shape):
    d = len(shape)
    g = _grid(shape)
    K = O.ktools(shape)
    from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField
    # compatible -> rho = 0
    u = list(SpectralGaussianField(shape, 3.0, 1.0, seed=7)[:3])
    beta = np.zeros((9,) + shape)
    for m in range(d):
        for j in range(3):
            beta[O.cidx(m, j)] = O.irfftn(1j * K.k[m] * O.rfftn(u[j]), shape)
    assert np.max(np.abs(O.rho_fourier(beta))) < 1e-10 * np.max(np.abs(beta))
    # single mode: beta_yz = a sin 2 pi x -> rho_zz = -2 pi a cos 2 pi x, rest 0
    a = 0.03
    beta = np.zeros((9,) + shape)
    beta[O.cidx(1, 2)] = a * np.sin(2 * np.pi * g[0])
    r = O.rho_fourier(beta)
    for c, (i, j) in enumerate(O.COMPS):
        expect = -2 * np.pi * a * np.cos(2 * np.pi * g[0]) if (i, j) == (2, 2) else 0.0
        np.testing.assert_allclose(r[c], expect, atol=1e-13)
    # d_i rho_ij = 0 for a random field
    from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField
    beta = SpectralGaussianField(shape, 3.0, 0.01, seed=8)
    r = O.rho_fourier(beta)
    for j in range(3):
        div = O.irfftn(sum(1j * K.k[i] * O.rfftn(r[O.cidx(i, j)]) for i in range(d)), shape)
        assert np.max(np.abs(div)) < 1e-10 * np.max(np.abs(r)) * shape[0]


def test_energy_closed_forms():
    mu, nu = 1.0, 0.3
    s = 0.01
    E = 2 * mu * (1 + nu)
    sig = np.zeros((9, 2, 2))
    sig[8] = s
    np.testing.assert_allclose(O.elastic_energy_density(sig, mu, nu), 0.5 * s * s / E, rtol=1e-14)
    sig = np.zeros((9, 2, 2))
    sig[1] = sig[3] = s
    np.testing.assert_allclose(O.elastic_energy_density(sig, mu, nu), s * s / (2 * mu), rtol=1e-14)


def test_loads_equal_stress_of_uniform_distortion():
    """External strain eps acts like a uniform plastic distortion -eps."""
    mu, nu = 1.0, 0.3
    shape = (8, 8)
    beta = np.zeros((9,) + shape)
    eps = np.array([0.01, -0.02, 0.005])
    load = O.Load("strain", initial=eps, rate=(0.001, 0.0, 0.002), t0=0.5)
    t = 2.0
    got = load.apply(np.zeros_like(beta), beta, t, mu, nu)
    e_t = eps + np.array([0.001, 0.0, 0.002]) * (t - 0.5)
    B = np.zeros((9,) + shape)
    for i in range(3):
        B[O.cidx(i, i)] = -e_t[i]
    np.testing.assert_allclose(got, O.sigma(B, mu, nu), atol=1e-15)
    # a strain load with its own constants (RunPython's ExternalStrain at
    # mu = 0.5, nu = 0.3) on a mu = 1 state: the stress of the uniform
    # distortion in a mu = 0.5 medium, whatever the state's mu
    own = O.Load("strain", initial=eps, rate=(0.001, 0.0, 0.002), t0=0.5, elastic=(0.5, 0.3))
    got = own.apply(np.zeros_like(beta), beta, t, mu, nu)
    np.testing.assert_allclose(got, O.sigma(B, 0.5, 0.3), atol=1e-15)
    assert np.abs(got - O.sigma(B, mu, nu)).max() > 1e-3
    # uniaxial strain-rate control: beta_E_zz = rate t - (<beta_zz> - init)
    beta = np.zeros((9,) + shape)
    beta[8] = 0.003
    load = O.Load("uniaxial", init_strain=0.001, uniaxial_rate=0.01)
    got = load.apply(np.zeros_like(beta), beta, 1.5, mu, nu)
    bzz = 0.01 * 1.5 - (0.003 - 0.001)
    B = np.zeros((9,) + shape)
    B[8] = -bzz
    np.testing.assert_allclose(got, O.sigma(B, mu, nu), atol=1e-15)


@pytest.mark.parametrize("kind", ["none", "strain", "uniaxial"])
def test_mean_stress_is_elastic_response_to_mean_distortion(kind):
    """Fixed-volume boundary: the mean elastic distortion is -<beta^P>, so
    the mean stress is -(lam tr<beta> d + 2 mu sym<beta>) plus the load
    shift, whatever the fluctuations (PlottingTools.py:69-100)."""
    from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField
    mu, nu, t = 0.5, 0.3, 1.7
    lam = 2 * mu * nu / (1 - 2 * nu)
    beta = SpectralGaussianField((8, 8, 8), 2.0, 0.01, seed=11)
    beta += np.arange(9).reshape(9, 1, 1, 1) * 1e-3  # nonzero means
    load = {"none": O.Load(),
            "strain": O.Load("strain", initial=(0.0, 0.0, 0.0), rate=(-0.005, -0.005, 0.01)),
            "uniaxial": O.Load("uniaxial", init_strain=0.001, uniaxial_rate=0.01)}[kind]
    got = O.mean_stress(beta, t, mu, nu, load)
    mb = beta.reshape(9, -1).mean(axis=1).reshape(3, 3)
    eps = 0.5 * (mb + mb.T)
    sig = -(lam * np.trace(mb) * np.eye(3) + 2 * mu * eps)
    if kind == "strain":  # StrainStress: + lam tr(e) d + 2 mu e of the imposed strain
        e = np.diag(np.array([-0.005, -0.005, 0.01]) * t)
        sig = sig + lam * np.trace(e) * np.eye(3) + 2 * mu * e
    elif kind == "uniaxial":  # RunLoadingMPI: (lam d + 2 mu d_iz d_jz) betaE_zz
        bzz = 0.01 * t - (mb[2, 2] - 0.001)
        sig = sig + bzz * (lam * np.eye(3) + 2 * mu * np.diag([0.0, 0.0, 1.0]))
    want = np.array([sig[0, 0], sig[1, 1], sig[2, 2], sig[0, 1], sig[0, 2], sig[1, 2]])
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-15)


@pytest.mark.parametrize("shape", [(16, 16), (8, 8, 8)])
def test_instability_measure_is_alternating_sum(shape):
    """The (N/2, ..., N/2) Fourier mode of |rho| equals the sum over cells of
    (-1)^(i+j[+k]) |rho| (the exponent e^{-i pi (i+j+k)})."""
    from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField
    beta = SpectralGaussianField(shape, 1.5, 0.01, seed=21)
    r = O.rho_fourier(beta)
    rhoM = np.sqrt((r * r).sum(axis=0))
    sign = np.ones(shape)
    for a, n in enumerate(shape):
        idx = [1] * len(shape)
        idx[a] = n
        sign = sign * ((-1.0) ** np.arange(n)).reshape(idx)
    want = abs((sign * rhoM).sum())
    assert abs(O.instability(beta) - want) <= 1e-12 * want
