"""Pins of the oracle's other stress and Nye-tensor forms: the sinktools
(k = N sin(2 pi n / N)), CalculateRhoFourier(type='sinK'),
CalculateRhoSymmetric and CalculateSigma(source='rho' and 'betaP', both
types) -- PlasticityState.py:85-304, FourierSpaceTools.py:13-61,
NumericalMethods.py:6-13."""
import numpy as np
import pytest

from oracle import cdd_oracle as O


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


def band_limited(shape, seed, kmax=3):
    """A random real field made of modes with |n_a| <= kmax on every axis:
    nothing at the Nyquist planes, where i k of a real field is not
    Hermitian and numpy's irfftn keeps only part of it."""
    rng = np.random.default_rng(seed)
    d = len(shape)
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    out = np.zeros((9,) + tuple(shape))
    for c in range(9):
        out[c] = rng.normal() * 1e-3
        for _ in range(6):
            m = rng.integers(-kmax, kmax + 1, size=d)
            ph = sum(2 * np.pi * m[a] * grids[a] / shape[a] for a in range(d))
            out[c] += 0.01 * (rng.normal() * np.cos(ph) + rng.normal() * np.sin(ph))
    return out


@pytest.mark.parametrize("shape", [(16, 16), (8, 8, 8), (12, 12, 12)])
def test_sinktools_closed_form(shape):
    K = O.ktools(shape, "sinK")
    Kk = O.ktools(shape, "k")
    N = shape[-1]
    d = len(shape)
    half = (shape[:-1] + (shape[-1] // 2 + 1,))
    idx = np.meshgrid(*[np.arange(h) for h in half], indexing="ij")
    for a in range(d):
        scale = N if d == 2 else shape[a]
        want = scale * np.sin(2 * np.pi * idx[a] / shape[a])
        np.testing.assert_array_equal(K.k[a], want)
    np.testing.assert_array_equal(K.kmask, Kk.kmask)
    # N sin(k / N) = k (1 - (k / N)^2 / 6 + ...) at the smallest wavenumber
    th = 2 * np.pi / shape[0]
    k1, s1 = Kk.k[0][(1,) + (0,) * (d - 1)], K.k[0][(1,) + (0,) * (d - 1)]
    assert abs((k1 - s1) / k1 - th * th / 6) < th ** 4 / 100


@pytest.mark.parametrize("shape", [(32, 32), (16, 16, 16)])
def test_rho_symmetric_equals_sink_rho_fourier(shape):
    """The central difference of a Fourier mode is i N sin(2 pi n / N) times
    it, so CalculateRhoSymmetric is CalculateRhoFourier('sinK') for every
    field (Nyquist modes: both give zero)."""
    rng = np.random.default_rng(5)
    beta = rng.normal(size=(9,) + shape)
    a = O.rho_symmetric(beta)
    b = O.rho_fourier(beta, "sinK")
    assert rel(a, b) < 1e-13


@pytest.mark.parametrize("shape", [(16, 16), (8, 8, 8)])
def test_rho_symmetric_is_the_levi_civita_form(shape):
    rng = np.random.default_rng(6)
    beta = rng.normal(size=(9,) + shape)
    d = len(shape)
    N = shape[-1]
    D = [0.5 * N * (np.roll(beta, -1, axis=1 + a) - np.roll(beta, 1, axis=1 + a)) for a in range(d)]
    want = np.zeros_like(beta)
    for i in range(3):
        for j in range(3):
            for k in range(d):
                for l in range(3):
                    want[O.cidx(i, j)] -= O.PERM[i, k, l] * D[k][O.cidx(l, j)]
    np.testing.assert_allclose(O.rho_symmetric(beta), want, rtol=0, atol=1e-12)


@pytest.mark.parametrize("shape,axis", [((16, 16), 0), ((16, 16), 1), ((8, 8, 8), 2)])
def test_sink_rho_single_mode_factor(shape, axis):
    """A mode along one axis: the 'sinK' Nye tensor is the 'k' one times
    sin(theta) / theta, theta = 2 pi m / N."""
    m = 3
    d = len(shape)
    g = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    ph = 2 * np.pi * m * g[axis] / shape[axis]
    rng = np.random.default_rng(7)
    amp = rng.normal(size=9)
    beta = amp[:, None] * np.cos(ph).reshape(1, -1)
    beta = beta.reshape((9,) + shape)
    th = 2 * np.pi * m / shape[axis]
    a = O.rho_fourier(beta, "sinK")
    b = O.rho_fourier(beta, "k") * np.sin(th) / th
    assert rel(a, b) < 1e-13


@pytest.mark.parametrize("shape", [(32, 32), (12, 12, 12)])
def test_rho_source_is_betaE_source_plus_mean_terms(shape):
    """Free-boundary ('rho') stress = fixed-volume ('betaE') stress + the mean
    terms lam tr<beta> d_ij + mu (<beta_ij> + <beta_ji>) of the latter's
    k = 0 mode (PlasticityState.py:108-117), on band-limited fields."""
    mu, nu = 0.7, 0.27
    beta = band_limited(shape, 8)
    lam = O.lame_lambda(mu, nu)
    want = O.sigma(beta, mu, nu)
    m = beta.reshape(9, -1).mean(axis=1)
    for i, j in O.COMPS:
        want[O.cidx(i, j)] += lam * (m[0] + m[4] + m[8]) * (i == j) + mu * (m[O.cidx(i, j)] + m[O.cidx(j, i)])
    got = O.sigma_rho_source(beta, mu, nu, "k")
    assert rel(got, want) < 1e-12


@pytest.mark.parametrize("shape", [(32, 32), (12, 12, 12)])
def test_sink_rho_source_is_sink_betaP_source_without_its_mean_correction(shape):
    """K(k) N(k) = M(k) for every k vector, the sinktools' ones included: the
    two independent 'sinK' formulas agree up to the betaP source's mean
    correction (PlasticityState.py:179-182)."""
    mu, nu = 1.0, 0.3
    beta = band_limited(shape, 9)
    lam = O.lame_lambda(mu, nu)
    want = O.sigma_betaP_source(beta, mu, nu, "sinK")
    m = beta.reshape(9, -1).mean(axis=1)
    for i, j in O.COMPS:
        want[O.cidx(i, j)] += lam * (m[0] + m[4] + m[8]) * (i == j) + mu * (m[O.cidx(i, j)] + m[O.cidx(j, i)])
    got = O.sigma_rho_source(beta, mu, nu, "sinK")
    assert rel(got, want) < 1e-12
    # and the sinK stress differs from the k one (the sources do not ignore type)
    assert rel(O.sigma_rho_source(beta, mu, nu, "k"), got) > 1e-3


@pytest.mark.parametrize("shape,axis", [((16, 16), 1), ((8, 8, 8), 0)])
def test_sink_betaP_source_single_axis_mode_equals_k(shape, axis):
    """M(k) is homogeneous of degree 0 in k: for a mode along one axis the
    sin k vector is parallel to k and the 'sinK' stress is the 'k' one."""
    m = 2
    g = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    ph = 2 * np.pi * m * g[axis] / shape[axis]
    amp = np.random.default_rng(10).normal(size=9)
    beta = (amp[:, None] * np.sin(ph).reshape(1, -1)).reshape((9,) + shape)
    assert rel(O.sigma_betaP_source(beta, 1.0, 0.3, "sinK"), O.sigma_betaP_source(beta, 1.0, 0.3, "k")) < 1e-12
