"""The state's other stress and Nye-tensor forms on the GPU against the
oracle: CalculateSigma(type in {'k', 'sinK'}, source in {'betaE', 'betaP',
'rho'}), CalculateRhoFourier('sinK') and CalculateRhoSymmetric
(PlasticityState.py:85-304), through the C ABI and through the reference
API's PlasticityState."""
import sys

import numpy as np
import pytest

from oracle import cdd_oracle as O
from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


@pytest.fixture(scope="module")
def L():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from Plasticity.amd import lib
    return lib


def _field(d, n, seed):
    # a rough field (L = 1 cell) puts weight on the Nyquist planes, where the
    # 'rho' source's real-space round trip matters
    return SpectralGaussianField((n,) * d, 1.0, 0.01, seed=seed)


@pytest.mark.parametrize("d,n", [(2, 32), (2, 64), (3, 16), (3, 32)])
@pytest.mark.parametrize("type", ["k", "sinK"])
def test_sigma_variants_match_oracle(L, d, n, type):
    say("sigma variants", d, n, type)
    mu, nu = 0.8, 0.28
    beta = _field(d, n, 31)
    p = L.Plan(d, n, mu=mu, nu=nu)
    want = {"betaE": O.sigma(beta, mu, nu), "betaP": O.sigma_betaP_source(beta, mu, nu, type),
            "rho": O.sigma_rho_source(beta, mu, nu, type)}
    for source, w in want.items():
        got = p.sigma_variant(beta, type, source).cpu().numpy()
        e = rel(got, w)
        say("  %s %s: rel L2 %.2e" % (type, source, e))
        assert e < 1e-12
        for i in range(3):
            for j in range(3):
                assert np.array_equal(got[3 * i + j], got[3 * j + i])


@pytest.mark.parametrize("d,n", [(2, 64), (3, 32)])
def test_rho_variants_match_oracle(L, d, n):
    say("rho variants", d, n)
    beta = _field(d, n, 32)
    p = L.Plan(d, n)
    for type in ("k", "sinK"):
        got = p.rho_variant(beta, type).cpu().numpy()
        e = rel(got, O.rho_fourier(beta, type))
        say("  rho %s: rel L2 %.2e" % (type, e))
        assert e < 1e-12
    got = p.rho_symmetric(beta).cpu().numpy()
    want = O.rho_symmetric(beta)
    e = rel(got, want)
    say("  rho symmetric: rel L2 %.2e" % e)
    assert e < 1e-14
    # central differences = the sinktools' spectral derivative
    assert rel(got, p.rho_variant(beta, "sinK").cpu().numpy()) < 1e-12


def test_reference_api_state_forms(L):
    """PlasticityState.CalculateSigma / CalculateRhoFourier /
    CalculateRhoSymmetric with the reference's arguments."""
    from Plasticity.PlasticityStates.PlasticityState import PlasticityState
    d, n = 3, 16
    beta = _field(d, n, 33)
    st = PlasticityState((n,) * d, field=beta, mu=1.0, nu=0.3)
    for type in ("k", "sinK"):
        s = st.CalculateSigma(type=type, source="rho").data.cpu().numpy()
        assert rel(s, O.sigma_rho_source(beta, 1.0, 0.3, type)) < 1e-12
        s = st.CalculateSigma(type=type, source="betaE").data.cpu().numpy()
        assert rel(s, O.sigma(beta, 1.0, 0.3)) < 1e-12  # the reference ignores type for betaE
        s = st.CalculateSigma(type=type, source="betaP").data.cpu().numpy()
        assert rel(s, O.sigma_betaP_source(beta, 1.0, 0.3, type)) < 1e-12
        r = st.CalculateRhoFourier(type).data.cpu().numpy()
        assert rel(r, O.rho_fourier(beta, type)) < 1e-12
    r = st.CalculateRhoSymmetric().data.cpu().numpy()
    assert rel(r, O.rho_symmetric(beta)) < 1e-14
    with pytest.raises(ValueError):
        st.CalculateSigma(type="cosK")
    with pytest.raises(L.CDDError):
        L.Plan(d, n, precision="fp32").sigma_variant(beta, "k", "rho")
