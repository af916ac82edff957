"""GPU parity of the CUDA executable's numerics (the LLF flux of
plasticity_kernel.cu centralHJ, simpleTVD forward Euler and TVD3rd) through
the C ABI against the oracle's restatement (cuda_llf_flux, run_cuda)."""
import sys

import numpy as np
import pytest

from oracle import cdd_oracle as O
from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField

pytestmark = pytest.mark.gpu

# plasticity.h: mu 0.5, nu 0.3; RunCUDA.py header_size: CFLsafeFactor 0.5
MU, NU, CFL = 0.5, 0.3, 0.5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


@pytest.fixture(scope="module")
def amd():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from Plasticity import amd as A
    return A


def _plan(amd, d, n, dyn, Lambda, rk=1, **kw):
    return amd.Plan(d, n, dynamics=dyn, rk_order=rk, Lambda=Lambda, mu=MU, nu=NU, cfl=CFL, scheme="llf", **kw)


def _odyn(dyn, Lambda, load=None):
    return O.Dynamics(O.GLIDE_ONLY if dyn == 1 else O.GLIDE_CLIMB, Lambda, MU, NU, load=load)


# gcd (lambda 0), mdp (lambda 1), a partial climb, lvp (NEWGLIDEONLY)
DYNS = [(0, 0.0), (0, 1.0), (0, 0.4), (1, 0.0)]


@pytest.mark.parametrize("dyn,Lambda", DYNS)
@pytest.mark.parametrize("d,n", [(2, 8), (2, 16), (2, 64), (3, 8), (3, 16), (3, 32), (3, 64)])
def test_llf_flux_matches_oracle(amd, d, n, dyn, Lambda):
    """3D from 32^3 up: the plane-marching kernel's LLF instances."""
    say("llf flux", d, n, dyn, Lambda)
    beta = SpectralGaussianField((n,) * d, 2.0, 0.01, seed=31)
    p = _plan(amd, d, n, dyn, Lambda)
    rhs, speed = p.flux(beta)
    ref_rhs, ref_speed = O.cuda_llf_flux(beta, O.sigma(beta, MU, NU), 0.0 if dyn == 1 else Lambda, dyn == 1)
    assert rel_l2(rhs.cpu().numpy(), ref_rhs) < 1e-12
    assert rel_l2(speed.cpu().numpy(), ref_speed) < 1e-13


@pytest.mark.parametrize("rk", [1, 3])
@pytest.mark.parametrize("d,n,dyn,Lambda", [(2, 32, 1, 0.0), (2, 32, 0, 1.0), (3, 16, 0, 0.0), (3, 16, 1, 0.0),
                                            (3, 16, 0, 1.0), (3, 32, 0, 0.4), (3, 32, 1, 0.0)])
def test_llf_steps_match_oracle(amd, d, n, dyn, Lambda, rk):
    nsteps = 10
    say("llf steps", d, n, dyn, Lambda, rk)
    beta = SpectralGaussianField((n,) * d, 2.0, 0.01, seed=32)
    p = _plan(amd, d, n, dyn, Lambda, rk=rk)
    p.set_state(beta)
    p.set_time(0.0)
    p.step(nsteps)
    got = p.get_state().cpu().numpy()
    t, dt, done, ns = p.clock()
    ref, tref, dts = O.run_cuda(_odyn(dyn, Lambda), beta, 0.0, nsteps, CFL, order=rk)
    assert ns == nsteps
    assert abs(t - tref) <= 1e-12 * tref
    assert abs(dt - dts[-1]) <= 1e-12 * dts[-1]
    assert rel_l2(got, ref) < 1e-11


@pytest.mark.parametrize("rk", [1, 3])
def test_llf_end_step_lands_me_past_the_end(amd, rk):
    """simpleTVD's end clamp dt = t_end - t + ME (plasticity_kernel.cu:1124-1125):
    the device stops after the step that passes t_end, at t_end + ME."""
    say("llf end time", rk)
    n = 16
    beta = SpectralGaussianField((n, n), 2.0, 0.01, seed=33)
    od = _odyn(0, 1.0)
    _, _, dts = O.run_cuda(od, beta, 0.0, 1, CFL, order=rk)
    t_end = 2.5 * dts[0]
    p = _plan(amd, 2, n, 0, 1.0, rk=rk)
    p.set_state(beta)
    p.set_time(0.0, t_end)
    p.step(20)
    t, dt, done, ns = p.clock()
    ref, tref, dref = O.run_cuda(od, beta, 0.0, 20, CFL, t_end=t_end, order=rk)
    assert done and ns == len(dref) >= 2
    assert abs(t - tref) <= 1e-15 and t > t_end
    assert abs(dref[-1] - (t_end - (dref[0] + dref[1]) + O.ME)) <= 1e-18
    assert rel_l2(p.get_state().cpu().numpy(), ref) < 1e-12


def test_llf_loaded_steps_match_oracle(amd):
    """The executable's loadSigma (plasticity_kernel.cu:1045-1066): sigma_ij +=
    (2 mu L_ij + lambda tr(L) d_ij)(LOAD_START + LOADING_RATE t), a linear
    stress-tensor load."""
    say("llf loaded")
    n, nsteps = 16, 6
    Ld = np.array([[0.0, 0.5, 0.0], [0.5, 0.0, 0.0], [0.0, 0.0, -0.2]])
    rate, start = 0.05, 0.01
    lamE = 2 * MU * NU / (1 - 2 * NU)
    S = 2 * MU * Ld + lamE * np.trace(Ld) * np.eye(3)
    beta = SpectralGaussianField((n,) * 3, 2.0, 0.01, seed=34)
    p = _plan(amd, 3, n, 0, 1.0, rk=1, load_kind=4, load_init=(S * start).ravel(), load_rate=(S * rate).ravel())
    p.set_state(beta)
    p.set_time(0.3)
    p.step(nsteps)
    load = O.Load(kind="tensor", tensor0=tuple((S * start).ravel()), tensor_rate=tuple((S * rate).ravel()))
    ref, tref, dts = O.run_cuda(_odyn(0, 1.0, load), beta, 0.3, nsteps, CFL)
    assert rel_l2(p.get_state().cpu().numpy(), ref) < 1e-11


def test_llf_refusals(amd):
    from Plasticity.amd.lib import CDDError
    say("llf refusals")
    with pytest.raises(CDDError):
        _plan(amd, 2, 16, 0, 0.0, rk=2)
    with pytest.raises(CDDError):
        _plan(amd, 2, 16, 0, 0.0, precision="fp32")
    with pytest.raises(CDDError):
        _plan(amd, 2, 16, 2, 0.5)  # woJzz
    with pytest.raises(CDDError):
        amd.Plan(2, 16, rk_order=1)  # Euler is the LLF scheme's integrator
    p = _plan(amd, 2, 16, 0, 0.0, rk=3)
    with pytest.raises(CDDError):
        p.set_vacancy(1.0, 1.0, 1.0)
    b = SpectralGaussianField((16, 16), 2.0, 0.01, seed=35)
    with pytest.raises(CDDError):
        p.flux_vjp(b, 0.0, b)
