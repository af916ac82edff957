"""GPU parity of the disorder coupling (DisorderState.CalculateSigma,
DisorderState.py:15-30; BetaP_DisorderDynamics, DisorderDynamics.py:17-44;
the advected ('s', 's') field of RunPython.py's 'x' runs) through the C ABI
against the oracle's disorder_sigma / DisorderDynamics."""
import sys

import numpy as np
import pytest

from oracle import cdd_oracle as O
from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField

pytestmark = pytest.mark.gpu

MU, NU = 0.5, 0.3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / np.linalg.norm(np.asarray(b).ravel())


@pytest.fixture(scope="module")
def amd():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from Plasticity import amd as A
    return A


def _fields(d, n, seed):
    beta = SpectralGaussianField((n,) * d, 2.0, 0.01, seed=seed)
    dis = SpectralGaussianField((n,) * d, 3.0, 0.05, seed=seed + 100)[0]
    return beta, dis


def _plan(amd, d, n, dyn=0, Lambda=0.0, rk=3):
    return amd.Plan(d, n, dynamics=dyn, rk_order=rk, Lambda=Lambda, mu=MU, nu=NU, cfl=0.25)


def _odyn(dyn, Lambda, disorder=None):
    return O.DisorderDynamics(O.GLIDE_ONLY if dyn == 1 else O.GLIDE_CLIMB, Lambda, MU, NU, disorder=disorder)


@pytest.mark.parametrize("d,n", [(2, 16), (2, 64), (3, 8), (3, 16), (3, 32)])
def test_disorder_sigma_matches_oracle(amd, d, n):
    say("disorder sigma", d, n)
    beta, dis = _fields(d, n, 51)
    p = _plan(amd, d, n)
    got = p.disorder_sigma(beta, dis).cpu().numpy()
    ref = O.disorder_sigma(beta, dis, MU, NU)
    assert rel_l2(got, ref) < 1e-12
    dsig = ref - O.sigma(beta, MU, NU)
    assert rel_l2(got - O.sigma(beta, MU, NU), dsig) < 1e-9  # the disorder term itself


@pytest.mark.parametrize("dyn,Lambda", [(0, 0.0), (0, 0.8), (1, 0.0)])
@pytest.mark.parametrize("d,n", [(2, 16), (2, 32), (3, 8), (3, 16)])
def test_disorder_flux_matches_oracle(amd, d, n, dyn, Lambda):
    say("disorder flux", d, n, dyn, Lambda)
    beta, dis = _fields(d, n, 52)
    p = _plan(amd, d, n, dyn, Lambda)
    rhs, rd, speed = p.disorder_flux(beta, dis)
    u = np.concatenate([beta, dis[None]])
    ref, sp = _odyn(dyn, Lambda).flux(0.0, u)
    assert rel_l2(rhs.cpu().numpy(), ref[:9]) < 1e-12
    assert rel_l2(rd.cpu().numpy(), ref[9]) < 1e-12
    assert rel_l2(speed.cpu().numpy(), sp) < 1e-12
    # the fixed field: same rhs of beta, no scalar
    rhs2, rd2, _ = p.disorder_flux(beta, dis, advect=False)
    ref2, _ = _odyn(dyn, Lambda, disorder=dis).flux(0.0, beta)
    assert rd2 is None and rel_l2(rhs2.cpu().numpy(), ref2) < 1e-12


@pytest.mark.parametrize("rk", [2, 3])
@pytest.mark.parametrize("d,n,dyn,Lambda", [(2, 32, 0, 0.8), (2, 32, 1, 0.0), (3, 16, 0, 0.8), (3, 16, 1, 0.0)])
def test_disorder_steps_match_oracle(amd, d, n, dyn, Lambda, rk):
    nsteps = 10
    say("disorder steps", d, n, dyn, Lambda, rk)
    beta, dis = _fields(d, n, 53)
    p = _plan(amd, d, n, dyn, Lambda, rk=rk)
    p.set_state(beta)
    p.set_disorder(dis, advect=True)
    p.set_time(0.0)
    p.step(nsteps)
    t, dt, done, ns = p.clock()
    u, tref, dts = O.run(_odyn(dyn, Lambda), np.concatenate([beta, dis[None]]), 0.0, nsteps, 0.25, order=rk)
    assert ns == nsteps and abs(t - tref) <= 1e-12 * tref
    assert rel_l2(p.get_state().cpu().numpy(), u[:9]) < 1e-11
    assert rel_l2(p.disorder_get().cpu().numpy(), u[9]) < 1e-11
    assert rel_l2(u[9], dis) > 1e-6  # the field moved


@pytest.mark.parametrize("d,n", [(2, 32), (3, 16)])
def test_fixed_disorder_steps_match_oracle(amd, d, n):
    nsteps = 8
    say("fixed disorder steps", d, n)
    beta, dis = _fields(d, n, 54)
    p = _plan(amd, d, n, 0, 0.5, rk=3)
    p.set_state(beta)
    p.set_disorder(dis, advect=False)
    p.set_time(0.0)
    p.step(nsteps)
    u, tref, dts = O.run(_odyn(0, 0.5, disorder=dis), beta, 0.0, nsteps, 0.25, order=3)
    assert rel_l2(p.get_state().cpu().numpy(), u) < 1e-11
    assert np.array_equal(p.disorder_get().cpu().numpy(), dis)


def test_zero_disorder_is_the_plain_run(amd):
    say("zero disorder")
    n = 32
    beta, _ = _fields(3, n, 55)
    a = _plan(amd, 3, n, 0, 0.9)
    a.set_state(beta)
    a.set_time(0.0)
    a.step(4)
    b = _plan(amd, 3, n, 0, 0.9)
    b.set_state(beta)
    b.set_disorder(np.zeros((n,) * 3), advect=True)
    b.set_time(0.0)
    b.step(4)
    assert rel_l2(b.get_state().cpu().numpy(), a.get_state().cpu().numpy()) < 1e-13
    assert not b.disorder_get().abs().max().item()


def test_disorder_refusals(amd):
    from Plasticity.amd.lib import CDDError
    say("disorder refusals")
    n = 16
    beta, dis = _fields(2, n, 56)
    p = _plan(amd, 2, n)
    p.set_disorder(dis)
    with pytest.raises(CDDError):
        p.set_vacancy(1.0, 1.0, 1.0)
    with pytest.raises(CDDError):
        p.flux_vjp(beta, 0.0, beta)
    q = amd.Plan(2, n, precision="fp32")
    with pytest.raises(CDDError):
        q.set_disorder(dis)
    r = amd.Plan(2, n, rk_order=1, scheme="llf")
    with pytest.raises(CDDError):
        r.set_disorder(dis)


def test_reference_api_disorder_state(amd):
    """DisorderState + BetaPDynamics(Lambda=1) + TVDRungeKutta_FieldMover
    through PlasticitySystem (the RunPython.py:196-257 'x' runs): the state's
    stress, the ten-component flux and the run against the oracle."""
    say("disorder state API")
    from Plasticity import PlasticitySystem
    from Plasticity.FieldDynamics import CentralUpwindHJBetaPDynamics
    from Plasticity.FieldMovers import FieldMover
    from Plasticity.Observers import Observer
    from Plasticity.PlasticityStates.DisorderState import DisorderState
    n = 32
    beta, dis = _fields(2, n, 57)
    state = DisorderState((n, n), field=np.concatenate([dis[None], beta]), mu=MU, nu=NU)
    assert rel_l2(state.CalculateSigma().numpy(), O.disorder_sigma(beta, dis, MU, NU)) < 1e-12
    dyn = CentralUpwindHJBetaPDynamics.BetaPDynamics(Lambda=1)
    rhs, speed = dyn.CalculateFlux(0.0, state, CFLCondition=True)
    u = np.concatenate([beta, dis[None]])
    ref, sp = _odyn(0, 1.0).flux(0.0, u)
    got = rhs.numpy()
    assert got.shape == (10, n, n)
    assert rel_l2(got[1:], ref[:9]) < 1e-12 and rel_l2(got[0], ref[9]) < 1e-12
    assert rel_l2(speed.cpu().numpy(), sp) < 1e-12
    rec = Observer.RecallStateObserver()
    mover = FieldMover.TVDRungeKutta_FieldMover(CFLsafeFactor=0.25)
    _, _, dts = O.run(_odyn(0, 1.0), u, 0.0, 5, 0.25)
    t_end = float(np.sum(dts[:4]) + 0.5 * dts[4])
    PlasticitySystem.PlasticitySystem((n, n), state, mover, dyn, [rec]).Run(startTime=0.0, endTime=t_end)
    ref_u, tref, dref = O.run(_odyn(0, 1.0), u, 0.0, 100, 0.25, t_end=t_end)
    assert mover.numberOfSteps == len(dref)
    g = rec.state.GetOrderParameterField().numpy()
    assert rel_l2(g[1:], ref_u[:9]) < 1e-11 and rel_l2(g[0], ref_u[9]) < 1e-11
    mover2 = FieldMover.TVDRungeKutta_FieldMover(CFLsafeFactor=0.25)
    mover2.Run(state, [], dyn, 0.0, t_end)
    g2 = mover2.lastState.GetOrderParameterField().numpy()
    assert rel_l2(g2[1:], ref_u[:9]) < 1e-11 and rel_l2(g2[0], ref_u[9]) < 1e-11


def test_reference_api_fixed_disorder_dynamics(amd):
    """BetaP_DisorderDynamics(disorder=d) on a PlasticityState
    (DisorderDynamics.py:17-44): GetSigma, CalculateFlux and a mover run."""
    say("fixed disorder API")
    from Plasticity.FieldDynamics.DisorderDynamics import BetaP_DisorderDynamics
    from Plasticity.FieldMovers import FieldMover
    from Plasticity.PlasticityStates.PlasticityState import PlasticityState
    n = 16
    beta, dis = _fields(3, n, 58)
    state = PlasticityState((n,) * 3, field=beta, mu=MU, nu=NU)
    dyn = BetaP_DisorderDynamics(Lambda=0.5, disorder=dis)
    assert rel_l2(dyn.GetSigma(state, 0.0).numpy(), O.disorder_sigma(beta, dis, MU, NU)) < 1e-12
    od = _odyn(0, 0.5, disorder=dis)
    ref, _ = od.flux(0.0, beta)
    assert rel_l2(dyn.CalculateFlux(0.0, state).numpy(), ref) < 1e-12
    _, _, dts = O.run(od, beta, 0.0, 3, 0.25)
    t_end = float(np.sum(dts[:2]) + 0.5 * dts[2])
    mover = FieldMover.TVDRungeKutta_FieldMover(CFLsafeFactor=0.25)
    mover.Run(state, [], dyn, 0.0, t_end)
    ref_u, _, dref = O.run(od, beta, 0.0, 100, 0.25, t_end=t_end)
    assert mover.numberOfSteps == len(dref)
    assert rel_l2(mover.lastState.betaP.numpy(), ref_u) < 1e-11
