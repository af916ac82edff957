"""Vacancy coupling on the GPU (cdd_set_vacancy: BetaP_VacancyDynamics with
the operator-splitting mover, FieldDynamics/VacancyDynamics.py:16-70,
FieldMovers/FieldMover.py:318-380) against the oracle's run_vacancy."""
import sys

import numpy as np
import pytest
import torch

from oracle import cdd_oracle as O
from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField

pytestmark = pytest.mark.gpu
MU, NU = 1.0, 0.3
LAM, ALPHA, VBETA, GAMMA = 0.5, 0.8, 1.0, 0.05


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / max(np.linalg.norm(np.asarray(b).ravel()), 1e-300)


@pytest.fixture(scope="module")
def L():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from Plasticity.amd import lib
    return lib


def _fields(shape, seed):
    beta = SpectralGaussianField(shape, 2.0, 0.01, seed=seed)
    c = 0.01 * SpectralGaussianField(shape, 2.0, 1.0, seed=seed + 1)[0]
    return beta, c


def _plan(L, shape, order, dt_bound=None):
    p = L.Plan(len(shape), shape[0], dynamics=L.GLIDE_CLIMB, rk_order=order, Lambda=LAM, mu=MU, nu=NU, cfl=0.25)
    p.set_vacancy(ALPHA, VBETA, GAMMA, dt_bound)
    return p


# 16^3 runs the per-tile flux kernel, 32^3 the plane-marching one
CASES = [((32, 32), 3, False, 8), ((32, 32), 2, True, 8), ((16, 16, 16), 3, False, 6), ((32, 32, 32), 3, True, 4),
         ((32, 32, 32), 2, False, 4)]


@pytest.mark.parametrize("shape,order,cap,nsteps", CASES)
def test_vacancy_steps_match_oracle(L, shape, order, cap, nsteps):
    say("vacancy steps", shape, order, cap, nsteps)
    beta, c = _fields(shape, 71)
    dyn = O.Dynamics(O.GLIDE_CLIMB, LAM, MU, NU)
    dt_bound = None
    if cap:  # half the first CFL step: binding on every step
        _, _, _, d0 = O.run_vacancy(dyn, O.Vacancy(ALPHA, VBETA, GAMMA), beta, c, 0.0, 1, 0.25, order=order)
        dt_bound = 0.5 * d0[0]
    ru, rc, rt, rdts = O.run_vacancy(dyn, O.Vacancy(ALPHA, VBETA, GAMMA, dt_bound), beta, c, 0.0, nsteps, 0.25,
                                     order=order)
    p = _plan(L, shape, order, dt_bound)
    p.set_state(beta)
    p.vacancy_set(c)
    p.set_time(0.0)
    p.step(nsteps)
    gu = p.get_state().cpu().numpy()
    gc = p.vacancy_get().cpu().numpy()
    t = p.clock()[0]
    say("  rel u %.2e  rel c %.2e  rel dc %.2e" % (rel(gu, ru), rel(gc, rc), rel(gc - c, rc - c)))
    assert abs(t - rt) <= 1e-12 * rt
    if cap:
        assert np.allclose(rdts, dt_bound, rtol=0, atol=0)
    assert rel(gu, ru) < 1e-11
    assert rel(gc, rc) < 1e-11
    assert rel(gc - c, rc - c) < 1e-9


@pytest.mark.parametrize("shape", [(32, 32), (16, 16, 16)])
def test_vacancy_flux_matches_oracle(L, shape):
    """cdd_flux of a vacancy plan: BetaP_VacancyDynamics.CalculateFlux."""
    say("vacancy flux", shape)
    beta, c = _fields(shape, 81)
    dyn = O.Dynamics(O.GLIDE_CLIMB, LAM, MU, NU)
    ref = O.central_upwind_flux(beta, O.vacancy_sigma(dyn, O.Vacancy(ALPHA), beta, c, 0.0), O.GLIDE_CLIMB, LAM)
    p = _plan(L, shape, 3)
    p.vacancy_set(c)
    rhs, speed = p.flux(beta)
    assert rel(rhs.cpu().numpy(), ref[0]) < 1e-12
    assert rel(speed.cpu().numpy(), ref[1]) < 1e-12
    # the state's own stress does not include the vacancies
    sig = p.sigma(beta).cpu().numpy()
    assert rel(sig, O.sigma(beta, MU, NU)) < 1e-13


def test_vacancy_off_is_the_plain_stepper(L):
    say("vacancy off")
    shape = (32, 32)
    beta, c = _fields(shape, 91)
    p = _plan(L, shape, 3)
    p.vacancy_set(c)
    p.set_vacancy(enable=False)
    q = L.Plan(2, 32, dynamics=L.GLIDE_CLIMB, rk_order=3, Lambda=LAM, mu=MU, nu=NU, cfl=0.25)
    for x in (p, q):
        x.set_state(beta)
        x.set_time(0.0)
        x.step(5)
    assert np.array_equal(p.get_state().cpu().numpy(), q.get_state().cpu().numpy())
    assert np.array_equal(p.vacancy_get().cpu().numpy(), c)


def test_adjoint_refuses_a_vacancy_plan(L):
    p = _plan(L, (16, 16), 2)
    beta, _ = _fields((16, 16), 95)
    with pytest.raises(L.CDDError):
        p.flux_vjp(beta, 0.0, np.zeros_like(beta))
    st = torch.empty(p.segment_store_bytes(1), dtype=torch.uint8, device="cuda")
    with pytest.raises(L.CDDError):
        p.segment_forward(1, st)


def test_reference_api_vacancy_mover(L):
    """VacancyState + BetaP_VacancyDynamics + OperatorSplittingTVDRK_FieldMover
    through PlasticitySystem (the reference's driver classes), with the
    mover's default dtBound = 0.01 binding, against the oracle."""
    say("vacancy reference API")
    from Plasticity import PlasticitySystem
    from Plasticity.FieldDynamics.VacancyDynamics import BetaP_VacancyDynamics
    from Plasticity.FieldMovers import FieldMover
    from Plasticity.Observers import Observer
    from Plasticity.PlasticityStates.VacancyState import VacancyState
    n = 32
    beta, c = _fields((n, n), 101)
    field = np.concatenate([c[None], beta])  # sorted order: ('s', 's') first
    state = VacancyState((n, n), field=field, mu=MU, nu=NU, alpha=ALPHA)
    rec = Observer.RecallStateObserver()
    dyn = BetaP_VacancyDynamics(Lambda=LAM, gamma=GAMMA, alpha=ALPHA, beta=VBETA)
    mover = FieldMover.OperatorSplittingTVDRK_FieldMover(CFLsafeFactor=0.25)
    sysm = PlasticitySystem.PlasticitySystem((n, n), state, mover, dyn, [rec])
    vac = O.Vacancy(ALPHA, VBETA, GAMMA, dt_bound=0.01)
    odyn = O.Dynamics(O.GLIDE_CLIMB, LAM, MU, NU)
    _, _, t4, dts = O.run_vacancy(odyn, vac, beta, c, 0.0, 4, 0.25)
    assert np.all(dts == 0.01)
    sysm.Run(startTime=0.0, endTime=t4)
    ru, rc, rt, _ = O.run_vacancy(odyn, vac, beta, c, 0.0, 4, 0.25, t_end=t4)
    got = rec.state.GetOrderParameterField().numpy()
    assert rel(got[1:], ru) < 1e-11
    assert rel(got[0], rc) < 1e-11
    # the flux through the reference signature CalculateFlux(time, state, cfield)
    bstate, cf = state.DecoupleState()
    rhs, speed = dyn.CalculateFlux(0.0, bstate, cf, CFLCondition=True)
    ref = O.central_upwind_flux(beta, O.vacancy_sigma(odyn, vac, beta, c, 0.0), O.GLIDE_CLIMB, LAM)
    assert rel(rhs.numpy(), ref[0]) < 1e-12


@pytest.mark.parametrize("shape,order", [((32, 32), 3), ((32, 32, 32), 2)])
def test_vacancy_alpha_field_matches_oracle(L, shape, order):
    """alpha as a field (the dynamics' alpha * cfield with an array alpha,
    VacancyDynamics.py:24-28): the GPU run and the reference-API mover
    against the oracle with the same field."""
    from Plasticity.PlasticityStates.VacancyState import VacancyState
    from Plasticity.FieldDynamics.VacancyDynamics import BetaP_VacancyDynamics
    from Plasticity.FieldMovers.FieldMover import OperatorSplittingTVDRK_FieldMover
    from Plasticity.PlasticitySystem import PlasticitySystem
    from Plasticity.Observers.Observer import RecallStateObserver
    say("vacancy alpha field", shape, order)
    beta, c = _fields(shape, 73)
    alpha = ALPHA * (1.0 + 0.5 * SpectralGaussianField(shape, 3.0, 1.0, seed=75)[4])
    dyn = O.Dynamics(O.GLIDE_CLIMB, LAM, MU, NU)
    ru, rc, rt, rdts = O.run_vacancy(dyn, O.Vacancy(alpha, VBETA, GAMMA), beta, c, 0.0, 5, 0.25, order=order)
    p = _plan(L, shape, order)
    p.vacancy_alpha(alpha)
    p.set_state(beta)
    p.vacancy_set(c)
    p.set_time(0.0)
    p.step(5)
    assert rel(p.get_state().cpu().numpy(), ru) < 1e-11 and rel(p.vacancy_get().cpu().numpy(), rc) < 1e-11
    # the scalar run differs (the field matters)
    su, _, _, _ = O.run_vacancy(dyn, O.Vacancy(ALPHA, VBETA, GAMMA), beta, c, 0.0, 5, 0.25, order=order)
    assert rel(su, ru) > 1e-8
    # through the reference API: the dynamics' alpha field
    state = VacancyState(shape, field=np.concatenate([c[None], beta]), mu=MU, nu=NU)
    obs = RecallStateObserver()
    mover = OperatorSplittingTVDRK_FieldMover(CFLsafeFactor=0.25, dtBound=None, order=order)
    PlasticitySystem(shape, state, mover, BetaP_VacancyDynamics(Lambda=LAM, alpha=alpha, beta=VBETA, gamma=GAMMA),
                     [obs]).Run(startTime=0.0, endTime=rt * (1 - 1e-12))
    assert mover.numberOfSteps == 5
    got = obs.state.GetOrderParameterField().numpy()
    assert rel(got[1:], ru) < 1e-11 and rel(got[0], rc) < 1e-11
