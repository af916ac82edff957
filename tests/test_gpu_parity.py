"""GPU parity: the HIP path through the C ABI against the oracle on the
same seeded fields (tolerances are relative L2 / relative max as noted)."""
import sys

import numpy as np
import pytest

from oracle import cdd_oracle as O
from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField

pytestmark = pytest.mark.gpu

MU, NU = 1.0, 0.3


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


def _plan(amd, d, n, dyn=0, Lambda=0.0, rk=3, **kw):
    return amd.Plan(d, n, dynamics=dyn, rk_order=rk, Lambda=Lambda, mu=MU, nu=NU, cfl=kw.pop("cfl", 0.25), **kw)


@pytest.mark.parametrize("d,n", [(2, 8), (2, 16), (2, 64), (2, 384), (2, 512), (2, 768), (2, 1024), (3, 8), (3, 16),
                                 (3, 32)])
def test_sigma_matches_oracle(amd, d, n):
    say("sigma", d, n)
    beta = SpectralGaussianField((n,) * d, 2.0, 0.01, seed=11)
    p = _plan(amd, d, n)
    got = p.sigma(beta).cpu().numpy()
    ref = O.sigma(beta, MU, NU)
    assert rel_l2(got, ref) < 1e-13


@pytest.mark.parametrize("d,n", [(2, 16), (2, 128), (2, 384), (2, 768), (3, 16)])
def test_rho_matches_oracle(amd, d, n):
    say("rho", d, n)
    beta = SpectralGaussianField((n,) * d, 2.0, 0.01, seed=12)
    p = _plan(amd, d, n)
    got = p.rho(beta).cpu().numpy()
    ref = O.rho_fourier(beta)
    assert rel_l2(got, ref) < 1e-13


@pytest.mark.parametrize("dyn,Lambda", [(0, 0.0), (0, 1.0), (0, 0.9), (1, 0.0)])
@pytest.mark.parametrize("d,n", [(2, 16), (2, 64), (3, 8), (3, 16)])
def test_flux_matches_oracle(amd, d, n, dyn, Lambda):
    say("flux", d, n, dyn, Lambda)
    beta = SpectralGaussianField((n,) * d, 2.0, 0.01, seed=13)
    p = _plan(amd, d, n, dyn, Lambda)
    rhs, speed = p.flux(beta)
    od = O.Dynamics(O.GLIDE_ONLY if dyn == 1 else O.GLIDE_CLIMB, Lambda, MU, NU)
    ref_rhs, ref_speed = od.flux(0.0, beta)
    assert rel_l2(rhs.cpu().numpy(), ref_rhs) < 1e-12
    assert rel_l2(speed.cpu().numpy(), ref_speed) < 1e-12


@pytest.mark.parametrize("rk", [2, 3])
@pytest.mark.parametrize("d,n,dyn,Lambda", [(2, 32, 1, 0.0), (2, 32, 0, 0.9), (3, 16, 0, 0.9), (3, 16, 1, 0.0)])
def test_steps_match_oracle(amd, d, n, dyn, Lambda, rk):
    nsteps = 10
    say("steps", d, n, dyn, Lambda, rk)
    beta = SpectralGaussianField((n,) * d, 2.0, 0.01, seed=14)
    p = _plan(amd, d, n, dyn, Lambda, rk=rk)
    p.set_state(beta)
    p.set_time(0.0)
    p.step(nsteps)
    got = p.get_state().cpu().numpy()
    t, dt, done, ns = p.clock()
    od = O.Dynamics(O.GLIDE_ONLY if dyn == 1 else O.GLIDE_CLIMB, Lambda, MU, NU)
    ref, tref, dts = O.run(od, beta, 0.0, nsteps, 0.25, order=rk)
    assert ns == nsteps
    assert abs(t - tref) <= 1e-12 * tref
    assert rel_l2(got, ref) < 1e-11


def test_energy_matches_oracle(amd):
    say("energy")
    for d, n in [(2, 32), (3, 16)]:
        beta = SpectralGaussianField((n,) * d, 2.0, 0.01, seed=15)
        p = _plan(amd, d, n)
        e = p.energy(beta)
        ref = O.elastic_energy(beta, MU, NU)
        assert abs(e - ref) < 1e-12 * abs(ref)


def test_end_time_stops_on_device(amd):
    say("end time")
    n = 16
    beta = SpectralGaussianField((n, n), 2.0, 0.01, seed=16)
    p = _plan(amd, 2, n, 1, rk=3)
    p.set_state(beta)
    od = O.Dynamics(O.GLIDE_ONLY, 0.0, MU, NU)
    _, t1, dts = O.run(od, beta, 0.0, 1, 0.25)
    t_end = 2.5 * dts[0]
    p.set_time(0.0, t_end)
    p.step(20)  # only the steps up to t_end execute
    t, dt, done, ns = p.clock()
    ref, tref, dts = O.run(od, beta, 0.0, 20, 0.25, t_end=t_end)
    assert done and ns == len(dts) and abs(t - t_end) < 1e-15
    assert rel_l2(p.get_state().cpu().numpy(), ref) < 1e-12


def test_reference_api_drop_in(amd):
    """PlasticitySystem -> TVDRungeKutta_FieldMover -> BetaPDynamics with a
    RecallStateObserver, the call chain of RunPython.simulation
    (RunPython.py:224-257), against the oracle."""
    say("api drop-in")
    from Plasticity import PlasticitySystem
    from Plasticity.FieldMovers import FieldMover
    from Plasticity.FieldDynamics import CentralUpwindHJBetaPGlideOnlyDynamics
    from Plasticity.Observers import Observer
    from Plasticity.PlasticityStates.PlasticityState import PlasticityState
    n = 32
    beta = SpectralGaussianField((n, n), 2.0, 0.01, seed=17)
    state = PlasticityState((n, n), field=beta, mu=MU, nu=NU)
    dyn = CentralUpwindHJBetaPGlideOnlyDynamics.BetaPDynamics()
    mover = FieldMover.TVDRungeKutta_FieldMover(CFLsafeFactor=0.25)
    obs = Observer.RecallStateObserver()
    system = PlasticitySystem.PlasticitySystem((n, n), state, mover, dyn, [obs])
    od = O.Dynamics(O.GLIDE_ONLY, 0.0, MU, NU)
    _, _, dts = O.run(od, beta, 0.0, 6, 0.25)
    t_end = float(np.sum(dts[:5]) + 0.5 * dts[5])
    system.Run(startTime=0.0, endTime=t_end)
    ref, tref, dref = O.run(od, beta, 0.0, 100, 0.25, t_end=t_end)
    assert mover.numberOfSteps == len(dref)
    assert rel_l2(obs.state.betaP.numpy(), ref) < 1e-12
    # the observer-free path (chunked launches) gives the same state
    mover2 = FieldMover.TVDRungeKutta_FieldMover(CFLsafeFactor=0.25)
    mover2.Run(state, [], dyn, 0.0, t_end)
    assert rel_l2(mover2.lastState.betaP.numpy(), ref) < 1e-12
    # stress through the state API
    s = state.CalculateSigma().numpy()
    assert rel_l2(s, O.sigma(beta, MU, NU)) < 1e-13


LOADS = {
    "strain": (dict(load_kind=1, load_init=(0.001, -0.002, 0.003) + (0,) * 6, load_rate=(-0.005, -0.005, 0.01) + (0,) * 6,
                    load_t0=0.1),
               dict(kind="strain", initial=(0.001, -0.002, 0.003), rate=(-0.005, -0.005, 0.01), t0=0.1)),
    # ExternalStrain's own Hooke's law (RunPython's mu = 0.5, nu = 0.3) on a mu = 1 plan
    "strain_runpython": (dict(load_kind=1, load_init=(0.001, -0.002, 0.003) + (0,) * 6,
                              load_rate=(-0.005, -0.005, 0.01) + (0,) * 6, load_t0=0.1, load_mu=0.5, load_nu=0.3),
                         dict(kind="strain", initial=(0.001, -0.002, 0.003), rate=(-0.005, -0.005, 0.01), t0=0.1,
                              elastic=(0.5, 0.3))),
    "stress": (dict(load_kind=2, load_init=(0.002, 0.0, -0.001) + (0,) * 6, load_rate=(0.0, 0.01, 0.02) + (0,) * 6),
               dict(kind="stress", initial=(0.002, 0.0, -0.001), rate=(0.0, 0.01, 0.02))),
    "uniaxial": (dict(load_kind=3, load_init_strain=0.001, load_uniaxial_rate=0.05),
                 dict(kind="uniaxial", init_strain=0.001, uniaxial_rate=0.05)),
    "tensor": (dict(load_kind=4, load_init=(0, 0.005, 0, 0.005, 0, 0, 0, 0, 0), load_rate=(0, 0, 0.01, 0, 0, 0, 0.01, 0, 0)),
               dict(kind="tensor", tensor0=(0, 0.005, 0, 0.005, 0, 0, 0, 0, 0), tensor_rate=(0, 0, 0.01, 0, 0, 0, 0.01, 0, 0))),
}


@pytest.mark.parametrize("kind", sorted(LOADS))
def test_loaded_steps_and_stress_log_match_oracle(amd, kind):
    """Loaded RK3 steps (3D 16^3, glide+climb) against the oracle, and the
    per-step stress-strain record against the oracle's mean stress of the
    GPU's own state at each step start."""
    import torch
    say("loaded steps", kind)
    pk, ok = LOADS[kind]
    n, nsteps = 16, 6
    beta = SpectralGaussianField((n,) * 3, 2.0, 0.01, seed=17)
    p = _plan(amd, 3, n, 0, 0.9, rk=3, **pk)
    p.set_state(beta)
    p.set_time(0.2)
    log = torch.zeros((nsteps, 8), dtype=torch.float64, device="cuda")
    p.stress_record(log)
    states = []
    for _ in range(nsteps):
        states.append(p.get_state().cpu().numpy())
        p.step(1)
    p.stress_record(None)
    got = p.get_state().cpu().numpy()
    load = O.Load(**ok)
    od = O.Dynamics(O.GLIDE_CLIMB, 0.9, MU, NU, load=load)
    ref, tref, dts = O.run(od, beta, 0.2, nsteps, 0.25, order=3)
    assert rel_l2(got, ref) < 1e-11
    rec = log.cpu().numpy()
    t = 0.2
    for j in range(nsteps):
        assert abs(rec[j, 0] - t) <= 1e-13 * t
        want = O.mean_stress(states[j], rec[j, 0], MU, NU, load)
        np.testing.assert_allclose(rec[j, 1:7], want, rtol=1e-11, atol=1e-15)
        assert abs(rec[j, 7] - states[j][8].mean()) <= 1e-15
        t += dts[j]


def test_config1_2d_128_lvp_rk2_100_steps(amd):
    """2D 128^2, seed 0, L = 4 cells, RMS 0.01, mu = 1, nu = 0.3, glide-only,
    TVD-RK2, CFL 0.25, 100 adaptive steps: 1e-11 relative L2 to the oracle."""
    say("config 1: 2D 128 lvp RK2 x100")
    n, nsteps = 128, 100
    beta = SpectralGaussianField((n, n), 4.0, 0.01, seed=0)
    p = _plan(amd, 2, n, 1, 0.0, rk=2)
    p.set_state(beta)
    p.set_time(0.0)
    p.step(nsteps)
    got = p.get_state().cpu().numpy()
    t, _, _, ns = p.clock()
    od = O.Dynamics(O.GLIDE_ONLY, 0.0, MU, NU)
    ref, tref, _ = O.run(od, beta, 0.0, nsteps, 0.25, order=2)
    assert ns == nsteps and abs(t - tref) <= 1e-12 * tref
    assert rel_l2(got, ref) < 1e-11


@pytest.mark.parametrize("d,n,dyn,Lambda", [(3, 32, 0, 0.9), (3, 128, 1, 0.0), (3, 128, 0, 0.0), (2, 64, 1, 0.0)])
def test_fp32_variant_error_against_fp64(amd, d, n, dyn, Lambda):
    """The float32 plan against the float64 plan on the same field: stress,
    flux and 20 RK3 steps, errors within a few hundred float32 ulps
    (relative L2; float32 eps = 1.19e-7).  In 3D the steps run the
    plane-marching kernel in both precisions (at 128^3 in 16-plane chunks,
    two fp32 blocks per CU)."""
    import torch
    say("fp32", d, n, dyn, Lambda)
    beta = SpectralGaussianField((n,) * d, 3.0, 0.01, seed=18)
    p64 = _plan(amd, d, n, dyn, Lambda, rk=3)
    p32 = _plan(amd, d, n, dyn, Lambda, rk=3, precision="fp32")
    s64 = p64.sigma(beta).cpu().numpy()
    s32 = p32.sigma(beta).cpu().numpy()
    assert p32.get_state().dtype == torch.float32
    assert rel_l2(s32, s64) < 1e-5
    r64, v64 = p64.flux(beta)
    r32, v32 = p32.flux(beta)
    assert rel_l2(r32.cpu().numpy(), r64.cpu().numpy()) < 1e-4
    assert rel_l2(v32.cpu().numpy(), v64.cpu().numpy()) < 1e-5
    for p in (p64, p32):
        p.set_state(beta)
        p.set_time(0.0)
        p.step(20)
    e = rel_l2(p32.get_state().cpu().numpy(), p64.get_state().cpu().numpy())
    say("  fp32 vs fp64 after 20 steps: %.2e" % e)
    assert e < 1e-4
    t64, t32 = p64.clock()[0], p32.clock()[0]
    assert abs(t32 - t64) < 1e-4 * t64
    with pytest.raises(amd.CDDError):
        p32.flux_vjp(beta, 0.0, beta)


@pytest.mark.parametrize("Lambda", [0.0, 0.6, 1.0])
def test_wojzz_flux_and_steps_match_oracle(amd, Lambda):
    say("woJzz", Lambda)
    n = 32
    beta = SpectralGaussianField((n, n), 2.5, 0.01, seed=19)
    p = _plan(amd, 2, n, 2, Lambda, rk=3)
    rhs, speed = p.flux(beta)
    od = O.Dynamics(O.GLIDE_CLIMB_WOJZZ, Lambda, MU, NU)
    ref_rhs, ref_speed = od.flux(0.0, beta)
    assert rel_l2(rhs.cpu().numpy(), ref_rhs) < 1e-12
    assert rel_l2(speed.cpu().numpy(), ref_speed) < 1e-12
    assert not rhs[8].any()
    p.set_state(beta)
    p.set_time(0.0)
    p.step(10)
    ref, tref, _ = O.run(od, beta, 0.0, 10, 0.25, order=3)
    assert rel_l2(p.get_state().cpu().numpy(), ref) < 1e-11
    with pytest.raises(amd.CDDError):
        p.flux_vjp(beta, 0.0, beta)
    with pytest.raises(amd.CDDError):
        _plan(amd, 3, 8, 2, Lambda)


def test_wojzz_reference_api(amd):
    from Plasticity import PlasticitySystem
    from Plasticity.FieldDynamics import CentralUpwindHJBetaPDynamics_woJzz as W
    from Plasticity.FieldMovers import FieldMover
    from Plasticity.Observers import Observer
    from Plasticity.PlasticityStates.PlasticityState import PlasticityState
    n = 16
    beta = SpectralGaussianField((n, n), 2.0, 0.01, seed=27)
    state = PlasticityState((n, n), field=beta, mu=MU, nu=NU)
    rec = Observer.RecallStateObserver()
    sysm = PlasticitySystem.PlasticitySystem((n, n), state, FieldMover.TVDRungeKutta_FieldMover(CFLsafeFactor=0.25),
                                             W.BetaPDynamics(Lambda=0.5), [rec])
    _, t1, dts = O.run(O.Dynamics(O.GLIDE_CLIMB_WOJZZ, 0.5, MU, NU), beta, 0.0, 3, 0.25)
    sysm.Run(startTime=0.0, endTime=t1)
    ref, _, _ = O.run(O.Dynamics(O.GLIDE_CLIMB_WOJZZ, 0.5, MU, NU), beta, 0.0, 3, 0.25, t_end=t1)
    assert rel_l2(rec.state.GetOrderParameterField().numpy(), ref) < 1e-11


def test_runpython_strain_loading_uses_its_own_elastic_constants(amd):
    """UpwindLoadingBetaPDynamics (RunPython.py:35-49) on a mu = 1 state: the
    external strain's stress takes RunPython's module constants mu = 0.5,
    nu = 0.3 (RunPython.py:19-26), not the state's; the state's stress solve
    keeps mu = 1."""
    from Plasticity.RunPython import UpwindLoadingBetaPDynamics
    from Plasticity.PlasticityStates.PlasticityState import PlasticityState
    from Plasticity.FieldMovers import FieldMover
    from Plasticity.PlasticitySystem import PlasticitySystem
    from Plasticity.Observers.Observer import RecallStateObserver
    say("runpython strain load at mu = 1")
    n = 16
    beta = SpectralGaussianField((n,) * 3, 2.0, 0.01, seed=19)
    state = PlasticityState((n,) * 3, field=beta, mu=MU, nu=NU)
    dyn = UpwindLoadingBetaPDynamics(rate=np.array([-0.005, -0.005, 0.01]), initial=np.array([0.001, 0.0, -0.002]))
    obs = RecallStateObserver()
    mover = FieldMover.TVDRungeKutta_FieldMover(CFLsafeFactor=0.25)
    t_end = 0.02
    PlasticitySystem((n,) * 3, state, mover, dyn, [obs]).Run(startTime=0.0, endTime=t_end)
    load = O.Load(kind="strain", initial=(0.001, 0.0, -0.002), rate=(-0.005, -0.005, 0.01), elastic=(0.5, 0.3))
    ref, tref, dts = O.run(O.Dynamics(O.GLIDE_CLIMB, 0.0, MU, NU, load=load), beta, 0.0, 10 ** 6, 0.25, t_end=t_end)
    assert mover.numberOfSteps == len(dts)
    assert rel_l2(obs.state.betaP.numpy(), ref) < 1e-11
    # the flux at the state: against the state's own constants it would differ
    rhs = dyn.CalculateFlux(0.01, state).numpy()
    want = O.Dynamics(O.GLIDE_CLIMB, 0.0, MU, NU, load=load).flux(0.01, beta)[0]
    wrong = O.Dynamics(O.GLIDE_CLIMB, 0.0, MU, NU, load=O.Load(kind="strain", initial=(0.001, 0.0, -0.002),
                                                                 rate=(-0.005, -0.005, 0.01))).flux(0.01, beta)[0]
    assert rel_l2(rhs, want) < 1e-12 and rel_l2(rhs, wrong) > 1e-6
