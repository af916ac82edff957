"""Device-side observers and failure detectors: the per-step energy record,
the large-k instability measure and the non-finite-speed guard."""
import sys

import numpy as np
import pytest

from oracle import cdd_oracle as O
from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from Plasticity.amd import lib
    return lib


@pytest.mark.parametrize("d,n", [(2, 32), (3, 16)])
def test_energy_record_matches_energy_of_each_step_start(L, d, n):
    import torch
    say("energy record", d, n)
    beta = SpectralGaussianField((n,) * d, 2.0, 0.01, seed=22)
    p = L.Plan(d, n, dynamics=0, rk_order=3, Lambda=0.9, load_kind=2, load_init=(0.002, 0.0, -0.001),
               load_rate=(0.0, 0.01, 0.0))
    p.set_state(beta)
    p.set_time(0.1)
    log = torch.zeros(5, dtype=torch.float64, device="cuda")
    p.energy_record(log)
    states, times = [], []
    for _ in range(5):
        states.append(p.get_state().clone())
        times.append(p.clock()[0])
        p.step(1)
    p.energy_record(None)
    got = log.cpu().numpy()
    for j in range(5):
        want = p.energy(states[j], times[j])
        assert abs(got[j] - want) <= 1e-12 * abs(want)


@pytest.mark.parametrize("d,n,fixed", [(2, 32, False), (3, 32, True)])
def test_speed_record_is_each_steps_cfl_speed(L, d, n, fixed):
    """Every step's logged speed is the oracle's largest one-sided corner
    speed of that step's start state (the max CalculateOneAdaptiveTimeStep
    turns into dt), under the CFL rule and under a fixed step with the
    uniaxial strain-rate control (3D 32^3: the plane-marching kernel)."""
    import torch
    say("speed record", d, n, fixed)
    beta = SpectralGaussianField((n,) * d, 2.0, 0.01, seed=24)
    init = float(beta[8].mean())
    kw = dict(load_kind=L.LOAD_UNIAXIAL, load_init_strain=init, load_uniaxial_rate=0.05) if d == 3 else {}
    p = L.Plan(d, n, dynamics=0, rk_order=3, Lambda=0.9, cfl=0.25, **kw)
    p.set_state(beta)
    if fixed:
        p.set_time(0.0, float("inf"), 2e-3)
    else:
        p.set_time(0.0)
    nsteps = 4
    log = torch.zeros(nsteps, dtype=torch.float64, device="cuda")
    dts = torch.zeros(nsteps, dtype=torch.float64, device="cuda")
    p.speed_record(log)
    p.dt_record(dts)
    load = O.Load("uniaxial", init_strain=init, uniaxial_rate=0.05) if d == 3 else None
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.9, 1.0, 0.3, load=load)
    want = []
    for _ in range(nsteps):
        t = p.clock()[0]
        _, speed = dyn.flux(t, p.get_state().cpu().numpy())
        want.append(speed.max())
        p.step(1)
    p.speed_record(None)
    p.dt_record(None)
    got = log.cpu().numpy()
    assert np.all(np.abs(got - np.array(want)) <= 1e-11 * np.abs(want))
    dt = dts.cpu().numpy()
    if fixed:
        assert np.all(dt == 2e-3)
    else:  # the adaptive step is CFL / (N max speed)
        assert np.all(np.abs(dt * n * got - 0.25) <= 1e-14)


@pytest.mark.parametrize("d,n", [(2, 64), (3, 16)])
def test_instability_matches_oracle(L, d, n):
    say("instability", d, n)
    beta = SpectralGaussianField((n,) * d, 1.5, 0.01, seed=23)
    p = L.Plan(d, n)
    got = p.instability(beta)
    want = O.instability(beta)
    assert abs(got - want) <= 1e-11 * want


def test_observer_flags_large_k_growth(L):
    import torch
    from Plasticity.Observers.Observer import LargeKInstabilityObserver
    from Plasticity.PlasticityStates.PlasticityState import PlasticityState
    n = 32
    beta = SpectralGaussianField((n, n), 3.0, 0.01, seed=24)
    obs = LargeKInstabilityObserver(limit=1e3)
    v0 = obs.Update(0.0, PlasticityState((n, n), field=beta))
    obs.Update(0.1, PlasticityState((n, n), field=beta))
    assert not obs.stop
    # a grid-scale (wavenumber N/2 - 1) mode in beta_xy: |rho| gains Nyquist content
    x = np.arange(n)
    mode = np.cos(2 * np.pi * (n // 2 - 1) * x / n)
    bad = beta.copy()
    bad[1] += 0.05 * np.outer(mode, mode)
    v1 = obs.Update(0.2, PlasticityState((n, n), field=bad))
    assert obs.stop
    assert abs(v0 - O.instability(beta)) <= 1e-11 * v0
    assert abs(v1 - O.instability(bad)) <= 1e-11 * v1 and v1 / v0 > 1e3


def test_non_finite_speed_stops_the_run(L):
    say("nan guard")
    n = 16
    beta = SpectralGaussianField((n, n), 2.0, 0.01, seed=25)
    beta[0, 3, 5] = np.nan
    p = L.Plan(2, n, dynamics=1, rk_order=3)
    p.set_state(beta)
    p.set_time(0.0)
    p.step(4)
    t, dt, done, ns = p.clock()
    assert p.status() == 1 and done and ns == 0 and t == 0.0
    # the state is the step's start state, untouched
    got = p.get_state().cpu().numpy()
    assert np.array_equal(np.isnan(got), np.isnan(beta)) and np.array_equal(got[~np.isnan(got)], beta[~np.isnan(beta)])
    # a fresh start clears the status
    p.set_state(SpectralGaussianField((n, n), 2.0, 0.01, seed=26))
    p.set_time(0.0)
    p.step(2)
    assert p.status() == 0 and p.clock()[3] == 2


def test_mover_raises_on_non_finite_state(L):
    """PlasticitySystem.Run (the reference API) stops with FloatingPointError
    when the device guard trips: no observer sees the unadvanced state again
    and Run does not return as if the interval had been covered."""
    from Plasticity.PlasticitySystem import PlasticitySystem
    from Plasticity.PlasticityStates.PlasticityState import PlasticityState
    from Plasticity.FieldMovers.FieldMover import TVDRungeKutta_FieldMover
    from Plasticity.FieldDynamics.CentralUpwindHJBetaPGlideOnlyDynamics import BetaPDynamics
    say("mover nan guard")
    n = 16
    beta = SpectralGaussianField((n, n), 2.0, 0.01, seed=27)
    beta[4, 7, 2] = np.nan
    state = PlasticityState((n, n), field=beta, mu=1.0, nu=0.3)
    seen = []

    class Count:
        def Update(self, t, st):
            seen.append(t)

    for obs in ([Count()], []):
        seen.clear()
        mover = TVDRungeKutta_FieldMover(CFLsafeFactor=0.25, order=2)
        system = PlasticitySystem((n, n), state, mover, BetaPDynamics(), obs)
        with pytest.raises(FloatingPointError):
            system.Run(startTime=0.0, endTime=1.0)
        assert seen == [] and mover.numberOfSteps == 0
