"""The reference's initialisers on the GPU path (FieldInitializer.py:153-209,
430-442, 488-503): runs started from GaussianRandomInitializer states, a
run continued after ResizeState (the larger-N continuation of the
reference's drivers), and the 90-degree rotation, each through the C ABI
against the oracle."""
import sys

import numpy as np
import pytest

from oracle import cdd_oracle as O
from Plasticity.FieldInitializers.FieldInitializer import (GaussianRandomInitializer, ResizeState,
                                                           RotateState_90degree)

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


def _gpu_run(L, beta, d, n, dyn, Lam, rk, nsteps, t0=0.0):
    p = L.Plan(d, n, dynamics=dyn, rk_order=rk, Lambda=Lam, mu=0.5, nu=0.3, cfl=0.1)
    p.set_state(beta)
    p.set_time(t0)
    p.step(nsteps)
    return p.get_state().cpu().numpy(), p.clock()[0]


@pytest.mark.parametrize("d,n,dyn,Lam,rk", [(2, 128, 1, 0.0, 3), (2, 64, 0, 1.0, 2), (3, 32, 0, 0.0, 3)])
def test_runs_from_the_reference_initializer(L, d, n, dyn, Lam, rk):
    """GaussianRandomInitializer (128^d generation, k filter, resampling to
    n^d) as the initial state of 10 GPU steps, against the oracle."""
    say("initializer run", d, n, dyn, Lam, rk)
    beta = GaussianRandomInitializer((n,) * d, 0.2, seed=3, device="cpu").GetOrderParameterField().numpy()
    got, t = _gpu_run(L, beta, d, n, dyn, Lam, rk, 10)
    od = O.Dynamics(O.GLIDE_ONLY if dyn == 1 else O.GLIDE_CLIMB, Lam, 0.5, 0.3)
    ref, tref, _ = O.run(od, beta, 0.0, 10, 0.1, order=rk)
    assert rel(got, ref) < 1e-11 and abs(t - tref) <= 1e-12 * tref


def test_resized_continuation(L):
    """A 2D run at 32^2, resized to 64^2 with ResizeState from the GPU state,
    continued on the GPU: equal to the oracle doing the same."""
    say("resized continuation")
    beta = GaussianRandomInitializer((32, 32), 0.2, seed=5, device="cpu").GetOrderParameterField().numpy()
    mid, t1 = _gpu_run(L, beta, 2, 32, 1, 0.0, 3, 8)
    from Plasticity.PlasticityStates.PlasticityState import PlasticityState
    big = ResizeState(PlasticityState((32, 32), field=mid, device="cpu"), 64, 2, device="cpu")
    got, t2 = _gpu_run(L, big.GetOrderParameterField().numpy(), 2, 64, 1, 0.0, 3, 8, t0=t1)
    od = O.Dynamics(O.GLIDE_ONLY, 0.0, 0.5, 0.3)
    r1, tr1, _ = O.run(od, beta, 0.0, 8, 0.1, order=3)
    rbig = ResizeState(PlasticityState((32, 32), field=r1, device="cpu"), 64, 2, device="cpu")
    ref, tr2, _ = O.run(od, rbig.GetOrderParameterField().numpy(), tr1, 8, 0.1, order=3)
    assert rel(mid, r1) < 1e-11 and rel(got, ref) < 1e-11 and abs(t2 - tr2) <= 1e-12 * tr2


def test_rotated_state_rotates_the_gpu_flux(L):
    """RotateState_90degree maps the GPU flux of a state to the flux of the
    rotated state (square grids are isotropic), to rounding."""
    say("rotation")
    n = 64
    st = GaussianRandomInitializer((n, n), 0.2, seed=9, device="cpu")
    rot = RotateState_90degree(st, device="cpu")
    p = L.Plan(2, n, dynamics=0, rk_order=3, Lambda=0.5, mu=0.5, nu=0.3)
    rhs, _ = p.flux(st.GetOrderParameterField().numpy())
    rhs_rot, _ = p.flux(rot.GetOrderParameterField().numpy())
    from Plasticity.PlasticityStates.PlasticityState import PlasticityState
    want = RotateState_90degree(PlasticityState((n, n), field=rhs.cpu().numpy(), device="cpu"),
                                device="cpu").GetOrderParameterField().numpy()
    assert rel(rhs_rot.cpu().numpy(), want) < 1e-12
