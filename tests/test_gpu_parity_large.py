"""GPU parity at larger grids and longer runs, against the C/OpenMP
restatement of the oracle (oracle/cdd_oracle.c, itself checked against the
NumPy oracle in tests/test_oracle_c.py).  Glide + climb, climb mobility 0.1
(Lambda = 0.9), TVD-RK3, CFL 0.25, the seeded field of the 3D workloads
(L = 6 cells, RMS 0.01); 1e-11 relative L2 after the run."""
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
def env():
    import os
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from oracle import cdd_oracle_c as OC
    if not os.path.exists(OC.LIB):
        pytest.skip("the C oracle is not built (run __graft_entry__.build())")
    from Plasticity import amd as A
    return A, OC


_FIELDS = {}


def _field(n, seed):
    key = (n, seed)
    if key not in _FIELDS:
        _FIELDS.clear()
        _FIELDS[key] = SpectralGaussianField((n, n, n), 6.0, 0.01, seed=seed)
    return _FIELDS[key]


def _compare(env, n, nsteps, seed=2, dynamics=0, Lambda=0.9, rk=3):
    A, OC = env
    beta = _field(n, seed)
    p = A.Plan(3, n, dynamics=dynamics, rk_order=rk, Lambda=Lambda, mu=1.0, nu=0.3, cfl=0.25)
    p.set_state(beta)
    p.set_time(0.0)
    p.step(nsteps)
    got = p.get_state().cpu().numpy()
    t, _, _, ns = p.clock()
    say("  GPU done; C oracle ...")
    import time
    dyn = O.Dynamics(O.GLIDE_ONLY if dynamics == 1 else O.GLIDE_CLIMB, Lambda, 1.0, 0.3)
    # the C oracle in chunks of 10 steps (same step sequence), with progress
    ref, tref, dts = beta, 0.0, []
    t0 = time.perf_counter()
    while len(dts) < nsteps:
        k = min(10, nsteps - len(dts))
        ref, tref, d = OC.run(dyn, ref, tref, k, 0.25, order=rk)
        dts.extend(d)
        say("  C oracle %d/%d steps, %.1f s" % (len(dts), nsteps, time.perf_counter() - t0))
    err = np.linalg.norm((got - ref).ravel()) / np.linalg.norm(ref.ravel())
    say("  rel L2 %.3e, t %.17g vs %.17g" % (err, t, tref))
    assert ns == nsteps and len(dts) == nsteps
    assert abs(t - tref) <= 1e-12 * tref
    assert err < 1e-11
    return err


@pytest.mark.parametrize("rk", [2, 3])
@pytest.mark.parametrize("dynamics,Lambda", [(0, 0.9), (0, 0.0), (1, 0.0)])
def test_3d_32_steps(env, dynamics, Lambda, rk):
    """32^3: the smallest grid the plane-marching flux kernel runs on (two
    16-plane chunks per column), both dynamics, both RK orders."""
    say("3D 32^3 x 12 steps vs C oracle", dynamics, Lambda, rk)
    _compare(env, 32, 12, seed=5, dynamics=dynamics, Lambda=Lambda, rk=rk)


def test_3d_64_50_steps(env):
    say("3D 64^3 x 50 RK3 steps vs C oracle")
    _compare(env, 64, 50)


def test_3d_128_5_steps(env):
    say("3D 128^3 x 5 RK3 steps vs C oracle")
    _compare(env, 128, 5)


@pytest.mark.parametrize("dynamics,Lambda", [(0, 0.9), (0, 0.0), (1, 0.0)])
def test_3d_256_one_step(env, dynamics, Lambda):
    """256^3, the default bench grid: every column's block marches the
    whole 256-plane ring as one chunk (both tile buffers, the fetches past
    the chunk wrapping round), through all three RK3 stages, for each
    instance of the plane-marching kernel (climb on / off, glide only)."""
    say("3D 256^3 x 1 RK3 step vs C oracle", dynamics, Lambda)
    _compare(env, 256, 1, seed=3, dynamics=dynamics, Lambda=Lambda)


@pytest.mark.timeout(1200)
def test_config2_3d_128_200_steps(env):
    """The full 3D workload check: 128^3, 200 RK3 steps (about four minutes
    of host CPU for the C oracle on 16 threads)."""
    say("config 2: 3D 128^3 x 200 RK3 steps vs C oracle")
    _compare(env, 128, 200)


def test_3d_256_fp32_vs_fp64(env):
    """The fp32 plan at the bench grid, where its plane-marching kernel marches
    128-plane chunks at two blocks per CU: 5 RK3 steps against the fp64 plan
    on the same field, within a few hundred float32 ulps (relative L2; float32
    eps = 1.19e-7), both dynamics."""
    import torch
    A, _ = env
    n = 256
    beta = torch.from_numpy(_field(n, 2)).cuda()
    for dynamics, Lambda in ((0, 0.9), (1, 0.0)):
        say("3D 256^3 fp32 vs fp64, dynamics %d" % dynamics)
        out = []
        for prec in ("fp64", "fp32"):
            p = A.Plan(3, n, dynamics=dynamics, rk_order=3, Lambda=Lambda, mu=1.0, nu=0.3, cfl=0.25, precision=prec)
            p.set_state(beta)
            p.set_time(0.0)
            p.step(5)
            out.append((p.get_state().double(), p.clock()[0]))
            p.close()
        (s64, t64), (s32, t32) = out
        e = float(torch.linalg.norm(s32 - s64) / torch.linalg.norm(s64))
        say("  rel L2 %.2e, t %.9g vs %.9g" % (e, t32, t64))
        assert e < 1e-5
        assert abs(t32 - t64) < 1e-5 * t64
        del out, s64, s32
        torch.cuda.empty_cache()
