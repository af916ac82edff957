"""The benchmark workloads at their own sizes against the oracle.

  * 3D 256^3 glide + climb (Lambda = 0.9, climb mobility 0.1), TVD-RK3,
    under the uniaxial strain-rate control ramped to a strain of 0.01 over
    1000 steps of the initial CFL step (RunLoadingMPI.py:68-83 with
    FieldMover.py:38-58): the first steps and their stress-strain records
    against the C oracle.
  * A batch of independent 2D realisations with sides 256, 384, 512, 768 and
    1024 (glide only, SSP-RK2, CFL 0.25, L ~ U[3, 8] cells, RMS 0.01): every
    step equal bit for bit to single-instance plans, and the 384 and 1024
    instances against the NumPy and C oracles (the reference's task farm,
    exec_batch.py:26-58, runs them one process per instance).
"""
import sys
import time

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
def env():
    import os
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from oracle import cdd_oracle_c as OC
    if not os.path.exists(OC.LIB):
        pytest.skip("the C oracle is not built (run __graft_entry__.build())")
    from Plasticity.amd import lib as L
    return L, OC


@pytest.mark.timeout(900)
def test_uniaxial_ramp_3d_256_vs_c_oracle(env):
    import torch
    L, OC = env
    n, total, nsteps = 256, 1000, 3
    say("3D 256^3 uniaxial ramp: %d CFL RK3 steps vs C oracle" % nsteps)
    beta = SpectralGaussianField((n,) * 3, 6.0, 0.01, seed=0)
    init_strain = float(beta[8].mean())
    beta_t = torch.from_numpy(beta).cuda()
    probe = L.Plan(3, n, dynamics=L.GLIDE_CLIMB, rk_order=3, Lambda=0.9, mu=1.0, nu=0.3, cfl=0.25)
    _, speed = probe.flux(beta_t)
    dt0 = 0.25 / (n * float(speed.max()))
    probe.close()
    del speed
    rate = 0.01 / (total * dt0)
    plan = L.Plan(3, n, dynamics=L.GLIDE_CLIMB, rk_order=3, Lambda=0.9, mu=1.0, nu=0.3, cfl=0.25,
                  load_kind=L.LOAD_UNIAXIAL, load_init_strain=init_strain, load_uniaxial_rate=rate)
    plan.set_state(beta_t)
    plan.set_time(0.0)
    slog = torch.zeros((nsteps, 8), dtype=torch.float64, device="cuda")
    plan.stress_record(slog)
    plan.step(nsteps)
    plan.stress_record(None)
    got = plan.get_state().cpu().numpy()
    t_gpu = plan.clock()[0]
    rec = slog.cpu().numpy()
    plan.close()
    del beta_t
    say("  GPU done; C oracle ...")
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.9, 1.0, 0.3, O.Load("uniaxial", init_strain=init_strain, uniaxial_rate=rate))
    ref, t = beta, 0.0
    t0 = time.perf_counter()
    for j in range(nsteps):
        sig = OC.sigma(dyn, t, ref)
        want = np.array([sig[O.cidx(a, b)].mean() for a, b in ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))])
        del sig
        assert abs(rec[j, 0] - t) <= 1e-13 * max(t, dt0)
        np.testing.assert_allclose(rec[j, 1:7], want, rtol=1e-11, atol=1e-15)
        assert abs(rec[j, 7] - ref[8].mean()) <= 1e-15
        ref, t, _ = OC.run(dyn, ref, t, 1, 0.25, order=3)
        say("  C oracle step %d, %.1f s" % (j + 1, time.perf_counter() - t0))
    err = rel(got, ref)
    say("  rel L2 %.3e; strain %.3e -> %.3e, mean sigma_zz %.6g -> %.6g" % (err, rec[0, 0] * rate, rec[-1, 0] * rate,
                                                                          rec[0, 3], rec[-1, 3]))
    assert abs(t_gpu - t) <= 1e-14 * t
    assert err < 1e-11


@pytest.mark.timeout(600)
def test_batch_of_benchmark_sides_vs_single_plans_and_oracles(env):
    import torch
    L, OC = env
    rng = np.random.default_rng(7)
    sizes = [1024, 256, 768, 384, 512, 256]
    lens = rng.uniform(3.0, 8.0, size=len(sizes))
    say("batch of sides", sizes)
    betas = [SpectralGaussianField((s, s), float(lens[i]), 0.01, seed=i) for i, s in enumerate(sizes)]
    kw = dict(dynamics=L.GLIDE_ONLY, rk_order=2, cfl=0.25)
    nsteps = 5
    # the multi-instance flux as the tile walk and as the row march
    for tile, per_cu in ((0, 2), (3, 3)):
        b = L.Batch(sizes, **kw)
        b.set_flux(tile, per_cu)
        plans = []
        for i, s in enumerate(sizes):
            b.set_state(i, betas[i])
            p = L.Plan(2, s, **kw)
            p.set_state(betas[i])
            p.set_time(0.0)
            plans.append(p)
        b.set_time(-1, 0.0)
        for step in range(nsteps):
            b.step(1)
            for i, p in enumerate(plans):
                p.step(1)
                assert torch.equal(b.get_state(i), p.get_state()), (tile, step, sizes[i])
                assert b.clock(i)[3] == step + 1 and b.clock(i)[0] == p.clock()[0]
        say("  batch (flux variant %d, %d blocks/CU) = single plans for %d steps" % (tile, per_cu, nsteps))
        if tile == 0:
            b.close()
            for p in plans:
                p.close()
    say("  oracles ...")
    for i, s in enumerate(sizes):
        if s not in (384, 1024) or (s == 384 and i != sizes.index(384)):
            continue
        dyn = O.Dynamics(O.GLIDE_ONLY, 0.0, 1.0, 0.3)
        if s == 384:  # NumPy oracle (the C oracle's FFT is radix 2)
            ref, t, _ = O.run(dyn, betas[i], 0.0, nsteps, 0.25, order=2)
        else:
            ref, t, _ = OC.run(dyn, betas[i], 0.0, nsteps, 0.25, order=2)
        got = b.get_state(i).cpu().numpy()
        err = rel(got, ref)
        say("  side %d: rel L2 %.3e vs the %s oracle" % (s, err, "NumPy" if s == 384 else "C"))
        assert err < 1e-11 and abs(b.clock(i)[0] - t) <= 1e-13 * t
    b.close()
    for p in plans:
        p.close()
