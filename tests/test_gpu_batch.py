"""Batched stepping: every instance of a mixed-size batch, advanced by the
captured hipGraph step, equals the single-plan (eager) path step for step,
and matches the oracle; per-instance end times; autotuned tiles give the
same results."""
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
def amd():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from Plasticity import amd as A
    return A


@pytest.mark.parametrize("rk,dyn,Lambda", [(2, 1, 0.0), (3, 0, 0.9)])
def test_batch_equals_single_plans(amd, rk, dyn, Lambda):
    sizes = [32, 16, 64, 32, 128, 16]
    say("batch", rk, dyn, sizes)
    L = amd.lib
    b = L.Batch(sizes, dynamics=dyn, rk_order=rk, Lambda=Lambda, cfl=0.25)
    betas = [SpectralGaussianField((n, n), 2.0 + 0.5 * i, 0.01, seed=100 + i) for i, n in enumerate(sizes)]
    plans = []
    for i, n in enumerate(sizes):
        b.set_state(i, betas[i])
        p = L.Plan(2, n, dynamics=dyn, rk_order=rk, Lambda=Lambda, cfl=0.25)
        p.set_state(betas[i])
        p.set_time(0.0)
        plans.append(p)
    b.set_time(-1, 0.0)
    for step in range(5):  # each replay checked against the eager path, with syncs between
        b.step(1)
        for i, p in enumerate(plans):
            p.step(1)
            got = b.get_state(i).cpu().numpy()
            ref = p.get_state().cpu().numpy()
            assert np.array_equal(got, ref), (step, i)
            assert b.clock(i)[3] == step + 1
    # and the oracle for the smallest instances
    for i in (1, 5):
        od = O.Dynamics(O.GLIDE_ONLY if dyn == 1 else O.GLIDE_CLIMB, Lambda, 1.0, 0.3)
        ref, t, _ = O.run(od, betas[i], 0.0, 5, 0.25, order=rk)
        got = b.get_state(i).cpu().numpy()
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-12
        assert abs(b.clock(i)[0] - t) < 1e-12 * t


def test_batch_per_instance_end_time(amd):
    say("batch end time")
    L = amd.lib
    sizes = [32, 32, 64]
    b = L.Batch(sizes, dynamics=1, rk_order=2, cfl=0.25)
    betas = [SpectralGaussianField((n, n), 2.0, 0.01, seed=200 + i) for i, n in enumerate(sizes)]
    for i in range(3):
        b.set_state(i, betas[i])
    b.set_time(-1, 0.0)
    od = O.Dynamics(O.GLIDE_ONLY, 0.0, 1.0, 0.3)
    _, _, dts = O.run(od, betas[1], 0.0, 3, 0.25, order=2)
    t_end = float(dts[0] + 0.5 * dts[1])  # instance 1 stops after 2 steps (odd buffer parity)
    b.set_time(1, 0.0, t_end)
    b.step(5)
    ref, t, d = O.run(od, betas[1], 0.0, 5, 0.25, order=2, t_end=t_end)
    assert b.clock(1)[2] and b.clock(1)[3] == 2 and b.clock(0)[3] == 5
    got = b.get_state(1).cpu().numpy()
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-12


def test_autotuned_tiles_same_result(amd):
    say("batch autotune")
    from Plasticity.amd.autotune import autotune_batch
    L = amd.lib
    sizes = [64, 128, 128]
    betas = [SpectralGaussianField((n, n), 3.0, 0.01, seed=300 + i) for i, n in enumerate(sizes)]
    outs = []
    for variant in (0, 1, 2, "auto"):
        b = L.Batch(sizes, dynamics=0, rk_order=2, Lambda=0.9, cfl=0.25)
        for i in range(3):
            b.set_state(i, betas[i])
        b.set_time(-1, 0.0)
        if variant == "auto":
            choices, timings = autotune_batch(b, reps=2)
            assert {64, 128} <= set(choices) and choices["multi"][1] in (0, 1, 2, 3)
            say("  auto", choices, timings.get("multi"))
        else:
            b.set_flux(0, 0)  # per-size launches with this tile
            for n in (64, 128):
                b.set_tile(n, variant)
        b.step(3)
        outs.append([b.get_state(i).cpu().numpy() for i in range(3)])
        b.close()
    for o in outs[1:]:
        for a, c in zip(outs[0], o):
            assert np.array_equal(a, c)


@pytest.mark.parametrize("rk", [2, 3])
def test_multi_instance_flux_configurations(amd, rk):
    """Every multi-instance flux configuration (tile walk variant x
    persistent blocks per CU, and the row march at 1-3 blocks per CU, with
    the sides a variant does not take left to per-size launches) gives the
    per-size launches' states and clocks bit for bit; one instance ends
    early."""
    L = amd.lib
    sizes = [16, 32, 64, 256, 16, 128]
    say("multi-instance configurations", rk, sizes)
    betas = [SpectralGaussianField((n, n), 2.5, 0.01, seed=400 + i) for i, n in enumerate(sizes)]
    p = L.Plan(2, 256, dynamics=0, rk_order=rk, Lambda=0.9, cfl=0.25)
    p.set_state(betas[3])
    p.set_time(0.0)
    p.step(1)
    t_end = 1.5 * p.clock()[1]  # the 256 instance stops after its second step
    p.close()
    ref = None
    for tile, per_cu in [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 1), (2, 2), (3, 1), (3, 2), (3, 3)]:
        b = L.Batch(sizes, dynamics=0, rk_order=rk, Lambda=0.9, cfl=0.25)
        for i in range(len(sizes)):
            b.set_state(i, betas[i])
        b.set_time(-1, 0.0)
        b.set_time(3, 0.0, t_end)
        b.set_flux(tile, per_cu)
        b.step(4)
        out = [b.get_state(i).cpu().numpy() for i in range(len(sizes))]
        clk = [b.clock(i) for i in range(len(sizes))]
        b.close()
        if ref is None:
            ref = (out, clk)
            assert clk[3][2] and clk[3][3] == 2 and clk[0][3] == 4
            continue
        for i in range(len(sizes)):
            assert np.array_equal(out[i], ref[0][i]), (tile, per_cu, sizes[i])
            assert clk[i] == ref[1][i], (tile, per_cu, sizes[i])


def test_spectral_row_tiles_same_result(amd):
    """The stress solve's half-row R2C / C2R tiles (the autotuner's second
    choice) give the default tiles' states bit for bit."""
    L = amd.lib
    sizes = [256, 512, 256]
    say("spectral row tiles", sizes)
    betas = [SpectralGaussianField((n, n), 3.0, 0.01, seed=500 + i) for i, n in enumerate(sizes)]
    outs = []
    for v in (0, 1):
        b = L.Batch(sizes, dynamics=0, rk_order=3, Lambda=0.9, cfl=0.25)
        for i in range(len(sizes)):
            b.set_state(i, betas[i])
        b.set_time(-1, 0.0)
        for n in set(sizes):
            b.set_spectral(n, v)
        b.step(3)
        outs.append([b.get_state(i).cpu().numpy() for i in range(len(sizes))])
        b.close()
    for a, c in zip(outs[0], outs[1]):
        assert np.array_equal(a, c)
