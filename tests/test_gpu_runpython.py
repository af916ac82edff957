"""RunPython.simulation on the GPU (Plasticity/RunPython.py; RunPython.py:
132-257) against the oracle over the same record intervals: the raw record
file of a relaxation, a strain-loaded run, and a DisorderState ('x') run."""
import os
import sys

import numpy as np
import pytest

from oracle import cdd_oracle as O

pytestmark = pytest.mark.gpu

MU, NU, CFL = 0.5, 0.3, 0.1


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / np.linalg.norm(np.asarray(b).ravel())


@pytest.fixture(scope="module")
def R():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from Plasticity import RunPython
    return RunPython


def records(path, ncomp, N, dim):
    raw = np.fromfile(path, dtype=np.float64).reshape(-1, 1 + ncomp * N ** dim)
    return [(r[0], r[1:].reshape((ncomp,) + (N,) * dim)) for r in raw]


def oracle_intervals(R, od, u, end):
    out, t = [(0.0, u)], 0.0
    while t <= end:
        t1 = t + R.record_increment(t)
        u, tt, _ = O.run(od, u, t, 10 ** 6, CFL, order=3, t_end=t1)
        t = t1
        out.append((t, u))
    return out


def _check(got, want, dsplit=False):
    assert len(got) == len(want)
    for (tg, ug), (tw, uw) in zip(got, want):
        assert abs(tg - tw) <= 1e-15
        if dsplit:  # the record holds ('s', 's') first, the oracle the field last
            ug = np.concatenate([ug[1:], ug[:1]])
        if tw > 0:
            assert rel_l2(ug, uw) < 1e-11


def _dct(tmp_path, method, seed, **kw):
    return dict(homedir=str(tmp_path / "store"), N=16, dim=2, previous="", postfix="", method=method, seed=seed, **kw)


def test_relaxation_records_match_oracle(R, tmp_path, monkeypatch):
    say("runpython lvp")
    from Plasticity.FieldInitializers import FieldInitializer
    monkeypatch.chdir(tmp_path)
    out = R.simulation(_dct(tmp_path, "lvp", 1), end_time=0.02, verbose=False)
    got = records(out, 9, 16, 2)
    beta0 = FieldInitializer.GaussianRandomInitializer((16, 16), 0.2, 1).GetOrderParameterField().numpy()
    want = oracle_intervals(R, O.Dynamics(O.GLIDE_ONLY, 0.0, MU, NU), beta0, 0.02)
    _check(got, want)
    assert os.path.isfile("lvp2d16_s1_.conf") and os.path.isfile("lvp2d16_s1_.ics")


def test_strain_loaded_run_matches_oracle(R, tmp_path, monkeypatch):
    """'gcd' with a load: UpwindLoadingBetaPDynamics, ExternalStrain of
    load_start + load_dir t (Lambda lands on coreEnergy, 0, as in the
    reference)."""
    say("runpython gcd loaded")
    from Plasticity.FieldInitializers import FieldInitializer
    monkeypatch.chdir(tmp_path)
    out = R.simulation(_dct(tmp_path, "gcd", 2, load=True, load_dir=[0.0, 0.0, 0.5], load_start=[0.001, 0.0, 0.0],
                            load_type="strain"), end_time=0.01, verbose=False)
    got = records(out, 9, 16, 2)
    beta0 = FieldInitializer.GaussianRandomInitializer((16, 16), 0.2, 2).GetOrderParameterField().numpy()
    load = O.Load(kind="strain", initial=(0.001, 0.0, 0.0), rate=(0.0, 0.0, 0.5), elastic=(0.5, 0.3))
    want = oracle_intervals(R, O.Dynamics(O.GLIDE_CLIMB, 0.0, MU, NU, load=load), beta0, 0.01)
    _check(got, want)


def test_disorder_run_matches_oracle(R, tmp_path, monkeypatch):
    """'lvpx': the DisorderState of RunPython.py:196-214 (disorder field
    from the continued generator), advected with the state."""
    say("runpython lvpx")
    from Plasticity.FieldInitializers import FieldInitializer
    monkeypatch.chdir(tmp_path)
    out = R.simulation(_dct(tmp_path, "lvpx", 3), end_time=0.01, verbose=False)
    got = records(out, 10, 16, 2)
    beta0 = FieldInitializer.GaussianRandomInitializer((16, 16), 0.2, 3).GetOrderParameterField().numpy()
    d = R.disorder_field((16, 16), 3)
    assert np.array_equal(got[0][1][0], d) and np.array_equal(got[0][1][1:], beta0)
    want = oracle_intervals(R, O.DisorderDynamics(O.GLIDE_ONLY, 0.0, MU, NU), np.concatenate([beta0, d[None]]), 0.01)
    _check(got, want, dsplit=True)
