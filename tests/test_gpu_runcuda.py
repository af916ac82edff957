"""The CUDA executable in process (Plasticity/RunCUDA.py run_header and
simulation; RunCUDA.py:120-254, plasticity.cu runTest :40-290) against the
oracle's restatement of the executable's run loop: per record interval,
run_cuda up to the interval's end (each interval closes with the step
dt = t_end - t + ME)."""
import os
import sys
import tarfile

import numpy as np
import pytest

from oracle import cdd_oracle as O
from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField, _reverse_axes

pytestmark = pytest.mark.gpu

MU, NU, CFL = 0.5, 0.3, 0.5


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
    from Plasticity import RunCUDA
    return RunCUDA


def read_plas(path, N, dim):
    """All [t, state] records of a .plas file (NumPy component layout)."""
    raw = np.fromfile(path, dtype=np.float64).reshape(-1, 1 + 9 * N ** dim)
    return [(r[0], np.stack([_reverse_axes(c) for c in r[1:].reshape((9,) + (N,) * dim)])) for r in raw]


def oracle_records(od, beta, inc, end, t0=0.0):
    recs, t, u = [(t0, beta)], t0, beta
    while t < end:
        u, t, _ = O.run_cuda(od, u, t, 10 ** 6, CFL, t_end=t + inc)
        recs.append((t, u))
    return recs


def _compare(got, want):
    assert len(got) == len(want)
    for (tg, ug), (tw, uw) in zip(got, want):
        assert abs(tg - tw) <= 1e-13 * max(tw, 1e-300)
        if tw > 0:
            assert rel_l2(ug, uw) < 1e-11


def _relax_header(R, dim, N, method, inc, end, out="run.plas"):
    h = {"blank": "nothing"}
    R.header_size(h, N, dim)
    R.header_dynamics(h, method)
    R.header_files(h, "run.ics", out)
    R.header_times(h, inc, end)
    return h


@pytest.mark.parametrize("dim,N,method", [(2, 16, "mdp"), (2, 32, "lvp"), (3, 16, "gcd")])
def test_run_header_records_match_oracle(R, tmp_path, dim, N, method):
    say("run_header", dim, N, method)
    beta = SpectralGaussianField((N,) * dim, 2.0, 0.01, seed=41)
    od = O.Dynamics(O.GLIDE_ONLY if method == "lvp" else O.GLIDE_CLIMB, 1.0 if method == "mdp" else 0.0, MU, NU)
    _, _, dts = O.run_cuda(od, beta, 0.0, 1, CFL)
    inc, end = 2.5 * dts[0], 7.0 * dts[0]
    np.stack([_reverse_axes(beta[c]) for c in range(9)]).tofile(str(tmp_path / "run.ics"))
    t, u = R.run_header(_relax_header(R, dim, N, method, inc, end), directory=str(tmp_path))
    got = read_plas(str(tmp_path / "run.plas"), N, dim)
    want = oracle_records(od, beta, inc, end)
    _compare(got, want)
    assert abs(t - want[-1][0]) <= 1e-13 * t and rel_l2(u.cpu().numpy(), want[-1][1]) < 1e-11


@pytest.mark.parametrize("dim,N,method", [(2, 16, "gcd"), (3, 16, "lvp")])
def test_run_header_dynamic_nucleation(R, tmp_path, dim, N, method):
    """DYNAMIC_NUCLEATION: the input field is the source beta0dot (every flux
    gains beta0dot / TIME_FINAL), the state starts at zero, the time step is
    capped at 1 / sqrt(max |sigma(beta0dot)|); the t = 0 record holds the
    input field (plasticity.cu:178-201,232-233)."""
    say("run_header nucleation", dim, N, method)
    src = SpectralGaussianField((N,) * dim, 2.0, 0.01, seed=43)
    od = O.Dynamics(O.GLIDE_ONLY if method == "lvp" else O.GLIDE_CLIMB, 0.0, MU, NU)
    cap = O.nucleation_cap(src, MU, NU)
    inc, end = 0.7 * cap, 2.0 * cap
    np.stack([_reverse_axes(src[c]) for c in range(9)]).tofile(str(tmp_path / "run.ics"))
    h = _relax_header(R, dim, N, method, inc, end)
    R.header_dynamic_nucleation(h)
    t, u = R.run_header(h, directory=str(tmp_path))
    got = read_plas(str(tmp_path / "run.plas"), N, dim)
    assert got[0][0] == 0.0 and np.array_equal(got[0][1], src)
    want, tw, uw = [(0.0, src)], 0.0, np.zeros_like(src)
    while tw < end:
        uw, tw, dts = O.run_cuda(od, uw, tw, 10 ** 6, CFL, t_end=tw + inc, source=src, rate=1.0 / end, dt_cap=cap)
        want.append((tw, uw))
    assert any(d == cap for d in dts) or len(want) > 2
    _compare(got, want)
    assert rel_l2(u.cpu().numpy(), want[-1][1]) < 1e-11


@pytest.mark.parametrize("dim,N", [(2, 16), (3, 16)])
def test_run_header_vacancies(R, tmp_path, dim, N):
    """VACANCIES: ten-component .ics input and .plas records (the tensor and
    c); the stress loses vacancycost c on the diagonal, c's rhs is the trace
    of the tensor rhs, c diffuses after every forward Euler step; a resumed
    run continues from the last ten-component record."""
    say("run_header vacancies", dim, N)
    beta = SpectralGaussianField((N,) * dim, 2.0, 0.01, seed=47)
    c = 0.01 * SpectralGaussianField((N,) * dim, 2.0, 1.0, seed=48)[0]
    od = O.Dynamics(O.GLIDE_CLIMB, 1.0, MU, NU)
    vac = O.Vacancy(0.6, 1.0, 0.05)
    _, _, _, dts = O.run_cuda_vacancy(od, vac, beta, c, 0.0, 1, CFL)
    inc = 2.5 * dts[0]
    fields = np.concatenate([beta, c[None]])
    np.stack([_reverse_axes(f) for f in fields]).tofile(str(tmp_path / "run.ics"))
    h = _relax_header(R, dim, N, "mdp", inc, 1.5 * inc)
    R.header_dynamics(h, "vac")
    h.update(vacancycost=0.6, vacancydiffusion=0.05)
    R.run_header(h, directory=str(tmp_path))
    h["TIME_FINAL"] = 3.5 * inc
    t, u = R.run_header(h, directory=str(tmp_path))
    raw = np.fromfile(str(tmp_path / "run.plas"), dtype=np.float64).reshape(-1, 1 + 10 * N ** dim)
    got = [(r[0], np.stack([_reverse_axes(f) for f in r[1:].reshape((10,) + (N,) * dim)])) for r in raw]
    want, tw, uw, cw = [(0.0, fields)], 0.0, beta, c
    while tw < 3.5 * inc:
        uw, cw, tw, _ = O.run_cuda_vacancy(od, vac, uw, cw, tw, 10 ** 6, CFL, t_end=tw + inc)
        want.append((tw, np.concatenate([uw, cw[None]])))
    assert len(got) == len(want) and np.array_equal(got[0][1], fields)
    for (tg, ug), (tw_, uw_) in zip(got[1:], want[1:]):
        assert abs(tg - tw_) <= 1e-13 * tw_
        assert rel_l2(ug[:9], uw_[:9]) < 1e-11 and rel_l2(ug[9], uw_[9]) < 1e-11
    assert rel_l2(u.cpu().numpy(), want[-1][1]) < 1e-11


def test_run_header_continues_an_existing_record_file(R, tmp_path):
    """CONTINUE_RUN: a second run with a later TIME_FINAL resumes from the last
    record and appends; the file equals one uninterrupted run's."""
    say("run_header resume")
    N, dim = 16, 2
    beta = SpectralGaussianField((N,) * dim, 2.0, 0.01, seed=42)
    od = O.Dynamics(O.GLIDE_CLIMB, 1.0, MU, NU)
    _, _, dts = O.run_cuda(od, beta, 0.0, 1, CFL)
    inc = 2.5 * dts[0]
    np.stack([_reverse_axes(beta[c]) for c in range(9)]).tofile(str(tmp_path / "run.ics"))
    R.run_header(_relax_header(R, dim, N, "mdp", inc, 1.5 * inc), directory=str(tmp_path))
    assert len(read_plas(str(tmp_path / "run.plas"), N, dim)) == 3
    R.run_header(_relax_header(R, dim, N, "mdp", inc, 3.5 * inc), directory=str(tmp_path))
    got = read_plas(str(tmp_path / "run.plas"), N, dim)
    _compare(got, oracle_records(od, beta, inc, 3.5 * inc))


def test_simulation_loading_chain(R, tmp_path, monkeypatch):
    """RunCUDA.simulation: a first loading run (postfix '0', from the seeded
    random state) and a continuation (postfix '2') from one of its records
    with LOAD_START = time_start LOADING_RATE; archives stored in homedir."""
    say("simulation chain")
    from Plasticity.FieldInitializers import FieldInitializer
    from Plasticity.TarFile import LoadTarICS, LoadTarState
    monkeypatch.chdir(tmp_path)
    store = tmp_path / "store"
    N, dim, seed, rate = 16, 2, 3, 0.05
    Ld = [[0.0, 0.5, 0.0], [0.5, 0.0, 0.0], [0.0, 0.0, 0.0]]
    beta0 = FieldInitializer.GaussianRandomInitializer((N, N), 0.2, seed).GetOrderParameterField().numpy()
    lam_e = 2 * MU * NU / (1 - 2 * NU)
    S = (2 * MU * np.asarray(Ld) + lam_e * np.trace(Ld) * np.eye(3)).ravel()

    def od(start):
        load = O.Load(kind="tensor", tensor0=tuple(S * start), tensor_rate=tuple(S * rate))
        return O.Dynamics(O.GLIDE_CLIMB, 1.0, MU, NU, load=load)
    _, _, dts = O.run_cuda(od(0.0), beta0, 0.0, 1, CFL)
    ts = 2.5 * dts[0]
    dct = dict(homedir=str(store), N=N, dim=dim, previous="", postfix="0", method="mdp", seed=seed, device=0,
               load_direction=Ld, load_rate=rate, time_step=ts, time_end=2 * ts)
    f1 = R.simulation(dct)
    assert os.path.isfile(store / "mdp2d16" / f1)
    with tarfile.open(f1) as tar:
        exts = sorted(os.path.splitext(m.name)[1] for m in tar.getmembers() if m.isfile())
    assert exts == [".conf", ".h", ".ics", ".json", ".plas"]
    _, ics = LoadTarICS(f1)
    assert np.array_equal(ics.GetOrderParameterField().numpy(), beta0)
    want1 = oracle_records(od(0.0), beta0, ts, 2 * ts)
    t1, st1 = LoadTarState(f1)
    assert abs(t1 - want1[-1][0]) <= 1e-13 * t1
    assert rel_l2(st1.GetOrderParameterField().numpy(), want1[-1][1]) < 1e-11
    # continuation from the first record at or after ts
    dct2 = dict(dct, previous="0", postfix="2", time_start=ts, time_end=ts)
    os.remove(f1)  # fetched back from the store
    f2 = R.simulation(dct2)
    tstart, start = want1[1]
    assert tstart >= ts
    want2 = oracle_records(od(ts * rate), start, ts, ts)
    t2, st2 = LoadTarState(f2)
    assert abs(t2 - want2[-1][0]) <= 1e-13 * t2
    assert rel_l2(st2.GetOrderParameterField().numpy(), want2[-1][1]) < 1e-11
