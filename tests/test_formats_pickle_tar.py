"""Pickled state records (Observer.RecordStateObserver / FieldInitializer.
LoadState, Observers/Observer.py:23-48, FieldInitializers/FieldInitializer.py:
29-59) and run archives (TarFile.py: .conf + .ics + .plas in a tar) -- host
only."""
import copyreg
import io
import json
import os
import pickle
import tarfile

import numpy as np
import pytest
import torch

from Plasticity.Fields.Fields import TensorField, TENSOR_COMPONENTS
from Plasticity.FieldInitializers import FieldInitializer as FI
from Plasticity.Observers.Observer import RecordStateObserver
from Plasticity.PlasticityStates.PlasticityState import PlasticityState
from Plasticity.PlasticityStates.VacancyState import VacancyState
from Plasticity import TarFile

CPU = torch.device("cpu")


def _state(seed, n=8, dim=2, cls=PlasticityState, ncomp=9):
    rng = np.random.default_rng(seed)
    return cls((n,) * dim, field=rng.standard_normal((ncomp,) + (n,) * dim), mu=0.7, nu=0.25, device=CPU)


def test_pickle_records_round_trip(tmp_path):
    fn = str(tmp_path / "run.save")
    obs = RecordStateObserver(fn)
    states = [_state(k) for k in range(3)]
    for k, s in enumerate(states):
        obs.Update(0.1 * (k + 1), s)
    obs.file.close()
    t, s = FI.LoadState(fn)
    assert t == pytest.approx(0.3)
    assert np.array_equal(s.GetOrderParameterField().numpy(), states[2].GetOrderParameterField().numpy())
    assert (s.mu, s.nu, s.gridShape) == (0.7, 0.25, (8, 8))
    t, s = FI.LoadState(fn, time=0.15)
    assert t == pytest.approx(0.2)
    assert np.array_equal(s.GetOrderParameterField().numpy(), states[1].GetOrderParameterField().numpy())
    with pytest.raises(ValueError):
        FI.LoadState(fn, time=1.0)


def test_vacancy_state_pickle_keeps_the_concentration(tmp_path):
    fn = str(tmp_path / "vac.save")
    s0 = VacancyState((8, 8), field=np.random.default_rng(4).standard_normal((10, 8, 8)), alpha=0.4, device=CPU)
    obs = RecordStateObserver(fn)
    obs.Update(1.5, s0)
    obs.file.close()
    t, s = FI.LoadState(fn)
    assert isinstance(s, VacancyState) and s.alpha == 0.4 and t == 1.5
    assert np.array_equal(s.GetOrderParameterField().numpy(), s0.GetOrderParameterField().numpy())
    assert np.array_equal(s.betaP.numpy(), s0.betaP.numpy())


def test_state_pickles_may_not_name_other_globals(tmp_path):
    fn = str(tmp_path / "evil.save")

    class Evil:
        def __reduce__(self):
            return (os.getcwd, ())
    with open(fn, "wb") as f:
        pickle.dump(0.0, f, protocol=2)
        pickle.dump(Evil(), f, protocol=2)
    with pytest.raises(pickle.UnpicklingError):
        FI.LoadState(fn)


def test_reference_layout_state_pickle_loads(tmp_path, monkeypatch):
    """A synthetic pickle with the reference's structure (no reference-written
    file is available): NEWOBJ of the state class, then its __dict__ (the
    reference's PlasticityState dict, PlasticityState.py:41-48, with 'field'
    a Field whose __dict__ is gridShape, components, data per component,
    Fields.py:17-27); extra keys such as 'dimension' and 'kspace' are
    tolerated."""
    rng = np.random.default_rng(9)
    arr = rng.standard_normal((9, 8, 8))
    comps = [tuple(c) for c in TENSOR_COMPONENTS]
    fdict = {"gridShape": (8, 8), "components": list(comps), "kspace": False,
             "data": {c: arr[k] for k, c in enumerate(comps)}}
    f = TensorField((8, 8), data=torch.zeros((9, 8, 8), dtype=torch.float64))
    sdict = {"gridShape": (8, 8), "dimension": 2, "nu": 0.3, "mu": 0.5, "field": f}
    monkeypatch.setattr(TensorField, "__reduce_ex__", lambda self, proto: (copyreg.__newobj__, (TensorField,), fdict))
    monkeypatch.setattr(PlasticityState, "__reduce_ex__",
                        lambda self, proto: (copyreg.__newobj__, (PlasticityState,), sdict))
    s0 = _state(1)
    fn = str(tmp_path / "ref.save")
    with open(fn, "wb") as fo:
        pickle.dump(2.5, fo, protocol=2)
        pickle.dump(s0, fo, protocol=2)
    monkeypatch.undo()
    t, s = FI.LoadState(fn)
    assert t == 2.5 and s.gridShape == (8, 8) and s.mu == 0.5 and s.nu == 0.3
    assert np.array_equal(s.GetOrderParameterField().numpy(), arr)


def _reversed(a):
    """The reference's record layout: every component with its grid axes
    reversed (RunCUDA.py:9-22)."""
    return np.stack([c.transpose() for c in a])


def test_run_archive(tmp_path):
    n, dim = 8, 2
    rng = np.random.default_rng(12)
    ics = rng.standard_normal((9, n, n))
    recs = [(0.5, rng.standard_normal((9, n, n))), (1.0, rng.standard_normal((9, n, n)))]
    run = tmp_path / "lvp2d8_s0"
    run.mkdir()
    conf = {"N": n, "dim": dim, "method": "lvp", "seed": 0}
    (run / "lvp2d8_s0.conf").write_text(json.dumps(conf))
    _reversed(ics).astype(np.float64).tofile(str(run / "lvp2d8_s0.ics"))
    with open(run / "lvp2d8_s0.plas", "wb") as f:
        for t, a in recs:
            f.write(np.array([t]).tobytes())
            f.write(_reversed(a).astype(np.float64).tobytes())
    tarname = str(tmp_path / "lvp2d8_s0.tar")
    with tarfile.open(tarname, "w") as tar:
        tar.add(str(run), arcname=run.name)
    assert TarFile.LoadTarJSON(tarname) == conf
    t, s = TarFile.LoadTarICS(tarname, device=CPU)
    assert t == 0.0 and np.array_equal(s.GetOrderParameterField().numpy(), ics)
    t, s = TarFile.LoadTarState(tarname, device=CPU)
    assert t == 1.0 and np.array_equal(s.GetOrderParameterField().numpy(), recs[1][1])
    t, s = TarFile.LoadTarState(tarname, time=0.2, device=CPU)
    assert t == 0.5 and np.array_equal(s.GetOrderParameterField().numpy(), recs[0][1])


def test_run_archive_refuses_unsafe_member_names(tmp_path):
    """A crafted archive whose record member climbs out of the target
    directory is refused before anything is written."""
    import io
    import tarfile
    from Plasticity.TarFile import RunArchive, tar_extract
    path = tmp_path / "evil.tar"
    with tarfile.open(path, "w") as tar:
        for name, data in (("../evil.plas", b"x" * 16), ("run/run.conf", b'{"N": 4, "dim": 2}')):
            info = tarfile.TarInfo(name)
            info.size = len(data)
            tar.addfile(info, io.BytesIO(data))
    dest = tmp_path / "sub"
    dest.mkdir()
    with tarfile.open(path) as tar:
        with pytest.raises(ValueError, match="unsafe"):
            tar_extract(tar, ".plas", path=str(dest))
    assert not (tmp_path / "evil.plas").exists()
    with RunArchive(str(path)) as arch:
        with pytest.raises(ValueError):
            arch.copy_to(".plas", str(dest / "x.plas"))
        with pytest.raises(KeyError, match=".ics"):
            arch.member(".ics")
        assert arch.configuration() == {"N": 4, "dim": 2}
