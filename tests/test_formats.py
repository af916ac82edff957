"""Reference file formats on the host: .plas records and .ics files in the
reversed component layout (RunCUDA.py:9-22, plasticity.cu:239,270,
assist.h:164-192), LoadStateRaw (FieldInitializer.py:61-101) and the
untransposed RecordStateObserverRaw records (Observer.py:51-69)."""
import numpy as np
import pytest

from Plasticity.FieldInitializers.FieldInitializer import (LoadStateRaw, LoadStateRawUntransposed,
                                                           SpectralGaussianField)
from Plasticity.Observers.Observer import RecordStateObserverRaw
from Plasticity.PlasticityStates.PlasticityState import PlasticityState
from Plasticity.RunCUDA import convertStateToCUDA


def _cuda_layout(field):
    """What the reference's executable writes for one state: component by
    component, the array with all axes reversed (x fastest)."""
    return np.stack([field[c].T for c in range(9)]).astype(np.float64)


@pytest.mark.parametrize("shape", [(8, 8), (4, 4, 4)])
def test_load_state_raw_reads_plas_records(tmp_path, shape):
    N, dim = shape[0], len(shape)
    a = SpectralGaussianField(shape, 2.0, 0.01, seed=1)
    b = SpectralGaussianField(shape, 2.0, 0.01, seed=2)
    path = tmp_path / "run.plas"
    with open(path, "wb") as f:
        for t, fld in ((0.5, a), (1.25, b)):
            np.array([t]).tofile(f)
            _cuda_layout(fld).tofile(f)
    t, st = LoadStateRaw(str(path), N, dim, device="cpu")
    assert t == 1.25 and np.array_equal(st.GetOrderParameterField().numpy(), b)
    t, st = LoadStateRaw(str(path), N, dim, time=0.3, device="cpu")
    assert t == 0.5 and np.array_equal(st.GetOrderParameterField().numpy(), a)


def test_ics_round_trip(tmp_path):
    a = SpectralGaussianField((4, 4, 4), 2.0, 0.01, seed=3)
    path = tmp_path / "init.ics"
    convertStateToCUDA(PlasticityState((4, 4, 4), field=a, device="cpu"), str(path))
    raw = np.fromfile(path, dtype=np.float64)
    assert np.array_equal(raw, _cuda_layout(a).ravel())
    t, st = LoadStateRaw(str(path), 4, 3, hastimes=False, device="cpu")
    assert t == 0.0 and np.array_equal(st.GetOrderParameterField().numpy(), a)


def test_record_state_observer_raw_round_trip(tmp_path):
    a = SpectralGaussianField((8, 8), 2.0, 0.01, seed=4)
    path = tmp_path / "obs.raw"
    obs = RecordStateObserverRaw(str(path))
    obs.Update(0.75, PlasticityState((8, 8), field=a, device="cpu"))
    obs.file.close()
    raw = np.fromfile(path, dtype=np.float64)
    assert raw[0] == 0.75 and np.array_equal(raw[1:], a.ravel())
    t, st = LoadStateRawUntransposed(str(path), 8, 2, device="cpu")
    assert t == 0.75 and np.array_equal(st.GetOrderParameterField().numpy(), a)


def test_incomplete_file_is_rejected(tmp_path):
    path = tmp_path / "short.plas"
    np.zeros(10).tofile(path)
    with pytest.raises(ValueError):
        LoadStateRaw(str(path), 8, 2, device="cpu")
