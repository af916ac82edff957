"""The C ABI's .plas append and resume (native output format of the
reference, plasticity.cu:102-160,239,270) against the host-side layout."""
import os
import sys

import numpy as np
import pytest

from Plasticity.FieldInitializers.FieldInitializer import LoadStateRaw, SpectralGaussianField

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


def _cuda_layout(field, dtype):
    return np.stack([field[c].T for c in range(9)]).astype(dtype)


@pytest.mark.parametrize("d,n,prec", [(2, 64, "fp64"), (3, 16, "fp64"), (3, 32, "fp64"), (2, 8, "fp32"),
                                      (3, 32, "fp32")])
def test_plas_append_and_resume(L, tmp_path, d, n, prec):
    say("plas", d, n, prec)
    dt = np.float32 if prec == "fp32" else np.float64
    beta = SpectralGaussianField((n,) * d, 2.0, 0.01, seed=20)
    p = L.Plan(d, n, dynamics=0, rk_order=3, Lambda=0.9, precision=prec)
    p.set_state(beta)
    p.set_time(0.0)
    path = str(tmp_path / "run.plas")
    recs = []
    for _ in range(2):
        p.step(2)
        t = p.clock()[0]
        p.plas_append(path)
        recs.append((t, p.get_state().cpu().numpy()))
    want = b"".join(np.array([t], dtype=dt).tobytes() + _cuda_layout(s, dt).tobytes() for t, s in recs)
    with open(path, "rb") as f:
        assert f.read() == want
    if prec == "fp64":
        t, st = LoadStateRaw(path, n, d, device="cpu")
        assert t == recs[-1][0] and np.array_equal(st.GetOrderParameterField().numpy(), recs[-1][1])
    # an interrupted append leaves a torn tail: resume uses the last whole record
    with open(path, "ab") as f:
        f.write(b"\x00" * 100)
    q = L.Plan(d, n, dynamics=0, rk_order=3, Lambda=0.9, precision=prec)
    t, nrec, torn = q.plas_resume(path)
    assert nrec == 2 and torn
    assert t == pytest.approx(recs[-1][0], rel=1e-6 if prec == "fp32" else 0)
    assert np.array_equal(q.get_state().cpu().numpy(), recs[-1][1])
    assert q.clock()[0] == t
    # stepping on from the resumed state equals stepping on from the original
    p.step(1)
    q.step(1)
    if prec == "fp64":
        assert np.array_equal(q.get_state().cpu().numpy(), p.get_state().cpu().numpy())


def test_plas_resume_rejects_wrong_size(L, tmp_path):
    path = str(tmp_path / "bad.plas")
    np.zeros(16).tofile(path)
    q = L.Plan(2, 16)
    with pytest.raises(L.CDDError):
        q.plas_resume(path)
