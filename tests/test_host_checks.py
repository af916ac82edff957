"""Host-side argument checks and mover logic that run before any kernel
launch (no GPU needed)."""
import numpy as np
import pytest
import torch

from Plasticity.amd import lib as L
from Plasticity.FieldMovers.FieldMover import TVDRungeKutta_FieldMover


class _FakePlan:
    """Stands in for a Plan's attributes; its handle is None, so a call that
    got past the checks into the library would fail differently."""
    torch = torch
    dev = torch.device("cuda", 0)
    h = None
    _f64_log = L.Plan._f64_log


@pytest.mark.parametrize("method", ["dt_record", "dt_replay", "energy_record", "stress_record"])
def test_float32_logs_are_refused_before_launch(method):
    fake = _FakePlan()
    shape = (6, 8) if method == "stress_record" else (6,)
    with pytest.raises(L.CDDError, match="float64"):
        getattr(L.Plan, method)(fake, torch.zeros(shape, dtype=torch.float32))


def test_stress_log_shape_and_device_are_checked():
    fake = _FakePlan()
    with pytest.raises(L.CDDError, match="shape"):
        L.Plan._f64_log(fake, torch.zeros((6, 8), dtype=torch.float64), "stress log", shape=(None, 9))
    with pytest.raises(L.CDDError, match="device"):
        L.Plan.stress_record(fake, torch.zeros((6, 8), dtype=torch.float64))
    with pytest.raises(L.CDDError, match="contiguous"):
        L.Plan.dt_record(fake, torch.zeros((4, 2), dtype=torch.float64).t())


class _GuardPlan:
    """A plan whose device guard trips on the second step."""

    def __init__(self):
        self.ns = 0

    def step(self, k):
        self.ns += 1

    def clock(self):
        return 0.1 * min(self.ns, 1), 0.1, self.ns >= 2, min(self.ns, 1)

    def status(self):
        return 1 if self.ns >= 2 else 0


@pytest.mark.parametrize("with_observer", [True, False])
def test_mover_stops_on_device_guard(with_observer):
    seen = []

    class Obs:
        def Update(self, t, state):
            seen.append(t)

    mover = TVDRungeKutta_FieldMover(CFLsafeFactor=0.25)
    with pytest.raises(FloatingPointError, match="non-finite"):
        mover._drive(_GuardPlan(), [Obs()] if with_observer else [], lambda: "state")
    # the observer saw the one good step only; the stopped step is not reported
    assert seen == ([0.1] if with_observer else [])
    assert mover.numberOfSteps == 1
