"""Per-step observers (Plasticity/Observers/Observer.py:7-104)."""
import sys

import numpy as np


class Observer:
    def __init__(self):
        pass

    def Update(self, time=None, state=None):
        pass


class RecallStateObserver(Observer):
    """Keeps the latest state (:15-20)."""

    def __init__(self):
        self.state = None

    def Update(self, time, state):
        self.state = state


class RecordStateObserver(Observer):
    """Pickles (time, state) after every step, protocol 2 (:23-48); read
    back with FieldInitializer.LoadState.  The state's field is pickled as
    host arrays per component, the reference's layout."""

    def __init__(self, filename=None, mode='wb'):
        self.file = open(filename, mode) if filename is not None else sys.stdout.buffer

    def Update(self, time, state):
        import pickle
        pickle.dump(time, self.file, protocol=2)
        pickle.dump(state, self.file, protocol=2)
        self.file.flush()


class RecordStateObserverRaw(Observer):
    """Appends [time, 9 components] as raw float64 (:51-69)."""

    def __init__(self, filename=None, mode='wb'):
        self.file = open(filename, mode) if filename is not None else sys.stdout.buffer

    def Update(self, time, state):
        self.file.write(np.array([time], dtype=np.float64).tobytes())
        self.file.write(np.ascontiguousarray(state.GetOrderParameterField().numpy()).tobytes())
        self.file.flush()


class VerboseTimestepObserver(Observer):
    def Update(self, time, state):
        print("t = %f" % time)


class EnergyDownhillObserver(Observer):
    """Counts steps where the total elastic energy rose (:89-104)."""

    def __init__(self, verbose=False):
        self.PreviousEnergy = None
        self.countTotal = 0
        self.countBad = 0
        self.verbose = verbose

    def Update(self, time, state):
        energy = state.TotalElasticEnergy()
        self.countTotal += 1
        if self.PreviousEnergy is not None and self.PreviousEnergy < energy:
            self.countBad += 1
        if self.verbose:
            print("%d/%d" % (self.countBad, self.countTotal))
        self.PreviousEnergy = energy


class LargeKInstabilityObserver(Observer):
    """The instability check of LargeKInstabilityChecking_FieldMover
    (Plasticity/FieldMovers/LargeKInstabilityChecking_FieldMover.py:73-89):
    |FFT(|rho|)| at the mode (N/2, ..., N/2), relative to its value at the
    first update, must stay below `limit`; ``stop`` turns True (and the
    offending state is kept in ``state``) once it does not.  Evaluated on the
    GPU (cdd_instability)."""

    def __init__(self, limit=1.e5):
        self.limit = limit
        self.initialChecking = None
        self.stop = False
        self.state = None

    def Update(self, time, state):
        from Plasticity.amd import plans
        beta = state.GetOrderParameterField().data
        p = plans.analysis_plan(state.gridShape, state.mu, state.nu, beta.device)
        value = p.instability(beta)
        if self.initialChecking is None:
            self.initialChecking = value
        elif not (value / self.initialChecking < self.limit):
            self.stop = True
            self.state = state
        return value
