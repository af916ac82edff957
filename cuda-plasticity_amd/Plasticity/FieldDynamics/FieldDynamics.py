"""Base class of field dynamics (Plasticity/FieldDynamics/FieldDynamics.py:6-14)."""


class FieldDynamics:
    def __init__(self, Lambda=0):
        self.Lambda = Lambda

    def GetSigma(self, state, time=0.0):
        return state.CalculateSigma(time=time)
