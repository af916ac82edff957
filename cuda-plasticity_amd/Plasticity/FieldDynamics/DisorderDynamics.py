"""BetaP dynamics with a fixed disorder field
(Plasticity/FieldDynamics/DisorderDynamics.py:17-44): the stress seen by the
flux gains the disorder term of `disorder` (DisorderState.CalculateSigma's,
with d held by the dynamics instead of the state), every stage, on the GPU."""
import numpy as np

from Plasticity.FieldDynamics.CentralUpwindHJBetaPDynamics3D import BetaPDynamics
from Plasticity.Fields.Fields import TensorField


class BetaP_DisorderDynamics(BetaPDynamics):
    def __init__(self, Dx=None, Lambda=0, coreEnergy=0, coreEnergyLog=0, disorder=None, **load):
        BetaPDynamics.__init__(self, Dx, Lambda, coreEnergy, coreEnergyLog, **load)
        self.disorder = disorder

    def _check_overrides(self):
        if type(self).GetSigma is not BetaP_DisorderDynamics.GetSigma:
            raise NotImplementedError(
                "%s overrides GetSigma in Python; declare the load with load_* arguments instead"
                % type(self).__name__)

    def _field(self, state):
        import torch
        if self.disorder is None:
            raise ValueError("BetaP_DisorderDynamics needs a disorder field")
        d = self.disorder
        if not torch.is_tensor(d):
            d = torch.from_numpy(np.ascontiguousarray(np.asarray(d, dtype=np.float64)))
        return d.to(device=state.betaP.data.device, dtype=torch.float64)

    def _plan(self, state, **extra):
        """A plan of this dynamics object with its disorder field set (kept
        out of the shared plan cache: the field lives on the plan)."""
        from Plasticity.PlasticityStates.DisorderState import DisorderState
        if isinstance(state, DisorderState):
            raise NotImplementedError("a DisorderState carries its own field; use the plain BetaP dynamics")
        p = self._private_plan(state, **extra)
        p.set_disorder(self._field(state), advect=False)
        return p

    def GetSigma(self, state, time=0.0):
        p = self._plan(state)
        return TensorField(state.gridShape, data=p.disorder_sigma(state.betaP.data, self._field(state), time))

    def CalculateFlux(self, time, state, CFLCondition=False):
        p = self._plan(state)
        rhs, _, speed = p.disorder_flux(state.betaP.data, self._field(state), time, advect=False)
        rhs = TensorField(state.gridShape, data=rhs)
        if CFLCondition:
            return rhs, speed
        return rhs
