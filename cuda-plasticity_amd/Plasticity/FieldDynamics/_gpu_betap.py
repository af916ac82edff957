"""GPU BetaP dynamics shared by the reference's three BetaP modules.

CalculateFlux(time, state, CFLCondition) has the signature and results of
CentralUpwindHJBetaPDynamics3D.BetaPDynamics.CalculateFlux
(Plasticity/FieldDynamics/CentralUpwindHJBetaPDynamics3D.py:214-221,
-> CentralUpwindHJ3D.py:63-271): rhs as a TensorField, plus the CFL speed
field a_tot + b_tot (+ c_tot) when CFLCondition is true.  The whole flux,
including the stress solve, runs in the HIP library.

The reference's loading subclasses override GetSigma in Python
(RunPython.py:35-66, MattTools/Run/RunLoadingMPI.py:62-83).  A uniform
load enters the GPU stress solve as its k = 0 mode, so loads are declared
as data (``load_*`` keyword arguments) instead; a subclass that overrides
GetSigma with arbitrary Python cannot run on the fused GPU path and is
refused.
"""
from Plasticity.FieldDynamics.FieldDynamics import FieldDynamics
from Plasticity.Fields.Fields import TensorField
from Plasticity.amd import lib as L
from Plasticity.amd import plans


class GpuBetaPDynamics(FieldDynamics):
    dynamics_kind = L.GLIDE_CLIMB

    def __init__(self, Dx=None, Lambda=0, coreEnergy=0, coreEnergyLog=0, load_kind=L.LOAD_NONE,
                 load_init=None, load_rate=None, load_t0=0.0, load_init_strain=0.0, load_uniaxial_rate=0.0,
                 load_mu=0.0, load_nu=0.0):
        if Dx is not None:
            raise NotImplementedError("Dx = 1/N_last (the reference default) is the only spacing on the GPU path")
        if coreEnergy or coreEnergyLog:
            raise NotImplementedError("core energy terms are unused by the reference BetaP dynamics")
        self.Dx = Dx
        self.Lambda = Lambda
        self.coreEnergy = coreEnergy
        self.coreEnergyLog = coreEnergyLog
        self.load = dict(load_kind=load_kind, load_init=tuple(load_init or (0.,) * 9),
                         load_rate=tuple(load_rate or (0.,) * 9), load_t0=load_t0,
                         load_init_strain=load_init_strain, load_uniaxial_rate=load_uniaxial_rate,
                         load_mu=float(load_mu), load_nu=float(load_nu))

    def _check_overrides(self):
        if type(self).GetSigma is not GpuBetaPDynamics.GetSigma:
            raise NotImplementedError(
                "%s overrides GetSigma in Python; declare the load with load_* arguments instead"
                % type(self).__name__)

    def plan_kwargs(self):
        self._check_overrides()
        kw = dict(dynamics=self.dynamics_kind,
                  Lambda=float(self.Lambda) if self.dynamics_kind != L.GLIDE_ONLY else 0.0)
        kw.update(self.load)
        return kw

    def _plan(self, state, **extra):
        n = state.gridShape[0]
        if any(s != n for s in state.gridShape):
            raise ValueError("the GPU stepper supports square grids only")
        kw = self.plan_kwargs()
        kw.update(extra)
        return plans.get_plan(len(state.gridShape), n, state.betaP.data.device, mu=state.mu, nu=state.nu, **kw)

    def _private_plan(self, state, **extra):
        """A plan owned by this dynamics object, outside the shared plan cache,
        for couplings whose fields live on the plan (disorder)."""
        n = state.gridShape[0]
        if any(s != n for s in state.gridShape):
            raise ValueError("the GPU stepper supports square grids only")
        kw = self.plan_kwargs()
        kw.update(extra)
        dev = state.betaP.data.device
        key = (len(state.gridShape), n, str(dev), state.mu, state.nu) + tuple(
            sorted((k, tuple(v) if isinstance(v, (list, tuple)) else v) for k, v in kw.items()))
        own = self.__dict__.setdefault("_private_plans", {})
        p = own.get(key)
        if p is None:
            p = L.Plan(len(state.gridShape), n, device=dev.index or 0, mu=state.mu, nu=state.nu, **kw)
            own[key] = p
        p.bind_stream()
        return p

    def GetSigma(self, state, time=0.0):
        """BetaPDynamics.GetSigma = state.CalculateSigma(): a DisorderState's
        includes its disorder term (DisorderState.py:15-30)."""
        p = self._plan(state)
        d = _disorder_field(state)
        if d is not None:
            return TensorField(state.gridShape, data=p.disorder_sigma(state.betaP.data, d, time))
        return TensorField(state.gridShape, data=p.sigma(state.betaP.data, time))

    def CalculateFlux(self, time, state, CFLCondition=False):
        """For a DisorderState the flux covers all ten components of its
        field: the disorder field is carried with H_s = v . grad d (H_2D/H_3D
        apply to every component, CentralUpwindHJBetaPDynamics3D.py:186-188)."""
        p = self._plan(state)
        d = _disorder_field(state)
        if d is not None:
            import torch
            from Plasticity.Fields.Fields import TensorScalarMixedField
            rhs, rd, speed = p.disorder_flux(state.betaP.data, d, time, advect=True)
            rhs = TensorScalarMixedField(state.gridShape, data=torch.cat([rd[None], rhs]))
        else:
            rhs, speed = p.flux(state.betaP.data, time)
            rhs = TensorField(state.gridShape, data=rhs)
        if CFLCondition:
            return rhs, speed
        return rhs


def _disorder_field(state):
    from Plasticity.PlasticityStates.DisorderState import DisorderState
    return state.DisorderField() if isinstance(state, DisorderState) else None


class GpuGlideOnlyDynamics(GpuBetaPDynamics):
    dynamics_kind = L.GLIDE_ONLY

    def __init__(self, Dx=None, coreEnergy=0, coreEnergyLog=0, **load):
        GpuBetaPDynamics.__init__(self, Dx, 0, coreEnergy, coreEnergyLog, **load)
