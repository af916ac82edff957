"""BetaP dynamics coupled to vacancies (Plasticity/FieldDynamics/
VacancyDynamics.py:16-70): glide + climb with climb suppression Lambda, the
stress seen by the flux shifted by the vacancies' chemical stress
sigma_ii -= alpha c.  The vacancy update of a step (c += beta (tr beta^P'
- tr beta^P), then diffusion exp(-gamma dt k^2)) is part of the GPU step of
FieldMovers.FieldMover.OperatorSplittingTVDRK_FieldMover."""
from Plasticity.FieldDynamics._gpu_betap import GpuBetaPDynamics
from Plasticity.Fields.Fields import TensorField
from Plasticity.amd import lib as L


class BetaP_VacancyDynamics(GpuBetaPDynamics):
    dynamics_kind = L.GLIDE_CLIMB

    def __init__(self, Dx=None, Lambda=0, coreEnergy=0, coreEnergyLog=0, gamma=1., alpha=1., beta=1., **load):
        GpuBetaPDynamics.__init__(self, Dx, Lambda, coreEnergy, coreEnergyLog, **load)
        self.gamma = gamma
        self.alpha = alpha
        self.beta = beta
        self._plans = {}

    def _check_overrides(self):
        if type(self).GetSigma is not BetaP_VacancyDynamics.GetSigma:
            raise NotImplementedError(
                "%s overrides GetSigma in Python; declare the load with load_* arguments instead"
                % type(self).__name__)

    def _vacancy_plan(self, state, cfl=0.25, rk_order=3, dt_bound=None):
        """A plan of this dynamics with the vacancy coupling on (one per grid,
        device, elastic constants and stepping; not shared with other users,
        since it holds the concentration field)."""
        n = state.gridShape[0]
        if any(s != n for s in state.gridShape):
            raise ValueError("the GPU stepper supports square grids only")
        dev = state.betaP.data.device
        key = (len(state.gridShape), n, str(dev), state.mu, state.nu, cfl, rk_order, dt_bound)
        p = self._plans.get(key)
        if p is None:
            kw = self.plan_kwargs()
            p = L.Plan(len(state.gridShape), n, device=dev.index or 0, mu=state.mu, nu=state.nu, cfl=cfl,
                       rk_order=rk_order, **kw)
            self._plans[key] = p
        import numpy as np
        import torch
        a = self.alpha
        field = torch.is_tensor(a) or (isinstance(a, np.ndarray) and a.ndim > 0)
        p.set_vacancy(1.0 if field else float(a), self.beta, self.gamma, dt_bound)
        p.vacancy_alpha(a if field else None)  # alpha may be a field, as alpha * cfield allows
        p.bind_stream()
        return p

    def GetSigma(self, state, time, cfield):
        """The state's stress with sigma_ii -= alpha c (:24-28)."""
        p = self._vacancy_plan(state)
        sig = p.sigma(state.betaP.data, time)
        a = self.alpha
        if not isinstance(a, (int, float)):
            import torch
            a = torch.as_tensor(a, dtype=torch.float64, device=sig.device)
        for k in (0, 4, 8):
            sig[k] -= a * cfield
        return TensorField(state.gridShape, data=sig)

    def CalculateFlux(self, time, betaPstate, cfield, CFLCondition=False):
        """The central-upwind flux with that stress (:30-35)."""
        p = self._vacancy_plan(betaPstate)
        p.vacancy_set(cfield)
        rhs, speed = p.flux(betaPstate.betaP.data, time)
        rhs = TensorField(betaPstate.gridShape, data=rhs)
        if CFLCondition:
            return rhs, speed
        return rhs
