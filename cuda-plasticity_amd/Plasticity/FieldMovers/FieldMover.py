"""TVD Runge-Kutta field mover on the GPU
(Plasticity/FieldMovers/FieldMover.py:8-141).

Run(state, observers, dynamics, startTime, endTime, fixedTimeStep) advances
the state with dt = CFLsafeFactor / (max(gridShape) max speed) from the
first-stage speeds, clamped to endTime, evaluating every stage at the step's
start time, and calls every observer after every step -- the reference
semantics.  The stepping itself (stress solves, fluxes, the CFL reduction
and the stage combinations) is one device-resident sequence per step in the
HIP library; when there are no observers, steps are launched in chunks and
the host only polls the device clock between chunks.

Like the reference, Run does not modify the state object it is given: the
result reaches the caller through observers (RecallStateObserver) and is
also kept in ``self.lastState``.
"""
from Plasticity.PlasticityStates.PlasticityState import PlasticityState
from Plasticity.Fields.Fields import TensorField


class FieldMover:
    def Run(self, state, observers, dynamics, startTime, endTime, dT):
        raise NotImplementedError


class TVDRungeKutta_FieldMover(FieldMover):
    """order=3: Shu-Osher TVD-RK3 (the reference); order=2: SSP-RK2."""

    def __init__(self, CFLsafeFactor=0.1, dtBound=0.01, order=3, chunk=64):
        self.CFLsafeFactor = CFLsafeFactor
        self.dtBound = dtBound  # unused by the reference too (FieldMover.py:107-108)
        self.order = order
        self.chunk = chunk
        self.lastState = None
        self.numberOfSteps = 0

    def _plan(self, state, dynamics):
        return dynamics._plan(state, cfl=float(self.CFLsafeFactor), rk_order=int(self.order))

    def Run(self, state, observers, dynamics=None, startTime=0., endTime=10., fixedTimeStep=None):
        from Plasticity.PlasticityStates.DisorderState import DisorderState
        if isinstance(state, DisorderState):
            return self._run_disorder(state, observers, dynamics, startTime, endTime, fixedTimeStep)
        p = self._plan(state, dynamics)
        p.set_state(state.betaP.data)
        p.set_time(float(startTime), float(endTime), float(fixedTimeStep) if fixedTimeStep else 0.0)

        def snapshot():
            return PlasticityState(state.gridShape, field=TensorField(state.gridShape, data=p.get_state()),
                                   inherit=state)
        return self._drive(p, observers, snapshot)

    def _drive(self, p, observers, snapshot):
        """Step the plan to its end time, calling every observer after every
        step (chunks of self.chunk steps when there are none).  A non-finite
        speed or right-hand side stops the run on the device with the state
        at that step's start: that raises FloatingPointError here, before
        any observer sees the unadvanced state again."""
        steps0 = p.clock()[3]
        while True:
            p.step(1 if observers else self.chunk)
            t, dt, done, ns = p.clock()
            if p.status():
                self.numberOfSteps = ns - steps0
                raise FloatingPointError("non-finite speed or flux at t = %.17g after %d steps; the state is "
                                         "left at that step's start" % (t, ns - steps0))
            if observers:
                cur = snapshot()
                for observer in observers:
                    observer.Update(t, cur)
            if done:
                break
        self.numberOfSteps = p.clock()[3] - steps0
        self.lastState = snapshot()
        return self.lastState

    def _run_disorder(self, state, observers, dynamics, startTime, endTime, fixedTimeStep):
        """A DisorderState with the plain BetaP dynamics: its disorder field
        joins every stage's stress and is carried by the flux and the RK
        stages like the tensor components (the mover's arithmetic on the
        state's ten-component field, FieldMover.py:115-135)."""
        import torch
        from Plasticity.Fields.Fields import TensorScalarMixedField
        from Plasticity.PlasticityStates.DisorderState import DisorderState
        if not hasattr(dynamics, "_private_plan") or hasattr(dynamics, "disorder"):
            raise NotImplementedError("a DisorderState runs with the plain BetaP dynamics")
        p = dynamics._private_plan(state, cfl=float(self.CFLsafeFactor), rk_order=int(self.order))
        p.set_state(state.betaP.data)
        p.set_disorder(state.DisorderField(), advect=True)
        p.set_time(float(startTime), float(endTime), float(fixedTimeStep) if fixedTimeStep else 0.0)

        def snapshot():
            data = torch.cat([p.disorder_get()[None], p.get_state()])
            return DisorderState(state.gridShape, field=TensorScalarMixedField(state.gridShape, data=data),
                                 inherit=state)
        return self._drive(p, observers, snapshot)


class OperatorSplittingTVDRK_FieldMover(TVDRungeKutta_FieldMover):
    """TVD-RK mover of a VacancyState with BetaP_VacancyDynamics
    (FieldMovers/FieldMover.py:318-380): each step advances beta^P with the
    concentration c of the step start, then c += beta (tr beta^P' -
    tr beta^P) and c diffuses by exp(-gamma dt k^2); the adaptive dt is
    capped at dtBound (here the cap is applied; the plain mover leaves it
    commented out).  The whole step runs on the device (cdd_set_vacancy).
    The reference's fixed-step variant writes back the step-start beta^P
    (:347, its F_2 commented out); here every step keeps the advanced one."""

    def Run(self, state, observers, dynamics=None, startTime=0., endTime=10., fixedTimeStep=None):
        from Plasticity.Fields.Fields import TensorScalarMixedField
        from Plasticity.PlasticityStates.VacancyState import VacancyState
        import torch
        p = dynamics._vacancy_plan(state, cfl=float(self.CFLsafeFactor), rk_order=int(self.order),
                                   dt_bound=float(self.dtBound) if self.dtBound else None)
        betaPstate, c = state.DecoupleState()
        p.set_state(betaPstate.betaP.data)
        p.vacancy_set(c)
        p.set_time(float(startTime), float(endTime), float(fixedTimeStep) if fixedTimeStep else 0.0)

        def snapshot():
            data = torch.cat([p.vacancy_get()[None], p.get_state()])
            return VacancyState(state.gridShape, field=TensorScalarMixedField(state.gridShape, data=data),
                                inherit=state)
        return self._drive(p, observers, snapshot)
