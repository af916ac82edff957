"""Branch-frozen replay of a TVD-RK run (for checking gradients).

The scheme is only piecewise smooth: the minmod limiter picks one of three
differences or zero at every site (CentralUpwindHJ3D.py:11-34) and the
one-sided speeds are maxima and minima over the corners and zero
(CentralUpwindHJ3D.py:189-241), and the CFL time step is a global maximum
(FieldMover.py:94-96).  The adjoint differentiates the map with all of these
held at the choices of the run it linearises.  A FrozenRun records those
choices along a base run (every stage of every step, cdd_flux_frozen with
record) and replays them for perturbed inputs, with the recorded time steps:
a smooth map whose central differences converge to the adjoint's directional
derivatives as eps^2.

tangent_run is the forward-mode counterpart: the tangent-linear map of a
run (cdd_flux_tangent at every stage), the exact directional derivative
whose transpose the adjoint computes.

Stages follow the plan's RK order: RK2 F0 = u + dt L(u),
u' = u/2 + (F0 + dt L(F0))/2; RK3 (Shu-Osher, FieldMover.py:115-135)
F1 = 3/4 u + 1/4 (F0 + dt L(F0)), u' = u/3 + 2/3 (F1 + dt L(F1)), every
stage at the step's start time.
"""
import torch


class FrozenRun:
    """A base run of `plan` from beta0 with the time steps `dts` (host
    floats, e.g. recorded by Plan.dt_record), its branch choices kept on the
    device (2 words per cell, stage and step)."""

    def __init__(self, plan, beta0, dts, t0=0.0):
        self.plan = plan
        self.dts = [float(x) for x in dts]
        self.t0 = float(t0)
        self.order = int(plan.cfg.rk_order)
        if self.order not in (2, 3):
            raise ValueError("a FrozenRun replays SSP-RK2 or TVD-RK3 steps")
        cells = beta0[0].numel()
        self.sel = torch.empty((len(self.dts), self.order, 2 * cells), dtype=torch.int64, device=plan.dev)
        self.base, self.t_end = self._run(plan, beta0, record=True)

    def _run(self, plan, beta0, record):
        u = plan._as_dev(beta0).clone()
        t = self.t0
        for k, dt in enumerate(self.dts):
            sel = self.sel[k]
            F0 = u + dt * plan.flux_frozen(u, t, sel[0], record)
            if self.order == 2:
                u = 0.5 * u + 0.5 * (F0 + dt * plan.flux_frozen(F0, t, sel[1], record))
            else:
                F1 = 0.75 * u + 0.25 * (F0 + dt * plan.flux_frozen(F0, t, sel[1], record))
                u = u / 3.0 + 2.0 / 3.0 * (F1 + dt * plan.flux_frozen(F1, t, sel[2], record))
            t += dt
        return u, t

    def replay(self, beta0, plan=None):
        """Final state of the branch-frozen map from beta0; `plan` may be
        another plan of the same grid and scheme (e.g. a different Lambda)."""
        return self._run(plan or self.plan, beta0, record=False)[0]


def tangent_run(plan, beta0, dbeta0, dts, t0=0.0, dLambda=0.0, observe=None):
    """Forward-mode differentiation of a run of `plan` from beta0 with the
    time steps `dts`: returns (u, du, t), the final state and its tangent
    along (dbeta0, dLambda) (cdd_flux_tangent at every stage, branches
    decided at the run's own states; the same stage formulas as FrozenRun).
    observe(u, du, t), if given, is called after every step."""
    order = int(plan.cfg.rk_order)
    if order not in (2, 3):
        raise ValueError("tangent_run covers SSP-RK2 and TVD-RK3 steps")
    u = plan._as_dev(beta0).clone()
    du = plan._as_dev(dbeta0).clone()
    t = float(t0)
    for dt in dts:
        dt = float(dt)
        L0, dL0 = plan.flux_tangent(u, du, t, dLambda)
        F0, dF0 = u + dt * L0, du + dt * dL0
        L1, dL1 = plan.flux_tangent(F0, dF0, t, dLambda)
        if order == 2:
            u, du = 0.5 * u + 0.5 * (F0 + dt * L1), 0.5 * du + 0.5 * (dF0 + dt * dL1)
        else:
            F1, dF1 = 0.75 * u + 0.25 * (F0 + dt * L1), 0.75 * du + 0.25 * (dF0 + dt * dL1)
            L2, dL2 = plan.flux_tangent(F1, dF1, t, dLambda)
            u, du = u / 3.0 + 2.0 / 3.0 * (F1 + dt * L2), du / 3.0 + 2.0 / 3.0 * (dF1 + dt * dL2)
        t += dt
        if observe is not None:
            observe(u, du, t)
    return u, du, t


def tangent_stored(plan, beta0, dbeta0, nsteps, t0=0.0, dLambda=0.0, fixed_dt=0.0, segment=20, observe=None):
    """tangent_run along the plan's own trajectory: the plan advances beta0
    by nsteps steps segment by segment (cdd_segment_forward, the stepper's
    own kernels and CFL time steps, every stage state kept), and the tangent
    is carried through each segment at exactly those stage states -- the
    states the adjoint linearises (cdd_segment_vjp_stored).  Returns
    (u, du, t).  observe(j, u, du, t) is called after every step j."""
    import torch
    order = int(plan.cfg.rk_order)
    if order not in (2, 3):
        raise ValueError("tangent_stored covers SSP-RK2 and TVD-RK3 steps")
    plan.set_state(beta0)
    plan.set_time(float(t0), float("inf"), float(fixed_dt))
    du = plan._as_dev(dbeta0).clone()
    shape = tuple(du.shape)
    t = float(t0)
    done = 0
    while done < nsteps:
        k = min(segment, nsteps - done)
        store = torch.empty(plan.segment_store_bytes(k), dtype=torch.uint8, device=plan.dev)
        log = torch.zeros(k, dtype=torch.float64, device=plan.dev)
        plan.dt_record(log)
        plan.segment_forward(k, store)
        plan.dt_record(None)
        states = store.view(torch.float64)[:k * order * du.numel()].view((k, order) + shape)
        for j, dt in enumerate(log.cpu().tolist()):
            u, F0 = states[j, 0], states[j, 1]
            _, dL0 = plan.flux_tangent(u, du, t, dLambda)
            dF0 = du + dt * dL0
            _, dL1 = plan.flux_tangent(F0, dF0, t, dLambda)
            if order == 2:
                du = 0.5 * du + 0.5 * (dF0 + dt * dL1)
            else:
                dF1 = 0.75 * du + 0.25 * (dF0 + dt * dL1)
                _, dL2 = plan.flux_tangent(states[j, 2], dF1, t, dLambda)
                du = du / 3.0 + 2.0 / 3.0 * (dF1 + dt * dL2)
            t += dt
            if observe is not None:
                nxt = states[j + 1, 0] if j + 1 < k else plan.get_state()
                observe(done + j, nxt, du, t)
        del states, store
        done += k
    return plan.get_state(), du, t
