"""torch.autograd interface to the hand-written adjoint.

``run_with_grad(beta0, Lambda, nsteps, ...)`` advances beta0 by nsteps TVD-RK
steps on the GPU and returns the final state as a tensor that autograd can
differentiate with respect to beta0 and Lambda (the climb-suppression
parameter; climb mobility = 1 - Lambda).  ``elastic_energy(beta)`` is the
total elastic energy with its hand-written gradient.

Forward: the time steps are chosen by the CFL rule and recorded on the
device; a checkpoint of the state is kept every ``checkpoint_every`` steps.
Backward: segments are processed last to first by cdd_segment_vjp: each is
recomputed from its checkpoint replaying the recorded time steps exactly
(the adjoint treats the schedule as frozen), keeping every stage state and
its stress planes, and the cotangent is carried back through it stage by
stage on the device with no host synchronisation.
"""
import torch

from . import lib as L


def _plan(n, d, dynamics, rk_order, Lambda, mu, nu, cfl, load, device):
    kw = dict(dynamics=dynamics, rk_order=rk_order, Lambda=float(Lambda), mu=mu, nu=nu, cfl=cfl, device=device)
    kw.update(load or {})
    return L.Plan(d, n, **kw)


class _Run(torch.autograd.Function):
    @staticmethod
    def forward(ctx, beta0, Lambda, nsteps, t0, ckpt, opts):
        d = beta0.dim() - 1
        n = beta0.shape[1]
        plan = _plan(n, d, opts["dynamics"], opts["rk_order"], float(Lambda), opts["mu"], opts["nu"], opts["cfl"],
                     opts.get("load"), beta0.device.index or 0)
        log = torch.zeros(nsteps, dtype=torch.float64, device=beta0.device)
        plan.set_state(beta0.detach())
        plan.set_time(float(t0))
        plan.dt_record(log)
        checkpoints = [beta0.detach().clone()]
        done_steps = 0
        while done_steps < nsteps:
            k = min(ckpt, nsteps - done_steps)
            plan.step(k)
            done_steps += k
            if done_steps < nsteps:
                checkpoints.append(plan.get_state())
        out = plan.get_state()
        plan.dt_record(None)
        ctx.plan = plan
        ctx.log = log
        ctx.checkpoints = checkpoints
        ctx.nsteps, ctx.t0, ctx.ckpt = nsteps, float(t0), ckpt
        return out

    @staticmethod
    def backward(ctx, gout):
        plan, log, cps = ctx.plan, ctx.log, ctx.checkpoints
        dts = log.cpu().tolist()
        times = [ctx.t0]
        for dt in dts[:-1]:
            times.append(times[-1] + dt)
        g = gout.detach().contiguous().clone()
        plan.lambda_bar(reset=True)
        for s in reversed(range(len(cps))):
            first = s * ctx.ckpt
            last = min(first + ctx.ckpt, ctx.nsteps)
            # recompute the segment from its checkpoint replaying the recorded
            # steps, then carry the cotangent back through it (in place)
            plan.segment_vjp(cps[s], times[first], log[first:last], g, out=g)
        lam_grad = torch.tensor(plan.lambda_bar(reset=True), dtype=torch.float64, device=gout.device)
        return g, lam_grad, None, None, None, None


def run_with_grad(beta0, Lambda, nsteps, t0=0.0, checkpoint_every=20, dynamics=L.GLIDE_CLIMB, rk_order=2, mu=1.0,
                  nu=0.3, cfl=0.25, load=None):
    """Final state after nsteps steps; differentiable w.r.t. beta0 and Lambda
    (a 0-dim float64 tensor).  load: optional dict of Plan load_* arguments
    (loads that do not depend on beta)."""
    if not torch.is_tensor(Lambda):
        Lambda = torch.tensor(float(Lambda), dtype=torch.float64, device=beta0.device)
    opts = dict(dynamics=dynamics, rk_order=rk_order, mu=mu, nu=nu, cfl=cfl, load=load)
    return _Run.apply(beta0, Lambda, int(nsteps), float(t0), int(checkpoint_every), opts)


class _Energy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, beta, t, opts):
        d = beta.dim() - 1
        n = beta.shape[1]
        plan = _plan(n, d, opts["dynamics"], opts["rk_order"], opts["Lambda"], opts["mu"], opts["nu"], opts["cfl"],
                     opts.get("load"), beta.device.index or 0)
        g, e = plan.energy_grad(beta.detach(), float(t))
        ctx.save_for_backward(g)
        return torch.tensor(e, dtype=torch.float64, device=beta.device)

    @staticmethod
    def backward(ctx, ge):
        (g,) = ctx.saved_tensors
        return ge * g, None, None


def elastic_energy(beta, t=0.0, mu=1.0, nu=0.3, load=None):
    """Total elastic energy sum_cells 1/2 S sigma sigma (PlasticityState.py:
    317-330, Observer.py:86) with the hand-written gradient."""
    opts = dict(dynamics=L.GLIDE_CLIMB, rk_order=2, Lambda=0.0, mu=mu, nu=nu, cfl=0.25, load=load)
    return _Energy.apply(beta, float(t), opts)
