"""torch.autograd interface to the hand-written adjoint.

``run_with_grad(beta0, Lambda, nsteps, ...)`` advances beta0 by nsteps TVD-RK
steps on the GPU and returns the final state as a tensor that autograd can
differentiate with respect to beta0 and Lambda (the climb-suppression
parameter; climb mobility = 1 - Lambda).  ``elastic_energy(beta)`` is the
total elastic energy with its hand-written gradient.

Forward: the time steps are chosen by the CFL rule and recorded on the
device; the run is cut into segments of ``checkpoint_every`` steps.  While
device memory allows (``store``), a segment keeps its whole trajectory (every
stage state and its stress planes, cdd_segment_forward); otherwise only a
checkpoint of its start state.
Backward: segments are processed last to first: a stored segment goes
straight into the reverse pass (cdd_segment_vjp_stored); a checkpointed one
is first recomputed from its checkpoint replaying the recorded time steps
exactly (cdd_segment_vjp; the adjoint treats the schedule as frozen).  The
cotangent is carried back stage by stage on the device with no host
synchronisation; both paths give bit-identical gradients.
"""
import torch

from . import lib as L


def _plan(n, d, dynamics, rk_order, Lambda, mu, nu, cfl, load, device):
    kw = dict(dynamics=dynamics, rk_order=rk_order, Lambda=float(Lambda), mu=mu, nu=nu, cfl=cfl, device=device)
    kw.update(load or {})
    return L.Plan(d, n, **kw)


def _store_budget(device, store):
    """Device bytes the forward may keep in stored segments: store=False keeps
    none, True or "auto" up to 60% of the free device memory (the rest serves
    the recomputation workspace of the other segments), an int that many."""
    if store is False or store is None:
        return 0
    if store is True or store == "auto":
        free, _ = torch.cuda.mem_get_info(device)
        return int(0.6 * free)
    return int(store)


def _forward_segments(plan, nsteps, ckpt, budget, device):
    """Advance the plan by nsteps steps in segments of ckpt steps; each
    segment keeps ("store", trajectory) while the budget allows, else
    ("ckpt", start state)."""
    segs, used, done = [], 0, 0
    while done < nsteps:
        k = min(ckpt, nsteps - done)
        nb = plan.segment_store_bytes(k)
        if used + nb <= budget:
            st = torch.empty(nb, dtype=torch.uint8, device=device)
            plan.segment_forward(k, st)
            used += nb
            segs.append(("store", st))
        else:
            segs.append(("ckpt", plan.get_state()))
            plan.step(k)
        done += k
    return segs


def _backward_segments(plan, segs, log, t0, ckpt, nsteps, g, src=None):
    """Carry the cotangent g (in place) back through every segment, last
    first; each store is released once used."""
    dts = log.cpu().tolist()
    times = [t0]
    for dt in dts[:-1]:
        times.append(times[-1] + dt)
    for s in reversed(range(len(segs))):
        first = s * ckpt
        last = min(first + ckpt, nsteps)
        sv = None if src is None else src[first:last].contiguous()
        kind, obj = segs[s]
        if kind == "store":
            plan.segment_vjp_stored(obj, log[first:last], g, out=g, src=sv)
        else:
            plan.segment_vjp(obj, times[first], log[first:last], g, out=g, src=sv)
        segs[s] = None


class _Run(torch.autograd.Function):
    @staticmethod
    def forward(ctx, beta0, Lambda, nsteps, t0, ckpt, opts):
        d = beta0.dim() - 1
        n = beta0.shape[1]
        plan = _plan(n, d, opts["dynamics"], opts["rk_order"], float(Lambda), opts["mu"], opts["nu"], opts["cfl"],
                     opts.get("load"), beta0.device.index or 0)
        log = torch.zeros(nsteps, dtype=torch.float64, device=beta0.device)
        plan.set_state(beta0.detach())
        plan.set_time(float(t0), float("inf"), float(opts.get("fixed_dt", 0.0)))
        plan.dt_record(log)
        segs = _forward_segments(plan, nsteps, ckpt, _store_budget(beta0.device, opts.get("store", "auto")),
                                 beta0.device)
        out = plan.get_state()
        plan.dt_record(None)
        ctx.plan = plan
        ctx.log = log
        ctx.segs = segs
        ctx.nsteps, ctx.t0, ctx.ckpt = nsteps, float(t0), ckpt
        return out

    @staticmethod
    def backward(ctx, gout):
        plan = ctx.plan
        g = gout.detach().contiguous().clone()
        plan.lambda_bar(reset=True)
        _backward_segments(plan, ctx.segs, ctx.log, ctx.t0, ctx.ckpt, ctx.nsteps, g)
        lam_grad = torch.tensor(plan.lambda_bar(reset=True), dtype=torch.float64, device=gout.device)
        return g, lam_grad, None, None, None, None


def run_with_grad(beta0, Lambda, nsteps, t0=0.0, checkpoint_every=20, dynamics=L.GLIDE_CLIMB, rk_order=2, mu=1.0,
                  nu=0.3, cfl=0.25, load=None, fixed_dt=0.0, store="auto"):
    """Final state after nsteps steps; differentiable w.r.t. beta0 and Lambda
    (a 0-dim float64 tensor).  load: optional dict of Plan load_* arguments
    (every load kind, the uniaxial strain-rate control included).  fixed_dt > 0:
    fixed time steps (CalculateOneFixedTimeStep) instead of the CFL rule.
    store: device memory for stored segment trajectories ("auto", True,
    False or a byte budget); segments beyond it are recomputed."""
    if not torch.is_tensor(Lambda):
        Lambda = torch.tensor(float(Lambda), dtype=torch.float64, device=beta0.device)
    opts = dict(dynamics=dynamics, rk_order=rk_order, mu=mu, nu=nu, cfl=cfl, load=load, fixed_dt=fixed_dt,
                store=store)
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


_COMPS = {"xx": (0, 0), "yy": (1, 1), "zz": (2, 2), "xy": (0, 1), "xz": (0, 2), "yz": (1, 2)}
_SYM6 = {(0, 0): 1, (1, 1): 2, (2, 2): 3, (0, 1): 4, (0, 2): 5, (1, 2): 6}


def _load_shift(load, t, mean_bzz, mu, nu):
    """The uniform stress of a Plan load (load_* keyword dict) at time t, as a
    3x3 nested list (the k = 0 mode the stress solve adds)."""
    lam = 2.0 * mu * nu / (1.0 - 2.0 * nu)
    load = load or {}
    kind = load.get("load_kind", L.LOAD_NONE)
    init = list(load.get("load_init", (0.0,) * 9)) + [0.0] * 9
    rate = list(load.get("load_rate", (0.0,) * 9)) + [0.0] * 9
    t0 = load.get("load_t0", 0.0)
    s = [[0.0] * 3 for _ in range(3)]
    if kind == L.LOAD_STRAIN:
        e = [init[i] + rate[i] * (t - t0) for i in range(3)]
        for i in range(3):
            s[i][i] = lam * sum(e) + 2.0 * mu * e[i]
    elif kind == L.LOAD_STRESS:
        for i in range(3):
            s[i][i] = init[i] + rate[i] * (t - t0)
    elif kind == L.LOAD_UNIAXIAL:
        bzz = load.get("load_uniaxial_rate", 0.0) * t - (mean_bzz - load.get("load_init_strain", 0.0))
        for i in range(3):
            s[i][i] = lam * bzz
        s[2][2] += 2.0 * mu * bzz
    elif kind == L.LOAD_TENSOR:
        for c in range(9):
            s[c // 3][c % 3] = init[c] + rate[c] * (t - t0)
    return s


def _mean_stress_grad(comp, load, mu, nu, cells):
    """d<sigma_ij>/d beta_c at one cell (uniform over cells), a 9-vector:
    <sigma> = shift(<beta_zz>) - lam tr<beta> d - mu (<beta> + <beta>^T)."""
    lam = 2.0 * mu * nu / (1.0 - 2.0 * nu)
    i, j = _COMPS[comp]
    g = [0.0] * 9
    for c in range(9):
        a, b = c // 3, c % 3
        v = 0.0
        if i == j and a == b:
            v -= lam
        if (a, b) == (i, j):
            v -= mu
        if (a, b) == (j, i):
            v -= mu
        if (load or {}).get("load_kind") == L.LOAD_UNIAXIAL and c == 8:
            v -= lam * (i == j) + 2.0 * mu * (i == 2 and j == 2)
        g[c] = v / cells
    return g


class _StressWindow(torch.autograd.Function):
    @staticmethod
    def forward(ctx, beta0, Lambda, nsteps, window, comp, t0, ckpt, opts):
        d = beta0.dim() - 1
        n = beta0.shape[1]
        plan = _plan(n, d, opts["dynamics"], opts["rk_order"], float(Lambda), opts["mu"], opts["nu"], opts["cfl"],
                     opts.get("load"), beta0.device.index or 0)
        log = torch.zeros(nsteps, dtype=torch.float64, device=beta0.device)
        slog = torch.zeros((nsteps, 8), dtype=torch.float64, device=beta0.device)
        plan.set_state(beta0.detach())
        plan.set_time(float(t0), float("inf"), float(opts.get("fixed_dt", 0.0)))
        plan.dt_record(log)
        plan.stress_record(slog)
        segs = _forward_segments(plan, nsteps, ckpt, _store_budget(beta0.device, opts.get("store", "auto")),
                                 beta0.device)
        final = plan.get_state()
        t_end = plan.clock()[0]
        plan.dt_record(None)
        plan.stress_record(None)
        # the window's states u_{n-W+1} .. u_{n-1} were recorded at their step
        # start; the final state's mean stress is the closed form of its k = 0 mode
        col = _SYM6[_COMPS[comp]]
        i, j = _COMPS[comp]
        mb = final.reshape(9, -1).mean(dim=1).tolist()
        lam = 2.0 * opts["mu"] * opts["nu"] / (1.0 - 2.0 * opts["nu"])
        sh = _load_shift(opts.get("load"), t_end, mb[8], opts["mu"], opts["nu"])
        m_final = sh[i][j] - (lam * (mb[0] + mb[4] + mb[8]) if i == j else 0.0) - opts["mu"] * (mb[3 * i + j] + mb[3 * j + i])
        J = (slog[nsteps - window + 1:, col].sum().item() + m_final) / window
        ctx.plan, ctx.log, ctx.segs = plan, log, segs
        ctx.nsteps, ctx.t0, ctx.ckpt, ctx.window = nsteps, float(t0), ckpt, window
        ctx.g = _mean_stress_grad(comp, opts.get("load"), opts["mu"], opts["nu"], final[0].numel())
        ctx.shape = tuple(final.shape)
        return torch.tensor(J, dtype=torch.float64, device=beta0.device)

    @staticmethod
    def backward(ctx, gJ):
        plan, log = ctx.plan, ctx.log
        dev = log.device
        w = gJ.item() / ctx.window
        gvec = torch.tensor(ctx.g, dtype=torch.float64, device=dev) * w
        g = gvec.reshape((9,) + (1,) * (len(ctx.shape) - 1)).expand(ctx.shape).contiguous()
        src = torch.zeros((ctx.nsteps, 9), dtype=torch.float64, device=dev)
        src[ctx.nsteps - ctx.window + 1:] = gvec
        plan.lambda_bar(reset=True)
        _backward_segments(plan, ctx.segs, log, ctx.t0, ctx.ckpt, ctx.nsteps, g, src=src)
        lam_grad = torch.tensor(plan.lambda_bar(reset=True), dtype=torch.float64, device=dev)
        return g, lam_grad, None, None, None, None, None, None


def mean_stress_window(beta0, Lambda, nsteps, window, comp="zz", t0=0.0, checkpoint_every=20,
                       dynamics=L.GLIDE_CLIMB, rk_order=3, mu=1.0, nu=0.3, cfl=0.25, load=None, fixed_dt=0.0,
                       store="auto"):
    """Stress-strain objective: the mean stress component `comp` (external load
    included, as PlottingTools.StrainStress reports it, PlottingTools.py:69-100)
    averaged over the states after each of the last `window` of nsteps steps;
    differentiable w.r.t. beta0 and Lambda through the hand-written adjoint."""
    if not 1 <= window <= nsteps:
        raise ValueError("window must be in [1, nsteps]")
    if comp not in _COMPS:
        raise ValueError("comp must be one of %s" % sorted(_COMPS))
    if not torch.is_tensor(Lambda):
        Lambda = torch.tensor(float(Lambda), dtype=torch.float64, device=beta0.device)
    opts = dict(dynamics=dynamics, rk_order=rk_order, mu=mu, nu=nu, cfl=cfl, load=load, fixed_dt=fixed_dt,
                store=store)
    return _StressWindow.apply(beta0, Lambda, int(nsteps), int(window), comp, float(t0), int(checkpoint_every), opts)


def elastic_energy(beta, t=0.0, mu=1.0, nu=0.3, load=None):
    """Total elastic energy sum_cells 1/2 S sigma sigma (PlasticityState.py:
    317-330, Observer.py:86) with the hand-written gradient."""
    opts = dict(dynamics=L.GLIDE_CLIMB, rk_order=2, Lambda=0.0, mu=mu, nu=nu, cfl=0.25, load=load)
    return _Energy.apply(beta, float(t), opts)
