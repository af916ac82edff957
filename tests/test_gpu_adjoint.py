"""Hand-written adjoint against (1) float64 torch autograd through the same
discretisation (tests/torch_reference.py, which equals the oracle) and
(2) central finite differences with the frozen time-step schedule."""
import sys

import numpy as np
import pytest
import torch

import torch_reference as TR
from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField

pytestmark = pytest.mark.gpu
MU, NU = 1.0, 0.3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


@pytest.fixture(scope="module")
def L():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from Plasticity.amd import lib
    return lib


SHEAR = dict(load_kind=4, load_init=(0, 0.005, 0, 0.005, 0, 0, 0, 0, 0))


def shear_shift():
    s = torch.zeros(9, dtype=torch.float64)
    s[1] = s[3] = 0.005
    return s


CASES = [((16, 16), 1, 0.0), ((16, 16), 0, 0.9), ((8, 8, 8), 0, 0.9), ((8, 8, 8), 1, 0.0),
         ((32, 32, 32), 0, 0.9), ((32, 32, 32), 1, 0.0)]  # 32^3: the plane-marching gather


@pytest.mark.parametrize("shape,dyn,Lam", CASES)
def test_flux_vjp_matches_autograd(L, shape, dyn, Lam):
    say("flux vjp", shape, dyn, Lam)
    beta = SpectralGaussianField(shape, 2.0, 0.01, seed=21)
    rng = np.random.default_rng(5)
    rbar = rng.standard_normal((9,) + shape)
    p = L.Plan(len(shape), shape[0], dynamics=dyn, rk_order=2, Lambda=Lam, mu=MU, nu=NU, **SHEAR)
    g, lg = p.flux_vjp(beta, 0.0, rbar)
    bt = torch.from_numpy(beta).requires_grad_(True)
    lt = torch.tensor(Lam, dtype=torch.float64, requires_grad=True)
    out = TR.flux(bt, TR.sigma(bt, MU, NU, shear_shift()), lt if dyn == 0 else None, dyn == 1)
    (out * torch.from_numpy(rbar)).sum().backward()
    assert rel(g.cpu().numpy(), bt.grad.numpy()) < 1e-10
    if dyn == 0:
        assert abs(lg - lt.grad.item()) <= 1e-9 * abs(lt.grad.item())


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("shape,dyn,Lam", CASES[:3])
def test_step_vjp_matches_autograd(L, shape, dyn, Lam, order):
    say("step vjp", shape, dyn, Lam, order)
    beta = SpectralGaussianField(shape, 2.0, 0.01, seed=22)
    ubar = np.random.default_rng(6).standard_normal((9,) + shape)
    p = L.Plan(len(shape), shape[0], dynamics=dyn, rk_order=order, Lambda=Lam, mu=MU, nu=NU, **SHEAR)
    _, speed = p.flux(beta)
    dt = 0.25 / (shape[0] * float(speed.max()))
    g, lg = p.step_vjp(beta, 0.0, dt, ubar)
    bt = torch.from_numpy(beta).requires_grad_(True)
    lt = torch.tensor(Lam, dtype=torch.float64, requires_grad=True)
    out = TR.rk_step(bt, dt, lt if dyn == 0 else None, dyn == 1, MU, NU, order, shear_shift())
    (out * torch.from_numpy(ubar)).sum().backward()
    assert rel(g.cpu().numpy(), bt.grad.numpy()) < 1e-10
    if dyn == 0:
        assert abs(lg - lt.grad.item()) <= 1e-9 * abs(lt.grad.item())


@pytest.mark.parametrize("shape", [(16, 16), (8, 8, 8)])
def test_energy_grad_matches_autograd(L, shape):
    say("energy grad", shape)
    beta = SpectralGaussianField(shape, 2.0, 0.01, seed=23)
    p = L.Plan(len(shape), shape[0], mu=MU, nu=NU, **SHEAR)
    g, e = p.energy_grad(beta, 0.0)
    bt = torch.from_numpy(beta).requires_grad_(True)
    et = TR.energy(bt, MU, NU, shear_shift())
    et.backward()
    assert abs(e - et.item()) <= 1e-12 * abs(et.item())
    assert rel(g.cpu().numpy(), bt.grad.numpy()) < 1e-11


def _run_energy(L, plan, beta0, nsteps, sched):
    plan.set_state(beta0)
    plan.set_time(0.0)
    plan.dt_replay(sched)
    plan.step(nsteps)
    plan.dt_replay(None)
    return plan.energy(plan.get_state(), plan.clock()[0])


def _fd_check(f, fm, f0, fp, eps, ad, rtol):
    """Central difference against the adjoint value `ad`.  The scheme is
    piecewise smooth (minmod and max/min speed selections): when a kink of
    one of them lies within +-eps of the base point the central difference
    is off by exactly half the gap between the one-sided differences (one
    kink, piecewise-linear model), so that half-gap is the difference's own
    uncertainty and is added to the tolerance."""
    fwd, bwd, cen = (fp - f0) / eps, (f0 - fm) / eps, (fp - fm) / (2 * eps)
    ok = abs(cen - ad) <= rtol * abs(ad) + 0.5 * abs(fwd - bwd)
    return ok, (f, cen, fwd, bwd, ad)


def test_checkpointed_gradient_vs_finite_differences(L):
    """Gradient of the final elastic energy after 8 RK2 steps (checkpoints
    every 3) w.r.t. beta0 and Lambda, against central differences replaying
    the recorded time steps, on unit-norm random directions (eps = 1e-8).
    A small grid and few steps keep the number of minmod / max-speed kinks
    within +-eps of the base point low."""
    say("checkpointed gradient vs FD")
    from Plasticity.amd.autograd import run_with_grad, elastic_energy
    n, nsteps, Lam = 16, 8, 0.9
    beta0 = torch.from_numpy(SpectralGaussianField((n, n), 2.5, 0.01, seed=24)).cuda()
    b = beta0.clone().requires_grad_(True)
    lt = torch.tensor(Lam, dtype=torch.float64, device="cuda", requires_grad=True)
    final = run_with_grad(b, lt, nsteps, checkpoint_every=3, rk_order=2, load=SHEAR)
    # the energy is evaluated at the final time (the shear load does not depend on time)
    E = elastic_energy(final, load=SHEAR)
    E.backward()
    grad = b.grad.detach()
    # recorded schedule for the FD replays
    plan = L.Plan(2, n, dynamics=0, rk_order=2, Lambda=Lam, mu=MU, nu=NU, **SHEAR)
    log = torch.zeros(nsteps, dtype=torch.float64, device="cuda")
    plan.set_state(beta0)
    plan.set_time(0.0)
    plan.dt_record(log)
    plan.step(nsteps)
    plan.dt_record(None)
    e0 = plan.energy(plan.get_state(), 0.0)
    assert abs(e0 - E.item()) <= 1e-13 * abs(e0)
    assert _run_energy(L, plan, beta0, nsteps, log) == e0
    rng = np.random.default_rng(7)
    eps = 1e-8
    for k in range(4):
        d = torch.from_numpy(rng.standard_normal((9, n, n))).cuda()
        d = d / torch.linalg.norm(d)
        ep = _run_energy(L, plan, beta0 + eps * d, nsteps, log)
        em = _run_energy(L, plan, beta0 - eps * d, nsteps, log)
        ok, info = _fd_check(k, em, e0, ep, eps, float((grad * d).sum()), 1e-6)
        assert ok, info
    # climb-suppression parameter
    h = 1e-8
    pp = L.Plan(2, n, dynamics=0, rk_order=2, Lambda=Lam + h, mu=MU, nu=NU, **SHEAR)
    pm = L.Plan(2, n, dynamics=0, rk_order=2, Lambda=Lam - h, mu=MU, nu=NU, **SHEAR)
    ok, info = _fd_check("Lambda", _run_energy(L, pm, beta0, nsteps, log), e0,
                         _run_energy(L, pp, beta0, nsteps, log), h, lt.grad.item(), 1e-6)
    assert ok, info


def test_end_to_end_matches_torch_autograd(L):
    """The whole checkpointed backward pass against torch autograd through
    the same discretisation, with the same (frozen) time steps."""
    say("end to end vs torch")
    from Plasticity.amd.autograd import run_with_grad, elastic_energy
    n, nsteps, Lam = 16, 6, 0.9
    beta0 = SpectralGaussianField((n, n), 2.0, 0.01, seed=25)
    b = torch.from_numpy(beta0).cuda().requires_grad_(True)
    final = run_with_grad(b, Lam, nsteps, checkpoint_every=4, rk_order=3, load=SHEAR)
    elastic_energy(final, load=SHEAR).backward()
    # time steps actually used
    plan = L.Plan(2, n, dynamics=0, rk_order=3, Lambda=Lam, mu=MU, nu=NU, **SHEAR)
    log = torch.zeros(nsteps, dtype=torch.float64, device="cuda")
    plan.set_state(beta0)
    plan.set_time(0.0)
    plan.dt_record(log)
    plan.step(nsteps)
    dts = log.cpu().tolist()
    bt = torch.from_numpy(beta0).requires_grad_(True)
    u = bt
    for dt in dts:
        u = TR.rk_step(u, dt, Lam, False, MU, NU, 3, shear_shift())
    TR.energy(u, MU, NU, shear_shift()).backward()
    assert rel(b.grad.cpu().numpy(), bt.grad.numpy()) < 1e-9


SEG_CASES = [((16, 16), 0, 0.9, 2), ((32, 32), 1, 0.0, 3), ((32, 32), 0, 0.9, 3), ((8, 8, 8), 0, 0.9, 2),
             ((8, 8, 8), 1, 0.0, 3)]


@pytest.mark.parametrize("shape,dyn,Lam,order", SEG_CASES)
def test_segment_vjp_matches_chained_step_vjps(L, shape, dyn, Lam, order):
    """cdd_segment_vjp (recompute with stored stages, fused RK combination,
    device Lambda accumulator) equals the chain of per-step VJPs, and leaves
    the plan's clock and state untouched."""
    say("segment vjp", shape, dyn, Lam, order)
    d, n, nsteps = len(shape), shape[0], 3
    plan = L.Plan(d, n, dynamics=dyn, rk_order=order, Lambda=Lam, mu=MU, nu=NU, **SHEAR)
    u0 = torch.from_numpy(SpectralGaussianField(shape, 2.0, 0.01, seed=40 + d)).cuda()
    log = torch.zeros(nsteps, dtype=torch.float64, device="cuda")
    plan.set_state(u0)
    plan.set_time(0.25)
    plan.dt_record(log)
    states = [u0]
    for _ in range(nsteps):
        plan.step(1)
        states.append(plan.get_state())
    plan.dt_record(None)
    clock_before = plan.clock()
    state_before = plan.get_state()
    dts = log.cpu().tolist()
    times = [0.25]
    for dt in dts[:-1]:
        times.append(times[-1] + dt)
    gbar = torch.from_numpy(np.random.default_rng(3).standard_normal((9,) + shape)).cuda()
    g, lam = gbar, 0.0
    for j in reversed(range(nsteps)):
        g, lb = plan.step_vjp(states[j], times[j], dts[j], g)
        lam += lb
    plan.lambda_bar(reset=True)
    seg = plan.segment_vjp(u0, 0.25, log, gbar)
    lam_seg = plan.lambda_bar(reset=True)
    assert rel(seg.cpu().numpy(), g.cpu().numpy()) < 1e-12
    if dyn == 0:
        assert abs(lam_seg - lam) <= 1e-11 * abs(lam)
    assert plan.clock() == clock_before
    assert torch.equal(plan.get_state(), state_before)
    # in place (out aliases ubar_end) gives the same
    g2 = gbar.clone()
    plan.segment_vjp(u0, 0.25, log, g2, out=g2)
    assert torch.equal(g2, seg)


def test_end_to_end_3d_rk2_matches_torch_autograd(L):
    """Checkpointed backward of a 3D glide+climb RK2 run against torch."""
    say("end to end 3D vs torch")
    from Plasticity.amd.autograd import run_with_grad, elastic_energy
    n, nsteps, Lam = 8, 5, 0.9
    beta0 = SpectralGaussianField((n, n, n), 2.0, 0.01, seed=26)
    b = torch.from_numpy(beta0).cuda().requires_grad_(True)
    lt = torch.tensor(Lam, dtype=torch.float64, device="cuda", requires_grad=True)
    final = run_with_grad(b, lt, nsteps, checkpoint_every=2, rk_order=2, load=SHEAR)
    elastic_energy(final, load=SHEAR).backward()
    plan = L.Plan(3, n, dynamics=0, rk_order=2, Lambda=Lam, mu=MU, nu=NU, **SHEAR)
    log = torch.zeros(nsteps, dtype=torch.float64, device="cuda")
    plan.set_state(beta0)
    plan.set_time(0.0)
    plan.dt_record(log)
    plan.step(nsteps)
    dts = log.cpu().tolist()
    bt = torch.from_numpy(beta0).requires_grad_(True)
    ltt = torch.tensor(Lam, dtype=torch.float64, requires_grad=True)
    u = bt
    for dt in dts:
        u = TR.rk_step(u, dt, ltt, False, MU, NU, 2, shear_shift())
    TR.energy(u, MU, NU, shear_shift()).backward()
    assert rel(b.grad.cpu().numpy(), bt.grad.numpy()) < 1e-9
    assert abs(lt.grad.item() - ltt.grad.item()) <= 1e-9 * abs(ltt.grad.item())


UNIAX = dict(load_kind=3, load_init_strain=0.001, load_uniaxial_rate=0.05)


@pytest.mark.parametrize("order", [2, 3])
def test_uniaxial_step_vjp_and_energy_grad_match_autograd(L, order):
    """The uniaxial strain-rate control's stress depends on <beta_zz>; its
    transpose (a uniform term in betabar_zz) in the step VJP and the energy
    gradient, against torch autograd."""
    say("uniaxial step vjp", order)
    n, t, dt, Lam = 16, 0.3, 2e-3, 0.9
    beta = SpectralGaussianField((n, n), 2.0, 0.01, seed=50)
    beta[8] += 0.002
    p = L.Plan(2, n, dynamics=0, rk_order=order, Lambda=Lam, mu=MU, nu=NU, **UNIAX)
    gbar = np.random.default_rng(5).standard_normal((9, n, n))
    got, lam = p.step_vjp(torch.from_numpy(beta).cuda(), t, dt, torch.from_numpy(gbar).cuda())
    bt = torch.from_numpy(beta).requires_grad_(True)
    lt = torch.tensor(Lam, dtype=torch.float64, requires_grad=True)
    shift = TR.uniaxial_shift(t, 0.05, 0.001, MU, NU)
    out = TR.rk_step(bt, dt, lt, False, MU, NU, order, shift)
    (out * torch.from_numpy(gbar)).sum().backward()
    assert rel(got.cpu().numpy(), bt.grad.numpy()) < 1e-10
    assert abs(lam - lt.grad.item()) <= 1e-10 * abs(lt.grad.item())
    g, e = p.energy_grad(torch.from_numpy(beta).cuda(), t)
    bt2 = torch.from_numpy(beta).requires_grad_(True)
    E = TR.energy(bt2, MU, NU, shift)
    E.backward()
    assert abs(e - E.item()) <= 1e-12 * abs(E.item())
    assert rel(g.cpu().numpy(), bt2.grad.numpy()) < 1e-11


def test_mean_stress_window_matches_torch_autograd(L):
    """Stress-strain objective under the uniaxial control: <sigma_zz> averaged
    over the last 3 of 6 RK3 steps, gradient w.r.t. beta0 and Lambda through
    the checkpointed adjoint (per-step sources), against torch autograd."""
    say("mean stress window vs torch")
    from Plasticity.amd.autograd import mean_stress_window
    n, nsteps, W, Lam = 16, 6, 3, 0.9
    beta0 = SpectralGaussianField((n, n), 2.0, 0.01, seed=51)
    b = torch.from_numpy(beta0).cuda().requires_grad_(True)
    lt = torch.tensor(Lam, dtype=torch.float64, device="cuda", requires_grad=True)
    J = mean_stress_window(b, lt, nsteps, W, "zz", t0=0.1, checkpoint_every=4, rk_order=3, load=UNIAX)
    J.backward()
    plan = L.Plan(2, n, dynamics=0, rk_order=3, Lambda=Lam, mu=MU, nu=NU, **UNIAX)
    log = torch.zeros(nsteps, dtype=torch.float64, device="cuda")
    plan.set_state(beta0)
    plan.set_time(0.1)
    plan.dt_record(log)
    plan.step(nsteps)
    dts = log.cpu().tolist()
    bt = torch.from_numpy(beta0).requires_grad_(True)
    ltt = torch.tensor(Lam, dtype=torch.float64, requires_grad=True)
    u, t, acc = bt, 0.1, 0.0
    for j, dt in enumerate(dts):
        u = TR.rk_step(u, dt, ltt, False, MU, NU, 3, TR.uniaxial_shift(t, 0.05, 0.001, MU, NU))
        t += dt
        if j + 1 > nsteps - W:
            acc = acc + TR.mean_stress(u, MU, NU, TR.uniaxial_shift(t, 0.05, 0.001, MU, NU))[2]
    Jt = acc / W
    Jt.backward()
    assert abs(J.item() - Jt.item()) <= 1e-12 * abs(Jt.item())
    assert rel(b.grad.cpu().numpy(), bt.grad.numpy()) < 1e-9
    assert abs(lt.grad.item() - ltt.grad.item()) <= 1e-9 * abs(ltt.grad.item())


@pytest.mark.parametrize("order,k", [(2, 3), (2, 4), (3, 3)])
def test_segment_forward_equals_step(L, order, k):
    """segment_forward advances the plan exactly like step (state, clock and
    the RK2 ping-pong buffer, odd and even segment lengths), and its store
    starts with the segment's start state."""
    say("segment_forward vs step", order, k)
    n = 16
    beta0 = SpectralGaussianField((n, n), 2.0, 0.01, seed=61)
    outs = []
    for stored in (False, True):
        p = L.Plan(2, n, dynamics=0, rk_order=order, Lambda=0.9, mu=MU, nu=NU, **SHEAR)
        p.set_state(beta0)
        p.set_time(0.0)
        p.step(1)
        start = p.get_state()
        if stored:
            st = torch.empty(p.segment_store_bytes(k), dtype=torch.uint8, device="cuda")
            p.segment_forward(k, st)
            assert torch.equal(st[: start.numel() * 8].view(torch.float64).reshape(start.shape), start)
        else:
            p.step(k)
        p.step(2)  # the plan keeps stepping from the right buffer
        outs.append((p.get_state().cpu().numpy(), p.clock()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert outs[0][1] == outs[1][1]


def test_stored_and_recomputed_segments_give_identical_gradients(L):
    """Stored trajectories (no recomputation) and checkpoint recomputation
    give bit-identical gradients, also with a budget that stores only some of
    the segments; likewise for the time-distributed stress objective."""
    say("stored vs recomputed segments")
    from Plasticity.amd.autograd import run_with_grad, mean_stress_window
    n, nsteps = 16, 7
    beta0 = SpectralGaussianField((n, n), 2.0, 0.01, seed=62)
    one = L.Plan(2, n, dynamics=0, rk_order=2, Lambda=0.9, mu=MU, nu=NU).segment_store_bytes(3)
    res = []
    for store in (False, True, one + 1):
        b = torch.from_numpy(beta0).cuda().requires_grad_(True)
        lt = torch.tensor(0.9, dtype=torch.float64, device="cuda", requires_grad=True)
        final = run_with_grad(b, lt, nsteps, checkpoint_every=3, rk_order=2, load=SHEAR, store=store)
        (final * final).sum().backward()
        res.append((final.detach().cpu().numpy(), b.grad.cpu().numpy(), lt.grad.item()))
    for r in res[1:]:
        assert np.array_equal(r[0], res[0][0])
        assert np.array_equal(r[1], res[0][1])
        assert r[2] == res[0][2]
    res = []
    for store in (False, True):
        b = torch.from_numpy(beta0).cuda().requires_grad_(True)
        lt = torch.tensor(0.9, dtype=torch.float64, device="cuda", requires_grad=True)
        J = mean_stress_window(b, lt, 6, 3, "zz", t0=0.1, checkpoint_every=4, rk_order=3, load=UNIAX, store=store)
        J.backward()
        res.append((J.item(), b.grad.cpu().numpy(), lt.grad.item()))
    assert res[0][0] == res[1][0]
    assert np.array_equal(res[0][1], res[1][1])
    assert res[0][2] == res[1][2]
