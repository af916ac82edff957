"""The hand-written adjoint checked at the workload's own size.

Two independent derivatives of the same discrete map:
  * central differences of the branch-frozen map (Plasticity.amd.frozen:
    the minmod selections, the max/min speed selections and the time steps
    of the base run held fixed -- the map the adjoint differentiates), which
    converge to its derivative as eps^2 (Richardson-extrapolated to eps^4);
  * the tangent-linear map (forward-mode differentiation of every operation,
    cdd_flux_tangent), exact up to rounding.
The adjoint must be the transpose of the tangent-linear map at every scale
(a dot-product test), and both must match the frozen differences where
those converge.  Workload: 2D 512^2, 400 SSP-RK2 steps, glide + climb
(Lambda = 0.9) under a uniform shear sigma_xy = 0.005, gradient of the final
total elastic energy with respect to beta0 (32 random unit directions) and
Lambda; the 3D 256^3 uniaxial ramp's last-100-step window for the final
mean sigma_zz."""
import sys

import numpy as np
import pytest

from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField

pytestmark = pytest.mark.gpu

SHEAR = None


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from Plasticity.amd import lib
    global SHEAR
    SHEAR = dict(load_kind=lib.LOAD_TENSOR, load_init=(0, 0.005, 0, 0.005, 0, 0, 0, 0, 0))
    return lib


def rel(a, b):
    return float((a - b).norm() / b.norm())


def _dot(a, b):
    return float((a * b).sum())


@pytest.mark.parametrize("d,n,dyn", [(2, 32, 0), (2, 32, 1), (3, 16, 0), (3, 16, 1)])
def test_frozen_and_tangent_flux(L, d, n, dyn):
    """The recorded flux is the flux, the replay reproduces it, and the
    tangent flux is the transpose of cdd_flux_vjp (dot product) and the
    derivative of the frozen flux (central differences converging as eps^2)."""
    import torch
    say("frozen and tangent flux", d, n, dyn)
    load = dict(load_kind=L.LOAD_UNIAXIAL, load_init_strain=0.001, load_uniaxial_rate=0.05) if d == 3 else SHEAR
    beta = torch.from_numpy(SpectralGaussianField((n,) * d, 2.0, 0.01, seed=51)).cuda()
    p = L.Plan(d, n, dynamics=dyn, rk_order=2, Lambda=0.9, **load)
    rhs, _ = p.flux(beta, 0.3)
    sel = torch.zeros(2 * n ** d, dtype=torch.int64, device="cuda")
    rec = p.flux_frozen(beta, 0.3, sel, True)
    rep = p.flux_frozen(beta, 0.3, sel, False)
    assert rel(rec, rhs) < 1e-13 and torch.equal(rec, rep)
    other = torch.from_numpy(SpectralGaussianField((n,) * d, 2.0, 0.01, seed=52)).cuda()
    assert rel(p.flux_frozen(other, 0.3, sel, False), p.flux(other, 0.3)[0]) > 1e-6
    g = torch.Generator(device="cpu").manual_seed(53)
    dvec = torch.randn(beta.shape, generator=g, dtype=torch.float64).cuda()
    dvec /= dvec.norm()
    w = torch.randn(beta.shape, generator=g, dtype=torch.float64).cuda()
    dlam = 0.0 if dyn == 1 else 0.7
    r2, dr = p.flux_tangent(beta, dvec, 0.3, dlam)
    assert rel(r2, rhs) < 1e-13
    bb, lb = p.flux_vjp(beta, 0.3, w)
    lhs, rhs_ = _dot(w, dr), _dot(bb, dvec) + lb * dlam
    say("  <w, J d> %.15e  <J^T w, d> %.15e" % (lhs, rhs_))
    assert abs(lhs - rhs_) <= 1e-12 * abs(lhs)
    _, dr0 = p.flux_tangent(beta, dvec, 0.3, 0.0)
    errs = []
    for eps in (1e-5, 1e-6):
        fd = (p.flux_frozen(beta + eps * dvec, 0.3, sel, False) - p.flux_frozen(beta - eps * dvec, 0.3, sel, False)) / (
            2 * eps)
        errs.append(rel(fd, dr0))
    say("  frozen central differences vs tangent: %.2e (1e-5), %.2e (1e-6)" % tuple(errs))
    assert errs[1] < 1e-6 and errs[1] < max(0.05 * errs[0], 1e-9)
    if dyn == 0:  # Lambda: the same branches replayed by plans at Lambda +- eps
        eps = 1e-6
        pp = L.Plan(d, n, dynamics=dyn, rk_order=2, Lambda=0.9 + eps, **load)
        pm = L.Plan(d, n, dynamics=dyn, rk_order=2, Lambda=0.9 - eps, **load)
        fd = (pp.flux_frozen(beta, 0.3, sel, False) - pm.flux_frozen(beta, 0.3, sel, False)) / (2 * eps)
        _, drl = p.flux_tangent(beta, torch.zeros_like(beta), 0.3, 1.0)
        say("  Lambda: frozen central difference vs tangent %.2e" % rel(fd, drl))
        assert rel(fd, drl) < 1e-8


def _energy_gradient(L, beta0, Lam, nsteps, rk, store="auto"):
    import torch
    from Plasticity.amd.autograd import run_with_grad, elastic_energy
    b = beta0.clone().requires_grad_(True)
    lt = torch.tensor(Lam, dtype=torch.float64, device="cuda", requires_grad=True)
    E = elastic_energy(run_with_grad(b, lt, nsteps, checkpoint_every=20, rk_order=rk, load=SHEAR, store=store),
                       load=SHEAR)
    E.backward()
    return E.item(), b.grad.detach(), lt.grad.item()


def _base(L, n, d, nsteps, rk, Lam, seed):
    import torch
    beta0 = torch.from_numpy(SpectralGaussianField((n,) * d, 4.0, 0.01, seed=seed)).cuda()
    plan = L.Plan(d, n, dynamics=L.GLIDE_CLIMB, rk_order=rk, Lambda=Lam, **SHEAR)
    log = torch.zeros(nsteps, dtype=torch.float64, device="cuda")
    plan.set_state(beta0)
    plan.set_time(0.0)
    plan.dt_record(log)
    plan.step(nsteps)
    plan.dt_record(None)
    return beta0, plan, log.cpu().tolist(), plan.energy(plan.get_state(), 0.0)


def _tangent_energy(plan, beta0, dvec, dts, dlam=0.0):
    """E(final) and its tangent along (dvec, dlam), the tangent carried on
    the stepper's own stage states (the ones the adjoint linearises)."""
    from Plasticity.amd.frozen import tangent_stored
    u, du, t = tangent_stored(plan, beta0, dvec, len(dts), dLambda=dlam)
    g, e = plan.energy_grad(u, 0.0)
    return e, _dot(g, du)


def _directions(beta0, ndir, seed):
    import torch
    rng = np.random.default_rng(seed)
    for _ in range(ndir):
        dvec = torch.from_numpy(rng.standard_normal(beta0.shape)).cuda()
        yield dvec / dvec.norm()


def test_adjoint_tangent_and_frozen_differences_small(L):
    """2D 32^2 x 20 RK2 steps: adjoint = tangent (dot products to 1e-11);
    Richardson-extrapolated frozen central differences (eps = 1e-6, 5e-7)
    match both to 1e-7 on white-noise directions."""
    import torch
    from Plasticity.amd.frozen import FrozenRun
    say("adjoint, tangent and frozen differences, 2D 32^2 x 20 RK2 steps")
    n, nsteps, Lam = 32, 20, 0.9
    beta0, plan, dts, e_run = _base(L, n, 2, nsteps, 2, Lam, 61)
    e_ad, grad, lam_grad = _energy_gradient(L, beta0, Lam, nsteps, 2)
    fr = FrozenRun(plan, beta0, dts)
    assert abs(plan.energy(fr.base, 0.0) - e_run) <= 1e-12 * e_run and abs(e_ad - e_run) <= 1e-12 * e_run

    def E(b, p=None):
        return plan.energy(fr.replay(b, p), 0.0)
    worst = 0.0
    for k, dvec in enumerate(_directions(beta0, 6, 1061)):
        ad = _dot(grad, dvec)
        e_t, tl = _tangent_energy(plan, beta0, dvec, dts)
        assert e_t == e_run
        D = {eps: (E(beta0 + eps * dvec) - E(beta0 - eps * dvec)) / (2 * eps) for eps in (1e-6, 5e-7)}
        rich = (4 * D[5e-7] - D[1e-6]) / 3
        say("  dir %d: adjoint %.12e tangent %.12e frozen FD %.12e (Richardson %.12e)" % (k, ad, tl, D[1e-6], rich))
        assert abs(ad - tl) <= 1e-11 * abs(ad)
        worst = max(worst, abs(rich - ad) / abs(ad))
    e_t, tl = _tangent_energy(plan, beta0, torch.zeros_like(beta0), dts, dlam=1.0)
    say("  Lambda: adjoint %.12e tangent %.12e; worst Richardson FD error %.2e" % (lam_grad, tl, worst))
    assert abs(lam_grad - tl) <= 1e-11 * abs(lam_grad)
    assert worst <= 1e-7


@pytest.mark.timeout(900)
def test_adjoint_is_the_transposed_tangent_512_400_steps(L):
    """The workload's gradient (2D 512^2, 400 SSP-RK2 steps, shear, final
    elastic energy): <grad E, d> from the checkpointed adjoint equals the
    tangent-linear derivative along 32 random unit directions d, and the
    Lambda derivative likewise."""
    import torch
    say("adjoint vs tangent-linear, 2D 512^2 x 400 RK2 steps, 32 directions")
    n, nsteps, Lam = 512, 400, 0.9
    beta0, plan, dts, e_run = _base(L, n, 2, nsteps, 2, Lam, 4)
    e_ad, grad, lam_grad = _energy_gradient(L, beta0, Lam, nsteps, 2)
    assert abs(e_ad - e_run) <= 1e-12 * e_run
    worst = 0.0
    for k, dvec in enumerate(_directions(beta0, 32, 32)):
        ad = _dot(grad, dvec)
        e_t, tl = _tangent_energy(plan, beta0, dvec, dts)
        err = abs(ad - tl) / abs(ad)
        worst = max(worst, err)
        say("  dir %2d: adjoint %.14e tangent %.14e rel %.2e" % (k, ad, tl, err))
        assert e_t == e_run
    e_t, tl = _tangent_energy(plan, beta0, torch.zeros_like(beta0), dts, dlam=1.0)
    errl = abs(lam_grad - tl) / abs(lam_grad)
    say("  Lambda: adjoint %.14e tangent %.14e rel %.2e; worst over directions %.2e" % (lam_grad, tl, errl, worst))
    assert worst <= 1e-11 and errl <= 1e-11


@pytest.mark.timeout(900)
def test_stress_window_adjoint_is_the_transposed_tangent_3d_256(L):
    """The 3D 256^3 uniaxial ramp (glide + climb, Lambda = 0.9, CFL steps,
    strain rate 0.01 / (1000 dt0), 1000 RK3 steps): the adjoint
    of the final mean sigma_zz (load included) with respect to the beta^P
    that starts the last 100 steps, and to Lambda, against the tangent-linear
    derivative carried on the stepper's own stage states, along 3 random unit
    directions."""
    import torch
    from Plasticity.amd.autograd import mean_stress_window, _mean_stress_grad
    from Plasticity.amd.frozen import tangent_stored
    n, total, window, ck = 256, 1000, 100, 10
    say("3D 256^3 ramp: adjoint of the final mean sigma_zz over the last %d steps vs tangent-linear" % window)
    beta = SpectralGaussianField((n,) * 3, 6.0, 0.01, seed=0)
    init_strain = float(beta[8].mean())
    beta0 = torch.from_numpy(beta).cuda()
    del beta
    probe = L.Plan(3, n, dynamics=L.GLIDE_CLIMB, rk_order=3, Lambda=0.9, mu=1.0, nu=0.3, cfl=0.25)
    _, speed = probe.flux(beta0)
    dt0 = 0.25 / (n * float(speed.max()))
    probe.close()
    del speed
    load = dict(load_kind=L.LOAD_UNIAXIAL, load_init_strain=init_strain, load_uniaxial_rate=0.01 / (total * dt0))
    plan = L.Plan(3, n, dynamics=L.GLIDE_CLIMB, rk_order=3, Lambda=0.9, mu=1.0, nu=0.3, cfl=0.25, **load)
    plan.set_state(beta0)
    plan.set_time(0.0)
    plan.step(total - window)
    t_mid = plan.clock()[0]
    beta_mid = plan.get_state()
    del beta0
    b = beta_mid.clone().requires_grad_(True)
    lt = torch.tensor(0.9, dtype=torch.float64, device="cuda", requires_grad=True)
    J = mean_stress_window(b, lt, window, 1, "zz", t0=t_mid, checkpoint_every=ck, rk_order=3, load=load)
    J.backward()
    grad, lam_grad = b.grad.detach(), lt.grad.item()
    del b
    say("  J = %.15g, |grad| %.6e, dJ/dLambda %.6e" % (J.item(), float(grad.norm()), lam_grad))
    gvec = torch.tensor(_mean_stress_grad("zz", load, 1.0, 0.3, n ** 3), dtype=torch.float64, device="cuda")

    def dJ(dvec, dlam):
        u, du, t = tangent_stored(plan, beta_mid, dvec, window, t0=t_mid, dLambda=dlam, segment=ck)
        return float((gvec * du.reshape(9, -1).sum(dim=1)).sum())
    rng = np.random.default_rng(256)
    worst = 0.0
    for k in range(3):
        dvec = torch.from_numpy(rng.standard_normal((9, n, n, n))).cuda()
        dvec /= dvec.norm()
        ad, tl = float((grad * dvec).sum()), dJ(dvec, 0.0)
        worst = max(worst, abs(ad - tl) / abs(ad))
        say("  dir %d: adjoint %.14e tangent %.14e rel %.2e" % (k, ad, tl, abs(ad - tl) / abs(ad)))
        del dvec
    tl = dJ(torch.zeros_like(beta_mid), 1.0)
    errl = abs(tl - lam_grad) / abs(lam_grad)
    say("  Lambda: adjoint %.14e tangent %.14e rel %.2e" % (lam_grad, tl, errl))
    assert worst <= 1e-11 and errl <= 1e-11
