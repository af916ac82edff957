"""Pins of the oracle's restatement of the CUDA executable's numerics
(oracle/cdd_oracle.py: cuda_derivatives, cuda_velocity_numerator,
cuda_llf_flux, cuda_euler_step; plasticity_kernel.cu) against the Python
path's own oracle functions and closed forms.  The executable cannot be
built here (nvcc, cutil, cuFFT); these checks tie every formula to the
independently pinned Python-path restatement."""
import numpy as np
import pytest

from oracle import cdd_oracle as O
from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField

MU, NU = 0.5, 0.3


def relmax(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("shape", [(16, 16), (8, 8, 8)])
def test_cuda_derivatives_are_the_limited_derivatives(shape):
    beta = SpectralGaussianField(shape, 2.0, 0.01, seed=1)
    for a in range(len(shape)):
        cp, cm = O.cuda_derivatives(beta, a, shape[-1])
        pp, pm = O.find_derivatives(beta, a)
        assert relmax(cp, pp) < 1e-12 and relmax(cm, pm) < 1e-12


def _corner(beta, d, k):
    N = beta.shape[-1]
    du = []
    for a in range(d):
        du += list(O.cuda_derivatives(beta, a, N))
    s = [k % 2, (k % 4) // 2, k // 4]
    return [du[2 * a + s[a]] for a in range(d)]


@pytest.mark.parametrize("lam,glide_only", [(0.0, False), (0.5, False), (1.0, False), (0.0, True)])
def test_cuda_velocity_numerator_2d_is_the_python_velocity(lam, glide_only):
    beta = SpectralGaussianField((16, 16), 2.0, 0.01, seed=2)
    sig = O.sigma(beta, MU, NU)
    for k in range(4):
        us = _corner(beta, 2, k)
        vc = O.cuda_velocity_numerator(us, sig, lam, glide_only)
        vp = O.velocity_lvp(us, sig) if glide_only else O.velocity_glide_climb(us, sig, lam)
        rn = O._rhomod(us)
        for l in range(3):
            assert relmax(vc[l], vp[l] * rn) < 1e-10


def test_cuda_velocity_numerator_3d_differs_by_its_vz_term():
    """3D glide + climb: v_x, v_y as the Python path; v_z differs by exactly
    sum_n uz[y,n] (sigma[y,n] - sigma[z,n]) (the executable subtracts
    uz[y,n] sigma[z,n], plasticity_kernel.cu:364-369)."""
    beta = SpectralGaussianField((8, 8, 8), 2.0, 0.01, seed=3)
    sig = O.sigma(beta, MU, NU)
    for k in range(8):
        us = _corner(beta, 3, k)
        vc = O.cuda_velocity_numerator(us, sig, 0.4, False)
        vp = O.velocity_glide_climb(us, sig, 0.4)
        rn = O._rhomod(us)
        assert relmax(vc[0], vp[0] * rn) < 1e-10 and relmax(vc[1], vp[1] * rn) < 1e-10
        dz = sum(us[2][O.cidx(1, n)] * (sig[O.cidx(1, n)] - sig[O.cidx(2, n)]) for n in range(3))
        assert relmax(vc[2], vp[2] * rn + dz) < 1e-10
        assert np.max(np.abs(dz)) > 1e-6  # the difference is real


@pytest.mark.parametrize("shape,lam,glide_only", [((16, 16), 0.5, False), ((16, 16), 0.0, True),
                                                  ((8, 8, 8), 1.0, False), ((8, 8, 8), 0.0, True)])
def test_llf_flux_is_the_central_upwind_combination_with_one_speed(shape, lam, glide_only):
    """With every one-sided speed equal to a = max |v| the central-upwind
    combination (built here from the Python path's Hamiltonian) is the LLF
    flux; the speed output is d (2a + ME)."""
    d = len(shape)
    beta = SpectralGaussianField(shape, 2.0, 0.01, seed=4)
    sig = O.sigma(beta, MU, NU)
    rhs, speed = O.cuda_llf_flux(beta, sig, lam, glide_only)
    N = shape[-1]
    du = []
    for a in range(d):
        du += list(O.cuda_derivatives(beta, a, N))
    V, US = [], []
    for k in range(1 << d):
        us = _corner(beta, d, k)
        v = O.cuda_velocity_numerator(us, sig, lam, glide_only)
        ir = O.cuda_rho_norm(us) + np.sqrt(O.ME)
        V.append([x / ir for x in v])
        US.append(us)
    amax = np.max(np.abs(np.stack([np.stack(v) for v in V])), axis=(0, 1))
    at = 2 * amax + O.ME
    want = np.zeros_like(beta)
    for k in range(1 << d):
        want = want - O.hamiltonian(US[k], V[k], 0.0 if glide_only else lam) * amax ** d / at ** d
    for a in range(d):
        want = want + (du[2 * a] - du[2 * a + 1]) * amax * amax / at
    assert relmax(rhs, want) < 1e-10
    assert relmax(speed, d * at) < 1e-15


def test_llf_flux_properties():
    beta = np.full((9, 16, 16), 0.01)
    rhs, speed = O.cuda_llf_flux(beta, O.sigma(beta, MU, NU), 0.3)
    assert np.max(np.abs(rhs)) == 0.0 and np.allclose(speed, 2 * O.ME, rtol=0, atol=0)
    b = SpectralGaussianField((16, 16), 2.0, 0.01, seed=5)
    r0, _ = O.cuda_llf_flux(b, O.sigma(b, MU, NU), 0.3)
    bs = np.roll(b, (3, -5), axis=(1, 2))
    r1, _ = O.cuda_llf_flux(bs, O.sigma(bs, MU, NU), 0.3)
    assert relmax(np.roll(r0, (3, -5), axis=(1, 2)), r1) < 1e-12


def test_euler_step_and_end_clamp():
    b = SpectralGaussianField((16, 16), 2.0, 0.01, seed=6)
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.0, MU, NU)
    rhs, speed = O.cuda_llf_flux(b, dyn.get_sigma(b, 0.0), 0.0)
    u1, t1, dts = O.run_cuda(dyn, b, 0.0, 1, 0.5)
    dt = 0.5 / speed.max() / 16
    assert dts[0] == dt and np.array_equal(u1, b + dt * rhs)
    u2, t2, dts2 = O.run_cuda(dyn, b, 0.0, 5, 0.5, t_end=0.5 * dt)
    assert len(dts2) == 1 and dts2[0] == 0.5 * dt + O.ME


def test_tvd3_step_is_the_python_rk3_with_the_executables_dt():
    """With a flux whose speed makes both dt rules agree (N a power of two)
    and no end clamp, TVD3rd equals the Python mover's RK3 step."""
    b = SpectralGaussianField((16, 16), 2.0, 0.01, seed=7)
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.3, MU, NU)

    def f(tt, uu):
        return O.cuda_llf_flux(uu, dyn.get_sigma(uu, tt), 0.3)
    u3, dt3 = O.cuda_tvd3_step(f, 0.0, b, np.inf, 0.5, 16)
    ur, dtr = O.tvd_rk3_step(f, 0.0, b, np.inf, 0.5)
    assert dt3 == dtr
    assert relmax(u3, ur) < 1e-14
    # end clamp lands ME past the end time
    u4, dt4 = O.cuda_tvd3_step(f, 0.0, b, 0.25 * dt3, 0.5, 16)
    assert dt4 == 0.25 * dt3 + O.ME
    # the RK3 amplification on a linear flux L(u) = -k u: 1 - z + z^2/2 - z^3/6
    k = 3.0
    lin = lambda tt, uu: (-k * uu, np.full(uu.shape[1:], 1.0))
    u5, dt5 = O.cuda_tvd3_step(lin, 0.0, b, np.inf, 0.5, 16)
    z = k * dt5
    assert relmax(u5, b * (1 - z + z * z / 2 - z ** 3 / 6)) < 1e-14


def test_nucleation_step_from_zero_is_the_capped_source():
    """DYNAMIC_NUCLEATION from the zero state: the flux of zero vanishes (all
    speeds are zero), so one Euler step is exactly dt * rate * source with
    dt = maxNucleationTimestep = 1 / sqrt(max |sigma(source)|)."""
    from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField
    src = SpectralGaussianField((16, 16), 2.0, 0.01, seed=44)
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.0, 0.5, 0.3)
    cap = O.nucleation_cap(src, 0.5, 0.3)
    s = O.sigma(src, 0.5, 0.3)
    assert cap == 1.0 / np.sqrt(np.abs(s).max()) and 1.0 < cap < 1e3
    rate = 1.0 / 7.0
    u, t, dts = O.run_cuda(dyn, np.zeros_like(src), 0.0, 1, 0.5, source=src, rate=rate, dt_cap=cap)
    assert list(dts) == [cap] and t == cap
    np.testing.assert_array_equal(u, cap * (rate * src))


def test_vacancy_variant_closed_forms():
    """The executable's VACANCIES step: with no vacancy cost it advances the
    tensor exactly as run_cuda, the concentration gains dt tr(rhs) before
    diffusing (diffusion keeps the mean: the total changes by dt sum tr(rhs));
    a uniform concentration acts as the uniform stress -vacancycost c."""
    from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField
    beta = SpectralGaussianField((16, 16), 2.0, 0.01, seed=45)
    c = 0.01 * SpectralGaussianField((16, 16), 2.0, 1.0, seed=46)[0]
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.3, 0.5, 0.3)
    u0, t0, d0 = O.run_cuda(dyn, beta, 0.0, 1, 0.5)
    u, c1, t, dts = O.run_cuda_vacancy(dyn, O.Vacancy(0.0, 1.0, 0.2), beta, c, 0.0, 1, 0.5)
    assert np.array_equal(u, u0) and list(dts) == list(d0)
    rhs = (u0 - beta) / d0[0]
    assert abs(c1.sum() - (c.sum() + d0[0] * (rhs[0] + rhs[4] + rhs[8]).sum())) < 1e-12 * abs(c).sum()
    cu = np.full((16, 16), 0.004)
    a = O.run_cuda_vacancy(dyn, O.Vacancy(0.7, 1.0, 0.2), beta, cu, 0.0, 1, 0.5)
    b = O.run_cuda(O.Dynamics(O.GLIDE_CLIMB, 0.3, 0.5, 0.3, O.Load("stress", initial=(-0.7 * 0.004,) * 3)), beta,
                   0.0, 1, 0.5)
    assert np.array_equal(a[0], b[0]) and list(a[3]) == list(b[2])
