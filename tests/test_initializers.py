"""The reference initialisers restated on the host (FieldInitializer.py)."""
import numpy as np
import pytest

from oracle import cdd_oracle as O
from Plasticity.FieldInitializers.FieldInitializer import (GaussianRandomInitializer, ResizeState,
                                                           RotateState_90degree, SpectralGaussianField,
                                                           zeropadding2D)
from Plasticity.PlasticityStates.PlasticityState import PlasticityState


def _f(state):
    return state.GetOrderParameterField().numpy()


def test_gaussian_random_initializer_spectrum_and_resolution_independence():
    s = GaussianRandomInitializer((128, 128), sigma=0.2, seed=7, device="cpu")
    f = _f(s)
    assert f.shape == (9, 128, 128) and np.isfinite(f).all()
    # same seed, same field; another seed, another field
    assert np.array_equal(f, _f(GaussianRandomInitializer((128, 128), sigma=0.2, seed=7, device="cpu")))
    assert not np.allclose(f, _f(GaussianRandomInitializer((128, 128), sigma=0.2, seed=8, device="cpu")))
    # the power spectrum falls as exp(-k^2 sigma^2 / 2) (white noise times the filter squared)
    P = np.mean([np.abs(np.fft.rfft2(f[c])) ** 2 for c in range(9)], axis=0)
    k1, k2 = 2 * np.pi * 3, 2 * np.pi * 9  # radians per unit length (N cells per unit)
    r = np.mean(P[9, :3]) / np.mean(P[3, :3])
    want = np.exp(-(k2 ** 2 - k1 ** 2) * 0.2 ** 2 / 2)
    assert 0.5 * want < r < 2.0 * want
    # a 64^2 request is the 128^2 field resampled (seed consistent across sizes)
    s64 = GaussianRandomInitializer((64, 64), sigma=0.2, seed=7, device="cpu")
    np.testing.assert_allclose(_f(s64), _f(ResizeState(s, 64, 2, device="cpu")), rtol=0, atol=1e-15)


def test_resize_round_trip_is_exact_for_band_limited_fields():
    a = SpectralGaussianField((32, 32), 4.0, 0.01, seed=3)
    s = PlasticityState((32, 32), field=a, device="cpu")
    back = ResizeState(ResizeState(s, 64, 2, device="cpu"), 32, 2, device="cpu")
    np.testing.assert_allclose(_f(back), a, rtol=0, atol=1e-14)


def test_rotate_90_is_the_flux_symmetry_and_has_order_four():
    a = SpectralGaussianField((16, 16), 2.5, 0.01, seed=5)
    s = PlasticityState((16, 16), field=a, device="cpu")
    r = s
    for _ in range(4):
        r = RotateState_90degree(r, device="cpu")
    np.testing.assert_allclose(_f(r), a, rtol=0, atol=1e-15)
    # the central-upwind flux commutes with the rotation (square grids are isotropic)
    rot = _f(RotateState_90degree(s, device="cpu"))
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.5)
    rhs_rot = dyn.flux(0.0, rot)[0]
    rhs = dyn.flux(0.0, a)[0]
    rhs_then_rot = _f(RotateState_90degree(PlasticityState((16, 16), field=rhs, device="cpu"), device="cpu"))
    np.testing.assert_allclose(rhs_rot, rhs_then_rot, rtol=0, atol=1e-12 * np.abs(rhs).max())


def test_zeropadding_interpolates(tmp_path):
    a = SpectralGaussianField((16, 16), 3.0, 0.01, seed=6)
    path = str(tmp_path / "state16.raw")
    a.tofile(path)
    out = zeropadding2D(16, path)
    assert out.endswith("state32.raw")
    b = np.fromfile(out).reshape(9, 32, 32)
    np.testing.assert_allclose(b[:, ::2, ::2], a, rtol=0, atol=1e-12)
