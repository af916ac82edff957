"""Seeded initial beta^P fields.

``SpectralGaussianField`` is the synthetic-input recipe of the benchmark
workloads (it is *not* the reference's ``GaussianRandomInitializer``,
Plasticity/FieldInitializers/FieldInitializer.py:153-209, which generates at
128^d and resamples): for each of the nine components in sorted order
(xx, xy, xz, yx, yy, yz, zx, zy, zz) draw white noise z_c ~ N(0, 1) from
``numpy.random.Generator(PCG64(seed))``, filter it in k-space by
exp(-|k|^2 L^2 / 2) with k = 2 pi n_wrapped / N in radians per cell and L
in cells, transform back and rescale the component to the requested RMS
(sqrt(mean(beta_c^2))); the mean is kept.
"""
import numpy as np


def _kgrid_sq(shape):
    ks = []
    for a, n in enumerate(shape):
        if a == len(shape) - 1:
            k = 2.0 * np.pi * np.arange(n // 2 + 1) / n
        else:
            k = 2.0 * np.pi * np.fft.fftfreq(n)
        ks.append(k)
    grids = np.meshgrid(*ks, indexing="ij")
    return sum(g * g for g in grids)


def SpectralGaussianField(gridShape, lengthscale, rms=0.01, seed=0):
    """Return a float64 array of shape (9, *gridShape)."""
    shape = tuple(int(s) for s in gridShape)
    rng = np.random.Generator(np.random.PCG64(seed))
    filt = np.exp(-_kgrid_sq(shape) * lengthscale * lengthscale / 2.0)
    out = np.empty((9,) + shape, dtype=np.float64)
    for c in range(9):
        z = rng.standard_normal(shape)
        b = np.fft.irfftn(np.fft.rfftn(z) * filt, s=shape, axes=tuple(range(len(shape))))
        r = np.sqrt(np.mean(b * b))
        out[c] = b * (rms / r)
    return out
