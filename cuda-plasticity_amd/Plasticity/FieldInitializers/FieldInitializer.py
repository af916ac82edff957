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


def _reverse_axes(a):
    """Every grid axis reversed: the component layout of the reference's .ics
    and .plas files (Python axis 0 fastest)."""
    return np.ascontiguousarray(a.transpose(tuple(reversed(range(a.ndim)))))


def LoadStateRaw(filename, N, dim, time=None, dtype='float64', hastimes=True, device=None):
    """Load (t, state) from a raw file of [t, 9 N^dim] records in the reversed
    component layout (FieldInitializer.py:61-101): the last record, or the
    first whose time is >= `time`; hastimes=False reads a bare .ics file
    (RunCUDA.py:9-22) with t = 0.  Like the reference, the records are
    transposed on load, so this reads the CUDA executable's .plas output
    and .ics files (and not the untransposed RecordStateObserverRaw output,
    for which LoadStateRawUntransposed is provided)."""
    from Plasticity.PlasticityStates.PlasticityState import PlasticityState
    t, data = _read_record(filename, N, dim, time, dtype, hastimes)
    field = np.stack([_reverse_axes(data[c]) for c in range(9)]).astype(np.float64)
    return t, PlasticityState(tuple([N] * dim), field=field, device=device)


def LoadStateRawUntransposed(filename, N, dim, time=None, dtype='float64', device=None):
    """(t, state) from a RecordStateObserverRaw file (Observer.py:51-69), which
    stores the components in their own (NumPy) layout."""
    from Plasticity.PlasticityStates.PlasticityState import PlasticityState
    t, data = _read_record(filename, N, dim, time, dtype, True)
    return t, PlasticityState(tuple([N] * dim), field=data.astype(np.float64), device=device)


def _read_record(filename, N, dim, time, dtype, hastimes):
    dtype = np.dtype(dtype)
    es = dtype.itemsize
    datasize = 9 * N ** dim
    rec = es * (datasize + (1 if hastimes else 0))
    with open(filename, 'rb') as f:
        f.seek(0, 2)
        cnt = f.tell() // rec
        if cnt < 1:
            raise ValueError("%s holds no complete record for N=%d, dim=%d" % (filename, N, dim))
        k = cnt - 1
        if hastimes and time is not None:
            for i in range(cnt):
                f.seek(i * rec)
                if np.frombuffer(f.read(es), dtype=dtype)[0] >= time:
                    k = i
                    break
        f.seek(k * rec)
        t = float(np.frombuffer(f.read(es), dtype=dtype)[0]) if hastimes else 0.0
        data = np.frombuffer(f.read(es * datasize), dtype=dtype).reshape((9,) + (N,) * dim)
    return t, data
