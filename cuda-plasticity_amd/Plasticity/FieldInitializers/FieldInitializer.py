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
    own = not hasattr(filename, "read")  # a path, or an open file (e.g. a tar member)
    f = open(filename, 'rb') if own else filename
    try:
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
    finally:
        if own:
            f.close()
    return t, data


# Globals a state pickle may name: the state and field classes of this
# package (also under the reference's module paths, which are the same) and
# numpy's array reconstruction; anything else is refused, so loading a
# pickle cannot run arbitrary code.
_PICKLE_ALLOWED = {
    ("Plasticity.PlasticityStates.PlasticityState", "PlasticityState"),
    ("Plasticity.PlasticityStates.VacancyState", "VacancyState"),
    ("Plasticity.Fields.Fields", "Field"),
    ("Plasticity.Fields.Fields", "TensorField"),
    ("Plasticity.Fields.Fields", "TensorScalarMixedField"),
    ("numpy.core.multiarray", "_reconstruct"),
    ("numpy._core.multiarray", "_reconstruct"),
    ("numpy.core.multiarray", "scalar"),
    ("numpy._core.multiarray", "scalar"),
    ("numpy", "ndarray"),
    ("numpy", "dtype"),
    ("copy_reg", "_reconstructor"),
    ("copyreg", "_reconstructor"),
    ("__builtin__", "object"),
    ("builtins", "object"),
    ("_codecs", "encode"),  # protocol-2 pickles written by Python 3 carry array bytes this way
}


def _state_unpickler(f):
    import importlib
    import pickle

    class _Restricted(pickle.Unpickler):
        def find_class(self, module, name):
            if (module, name) not in _PICKLE_ALLOWED:
                raise pickle.UnpicklingError("refusing %s.%s in a state pickle" % (module, name))
            mod = {"copy_reg": "copyreg", "__builtin__": "builtins"}.get(module, module)
            if mod == "numpy.core.multiarray":
                mod = "numpy._core.multiarray" if hasattr(np, "_core") else mod
            return getattr(importlib.import_module(mod), name)
    # Python 2 pickles (the reference's) carry numpy buffers as latin-1 strings
    return _Restricted(f, encoding="latin1")


def LoadState(filename, time=None):
    """(time, state) from a RecordStateObserver file of consecutive pickled
    (time, state) pairs (FieldInitializer.py:29-59): the last pair, or the
    first whose time is >= `time`.  Only state, field and numpy classes may
    appear in the pickle.  Reference-written files (Python 2 pickles of its
    PlasticityState, whose field holds per-component arrays) use the same
    module paths and layout; no such file is available here, so that path is
    exercised with a synthetic pickle of the same structure only."""
    T, state = None, None
    with open(filename, "rb") as f:
        while True:
            try:
                u = _state_unpickler(f)
                T = u.load()
                state = u.load()
            except EOFError:
                break
            if time is not None and T >= time:
                return T, state
    if state is None or time is not None:
        raise ValueError("no state%s found in %s" % ("" if time is None else " at or after t=%g" % time, filename))
    return T, state


# ---------------------------------------------------------------------------
# The reference's own initialisers (host side, NumPy/SciPy), so runs can start
# from the same fields as reference runs with the same seed.
# ---------------------------------------------------------------------------
def _ktools_kSq(gridShape):
    """kSq of FourierSpaceTools (GridArray/FourierSpaceTools.py:25-53) on the
    half spectrum: 2D scales both axes by gridShape[1], 3D each axis by its
    own size; theta - 2 pi [theta > pi] keeps the Nyquist at +pi."""
    def ident(a):
        return a - 2 * np.pi * (a > np.pi).astype(float)
    shape = tuple(gridShape)
    half = shape[:-1] + (shape[-1] // 2 + 1,)
    idx = np.meshgrid(*[np.arange(s) for s in half], indexing="ij")
    if len(shape) == 2:
        ks = [shape[1] * ident(2.0 * np.pi * idx[a] / float(shape[a])) for a in range(2)]
    else:
        ks = [shape[a] * ident(2.0 * np.pi * idx[a] / float(shape[a])) for a in range(len(shape))]
    return sum(k * k for k in ks)


def ResampleArray(arr, N, Dim=2):
    """scipy.signal.resample along every axis (FieldInitializer.py:421-428)."""
    from scipy import signal
    out = np.asarray(arr)
    for a in range(Dim):
        out = signal.resample(out, N, axis=a)
    return out


def _state_like(gridShape, field, device=None):
    from Plasticity.PlasticityStates.PlasticityState import PlasticityState
    return PlasticityState(tuple(gridShape), field=np.ascontiguousarray(field, dtype=np.float64), device=device)


def ResizeState(state, N, Dim=2, device=None):
    """Resample every component to N^Dim (FieldInitializer.py:430-442)."""
    f = state.GetOrderParameterField().numpy()
    new = np.stack([ResampleArray(f[c], N, Dim) for c in range(9)])
    return _state_like((N,) * Dim, new, device)


def GaussianRandomInitializer(gridShape, sigma=0.2, seed=None, device=None):
    """The reference's random initial state (FieldInitializer.py:153-209):
    the fields are always generated on a 128^d grid -- per component in
    sorted order, white noise numpy.random.normal(scale=128) from the legacy
    global generator seeded with `seed` (0 when None), filtered in k-space
    by (sqrt(pi) sigma)^d exp(-kSq sigma^2 / 4) -- and resampled to the
    requested size, so a seed gives the same field at every resolution.
    Only the plain PlasticityState branch is restated (no slip systems,
    vacancies or smectic states)."""
    d = len(gridShape)
    g = (128,) * d
    Ksq = _ktools_kSq(g)
    rs = np.random.RandomState(0 if seed is None else seed)
    field = np.empty((9,) + g)
    for c in range(9):
        temp = rs.normal(scale=g[0], size=g)
        ktemp = np.fft.rfftn(temp) * (np.sqrt(np.pi) * sigma) ** d * np.exp(-Ksq * sigma ** 2 / 4.)
        field[c] = np.real(np.fft.irfftn(ktemp, s=g, axes=tuple(range(d))))
    if tuple(gridShape)[0] != 128:
        field = np.stack([ResampleArray(field[c], gridShape[0], d) for c in range(9)])
    return _state_like(tuple(gridShape), field, device)


def RotateState_90degree(state, device=None):
    """Rotate by 90 degrees about z (FieldInitializer.py:488-503): the
    component map of R = [[0, 1, 0], [-1, 0, 0], [0, 0, 1]] followed by
    scipy.ndimage.rotate(., 90, mode='wrap') of every component."""
    from scipy import ndimage as ndi
    b = state.GetOrderParameterField().numpy()
    src = {0: (4, 1), 1: (3, -1), 2: (5, 1), 3: (1, -1), 4: (0, 1), 5: (2, -1), 6: (7, 1), 7: (6, -1), 8: (8, 1)}
    out = np.stack([ndi.rotate(sgn * b[s], 90, mode='wrap') for s, sgn in (src[c] for c in range(9))])
    return _state_like(state.gridShape, out, device)


def zeropadding2D(N, filename):
    """Double a raw 2D state file by zero-padding its spectrum
    (FieldInitializer.py:321-334): reads 9 N^2 float64 values, writes the
    2N^2 field to filename with str(N) replaced by str(2N)."""
    a = np.fromfile(filename).reshape(9, N, N)
    b = np.zeros((9, 2 * N, 2 * N), float)
    h = N // 2
    for i in range(9):
        ka = np.fft.fft2(a[i])
        kb = np.zeros((2 * N, 2 * N), complex)
        kb[:h, :h] = ka[:h, :h]
        kb[-h:, :h] = ka[-h:, :h]
        kb[:h, -h:] = ka[:h, -h:]
        kb[-h:, -h:] = ka[-h:, -h:]
        b[i] = np.real(np.fft.ifft2(kb))
    b *= 4.
    out = filename.replace(str(N), str(2 * N))
    b.tofile(out)
    return out
