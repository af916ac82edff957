"""The file boundary of the reference's native path (Plasticity/RunCUDA.py).

The reference writes the initial state to a .ics file, compiles and runs a
separate executable and reads its .plas output back (RunCUDA.py:120-254).
Here the stepper runs in process (Plasticity.amd); these helpers keep the
file formats for interoperability with runs made by the reference.
"""
import numpy as np

from Plasticity.FieldInitializers.FieldInitializer import LoadStateRaw, _reverse_axes


def convertStateToCUDA(state, filename):
    """Write the nine components, each with all grid axes reversed, as raw
    float64 (RunCUDA.py:9-22)."""
    f = state.GetOrderParameterField()
    data = f.numpy() if hasattr(f, "numpy") else np.asarray(f)
    np.stack([_reverse_axes(np.asarray(data[c], dtype=np.float64)) for c in range(9)]).tofile(filename)


def convertCUDAToState(filename, N, dim, time=None):
    """(t, state) of a .plas record (the last, or the first with t >= time)."""
    return LoadStateRaw(filename, N, dim, time=time)
