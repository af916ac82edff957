"""Tensor fields backed by one device array.

The reference ``Field`` is a dict of per-component NumPy arrays in sorted
component order (Plasticity/Fields/Fields.py:7-291).  Here a TensorField is
a view of a single (ncomp, *gridShape) float64 torch tensor on the GPU -- the
layout the HIP kernels read -- with the same indexing: ``f['x', 'y']`` or
``f[x, y]`` returns that component's (device) array, and ``.components``
lists the sorted component keys.
"""
import numpy as np

from Plasticity.Constants import x, y, z

TENSOR_COMPONENTS = [(i, j) for i in (x, y, z) for j in (x, y, z)]
VECTOR_COMPONENTS = [x, y, z]


class Field(object):
    def __init__(self, gridShape, components, data=None, device=None):
        import torch
        self.gridShape = tuple(gridShape)
        self.components = sorted(components)
        self._index = {c: k for k, c in enumerate(self.components)}
        if data is None:
            dev = device if device is not None else torch.device("cuda", 0)
            data = torch.zeros((len(self.components),) + self.gridShape, dtype=torch.float64, device=dev)
        self.data = data

    def __getitem__(self, index):
        return self.data[self._index[index]]

    # pickling: host arrays per component, the reference's Field layout
    # (Fields.py:7-50: gridShape, components, data = {component: array})
    def __getstate__(self):
        arr = self.data.detach().cpu().numpy()
        return {"gridShape": self.gridShape, "components": list(self.components),
                "data": {c: arr[k].copy() for c, k in self._index.items()}}

    def __setstate__(self, d):
        import torch
        comps = sorted(d["components"])
        data = d["data"]
        arr = np.stack([np.asarray(data[c], dtype=np.float64) for c in comps])
        self.gridShape = tuple(d["gridShape"])
        self.components = comps
        self._index = {c: k for k, c in enumerate(comps)}
        dev = torch.device("cuda", 0) if torch.cuda.is_available() else torch.device("cpu")
        self.data = torch.from_numpy(arr).to(dev)

    def __setitem__(self, index, value):
        self.data[self._index[index]] = value

    def GridShape(self):
        return self.gridShape

    def GridDimension(self):
        return len(self.gridShape)

    def numpy(self):
        """Host copy as an array (ncomp, *gridShape)."""
        return self.data.detach().cpu().numpy()

    def numpy_arrays(self):
        a = self.numpy()
        return {c: a[k] for c, k in self._index.items()}

    def _new(self, data):
        return self.__class__(self.gridShape, components=self.components, data=data)

    def _other(self, other):
        return other.data if isinstance(other, Field) else other

    def __add__(self, other):
        return self._new(self.data + self._other(other))

    __radd__ = __add__

    def __sub__(self, other):
        return self._new(self.data - self._other(other))

    def __rsub__(self, other):
        return self._new(self._other(other) - self.data)

    def __mul__(self, other):
        return self._new(self.data * self._other(other))

    __rmul__ = __mul__

    def __truediv__(self, other):
        return self._new(self.data / self._other(other))

    def __neg__(self):
        return self._new(-self.data)

    def max(self):
        return float(self.data.max())

    def min(self):
        return float(self.data.min())

    def sum(self):
        return float(self.data.sum())

    def modulus(self):
        return (self.data * self.data).sum(dim=0).sqrt()


class TensorField(Field):
    """Second-rank tensor field, nine components xx xy xz yx yy yz zx zy zz
    (Fields.py:235-251)."""

    def __init__(self, gridShape, components=None, data=None, device=None, kspace=False):
        if kspace:
            raise NotImplementedError("k-space fields stay inside the HIP stress solve")
        super(TensorField, self).__init__(gridShape, components or list(TENSOR_COMPONENTS), data, device)

    @classmethod
    def from_numpy(cls, arr, device=None):
        import torch
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        dev = device if device is not None else torch.device("cuda", 0)
        return cls(arr.shape[1:], data=torch.from_numpy(arr).to(dev))


MIXED_COMPONENTS = list(TENSOR_COMPONENTS) + [('s', 's')]


class TensorScalarMixedField(Field):
    """Nine tensor components plus the scalar ('s', 's') of a VacancyState
    (Fields.py:348-358); in sorted order ('s', 's') is component 0."""

    def __init__(self, gridShape, components=None, data=None, device=None, kspace=False):
        if kspace:
            raise NotImplementedError("k-space fields stay inside the HIP stress solve")
        super(TensorScalarMixedField, self).__init__(gridShape, components or list(MIXED_COMPONENTS), data, device)
