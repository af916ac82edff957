"""PlasticityState on the GPU (Plasticity/PlasticityStates/PlasticityState.py:19-551).

The order parameter field ``betaP`` is a TensorField over one device tensor
(9, *gridShape).  Stress, Nye tensor and elastic energy come from the HIP
stress solve through the C ABI; arithmetic (``state * a``, ``state + f``,
used by RK stage combinations in the reference mover) stays available for
callers but the GPU mover never needs it.
"""
from Plasticity.Fields.Fields import TensorField
from Plasticity.amd import plans


class PlasticityState:
    def __init__(self, gridShape, field=None, nu=0.3, mu=0.5, inherit=None, device=None):
        self.gridShape = tuple(gridShape)
        self.dimension = len(self.gridShape)
        if inherit is not None:
            self.nu, self.mu = inherit.nu, inherit.mu
        else:
            self.nu, self.mu = nu, mu
        self.UpdateOrderParameterField(field, device)

    # -- order parameter (PlasticityState.py:69-83)
    def GetOrderParameterField(self):
        return self.betaP

    def UpdateOrderParameterField(self, field, device=None):
        if field is None:
            self.betaP = TensorField(self.gridShape, device=device)
        elif isinstance(field, TensorField):
            self.betaP = field
        else:  # NumPy array (9, *gridShape)
            self.betaP = TensorField.from_numpy(field, device=device)

    # pickling (PlasticityState.py:41-60): the order parameter field as host
    # arrays under 'field'; a pickle written by the reference (its __dict__
    # with 'field' a Field of per-component arrays) loads the same way
    def __getstate__(self):
        d = {k: v for k, v in self.__dict__.items() if k not in ("betaP", "betaP_V")}
        d["field"] = self.GetOrderParameterField()
        return d

    def __setstate__(self, d):
        d = dict(d)
        field = d.pop("field")
        for k in ("ktools", "sinktools"):
            d.pop(k, None)
        self.__dict__.update(d)
        self.gridShape = tuple(self.gridShape)
        self.dimension = len(self.gridShape)
        self.UpdateOrderParameterField(field)

    def _plan(self):
        return plans.analysis_plan(self.gridShape, self.mu, self.nu, self.betaP.data.device)

    # -- physics (PlasticityState.py:85-330)
    def CalculateSigma(self, type='k', source='betaE', time=0.0):
        """CalculateSigma(type, source) (:85-230): source 'betaE' (the
        stepper's stress; the reference ignores type here), 'betaP' (the
        M-tensor form) or 'rho' (free boundary, from the Nye tensor), with the
        ktools ('k') or sinktools ('sinK') k vectors."""
        if type not in ('k', 'sinK') or source not in ('betaE', 'betaP', 'rho'):
            raise ValueError("type must be 'k' or 'sinK', source 'betaE', 'betaP' or 'rho'")
        p = self._plan()
        if source in ('rho', 'betaP'):
            return TensorField(self.gridShape, data=p.sigma_variant(self.betaP.data, type, source))
        return TensorField(self.gridShape, data=p.sigma(self.betaP.data, time))

    def CalculateRhoFourier(self, type='k'):
        """Nye tensor rho_ij = -eps_ilm d_l beta_mj (:264-304), spectral
        derivatives with the ktools ('k') or sinktools ('sinK')."""
        if type not in ('k', 'sinK'):
            raise ValueError("type must be 'k' or 'sinK'")
        if type == 'sinK':
            return TensorField(self.gridShape, data=self._plan().rho_variant(self.betaP.data, type))
        return TensorField(self.gridShape, data=self._plan().rho(self.betaP.data))

    def CalculateRhoSymmetric(self):
        """Nye tensor from central differences (:232-262)."""
        return TensorField(self.gridShape, data=self._plan().rho_symmetric(self.betaP.data))

    def CalculateElasticEnergy(self, type='k', source='betaE'):
        """Energy density 1/2 S_ijkm sigma_ij sigma_km (:317-330), a device array."""
        s = self.CalculateSigma(type, source).data
        nu, mu = self.nu, self.mu
        ss = (s * s).sum(dim=0)
        sst = (s.view((3, 3) + self.gridShape) * s.view((3, 3) + self.gridShape).transpose(0, 1)).sum(dim=(0, 1))
        tr = s[0] + s[4] + s[8]
        return 0.5 * (2.0 * ss - sst - nu / (1.0 + nu) * tr * tr) / (2.0 * mu)

    def TotalElasticEnergy(self, time=0.0):
        """Sum of the density over cells (Observer.py:86), reduced on the device."""
        return self._plan().energy(self.betaP.data, time)

    # -- arithmetic (PlasticityState.py:517-551)
    def _wrap(self, data):
        return self.__class__(self.gridShape, field=TensorField(self.gridShape, data=data), inherit=self)

    def _o(self, other):
        if isinstance(other, PlasticityState):
            return other.betaP.data
        if isinstance(other, TensorField):
            return other.data
        return other

    def __mul__(self, other):
        return self._wrap(self.betaP.data * self._o(other))

    __rmul__ = __mul__

    def __add__(self, other):
        return self._wrap(self.betaP.data + self._o(other))

    __radd__ = __add__

    def __sub__(self, other):
        return self._wrap(self.betaP.data - self._o(other))

    def __rsub__(self, other):
        return self._wrap(self._o(other) - self.betaP.data)

    def __truediv__(self, other):
        return self._wrap(self.betaP.data / self._o(other))

    def copy(self):
        return self._wrap(self.betaP.data.clone())
