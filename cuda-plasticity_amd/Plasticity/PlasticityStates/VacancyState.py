"""VacancyState on the GPU (Plasticity/PlasticityStates/VacancyState.py:9-200):
beta^P plus the vacancy concentration c, one (10, *gridShape) device tensor
in the reference's sorted component order (('s', 's') first).  ``betaP`` is a
view of the nine tensor components, so the PlasticityState physics
(CalculateSigma, CalculateRhoFourier, energies) applies unchanged."""
from Plasticity.Fields.Fields import Field, TensorField, TensorScalarMixedField
from Plasticity.PlasticityStates.PlasticityState import PlasticityState

S = ('s', 's')


class VacancyState(PlasticityState):
    def __init__(self, gridShape, field=None, nu=0.3, mu=0.5, inherit=None, alpha=1., device=None):
        self.alpha = inherit.alpha if isinstance(inherit, VacancyState) else alpha
        PlasticityState.__init__(self, gridShape, field=field, nu=nu, mu=mu, inherit=inherit, device=device)

    def GetOrderParameterField(self):
        return self.betaP_V

    def UpdateOrderParameterField(self, field, device=None):
        if field is None:
            self.betaP_V = TensorScalarMixedField(self.gridShape, device=device)
        elif isinstance(field, Field):
            if len(field.components) != 10:
                raise ValueError("a VacancyState holds nine tensor components and ('s', 's')")
            self.betaP_V = field
        else:  # NumPy array (10, *gridShape), sorted order
            import numpy as np
            import torch
            arr = np.ascontiguousarray(field, dtype=np.float64)
            dev = device if device is not None else torch.device("cuda", 0)
            self.betaP_V = TensorScalarMixedField(self.gridShape, data=torch.from_numpy(arr).to(dev))
        i0 = self.betaP_V.components.index(('x', 'x'))
        self.betaP = TensorField(self.gridShape, data=self.betaP_V.data[i0:i0 + 9])

    def ExtractBetaPField(self):
        """A copy of the nine tensor components (:32-36)."""
        return TensorField(self.gridShape, data=self.betaP.data.clone())

    def DecoupleState(self):
        """(PlasticityState of beta^P, the concentration c) (:38-40)."""
        return PlasticityState(self.gridShape, field=self.ExtractBetaPField(), inherit=self), self.betaP_V[S]

    def RecoupleState(self, betaPstate, cfield):
        """Write beta^P and c back (:42-45)."""
        self.betaP.data.copy_(betaPstate.betaP.data)
        self.betaP_V[S] = cfield

    def CalculateSigmawithVacancy(self, type='k', source='betaE', cfield=None):
        """sigma of the state with -alpha c on the diagonal (:186-195)."""
        sig = self.CalculateSigma(type=type, source=source)
        c = self.betaP_V[S] if cfield is None else cfield
        a = self._alpha_like(sig.data)
        for k in (0, 4, 8):
            sig.data[k] -= a * c
        return sig

    def _alpha_like(self, t):
        """alpha as a number, or an alpha field as a tensor beside t."""
        a = self.alpha
        if isinstance(a, (int, float)):
            return a
        import torch
        return torch.as_tensor(a, dtype=torch.float64, device=t.device)

    def CalculateVacancyEnergy(self):
        """Energy density 1/2 alpha c^2 (:198-199)."""
        c = self.betaP_V[S]
        return 0.5 * self._alpha_like(c) * c * c
