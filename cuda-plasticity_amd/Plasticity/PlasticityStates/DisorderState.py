"""DisorderState on the GPU (Plasticity/PlasticityStates/DisorderState.py:11-34):
a VacancyState whose ('s', 's') component is a disorder field d.  Its stress
gains the disorder term sum_{l,t} eps_lit irfftn(k_l rfftn(rho'_tj)),
rho' = d rho / (ME + 2 |rho|^2) with rho the spectral Nye tensor, computed in
the HIP library.  With the BetaP dynamics d is carried by the flux and the
RK stages like the tensor components (RunPython.py:196-214 'x' runs)."""
from Plasticity.Fields.Fields import TensorField
from Plasticity.PlasticityStates.VacancyState import VacancyState, S


class DisorderState(VacancyState):
    def __init__(self, gridShape, field=None, nu=0.3, mu=0.5, inherit=None, device=None):
        VacancyState.__init__(self, gridShape, field=field, nu=nu, mu=mu, inherit=inherit, device=device)

    def DisorderField(self):
        """The disorder field d (the ('s', 's') component)."""
        return self.betaP_V[S]

    def CalculateSigma(self, type='k', source='betaE', dodisorder=True):
        """The state's stress plus the disorder term (:15-30).  Like the
        reference, the term is added whatever dodisorder says."""
        from Plasticity.amd import plans
        p = plans.analysis_plan(self.gridShape, self.mu, self.nu, self.betaP.data.device)
        if type == 'k' and source == 'betaE':
            return TensorField(self.gridShape, data=p.disorder_sigma(self.betaP.data, self.DisorderField()))
        base = VacancyState.CalculateSigma(self, type=type, source=source)
        extra = p.disorder_sigma(self.betaP.data, self.DisorderField()) - p.sigma(self.betaP.data)
        return TensorField(self.gridShape, data=base.data + extra)
