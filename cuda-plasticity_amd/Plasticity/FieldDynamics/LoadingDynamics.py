"""Externally loaded BetaP dynamics, expressed as declarative uniform loads.

  * UpwindLoadingBetaPDynamics / NewGlideOnlyLoadDynamics
    (Plasticity/RunPython.py:35-66): type 'strain' adds ExternalStrain of
    initial + rate (t - initT) (:21-26), type 'stress' ExternalStress (:28-32).
    The reference passes (Dx, coreEnergy, coreEnergyLog) positionally into
    (Dx, Lambda, coreEnergy), dropping Lambda (:36-38); here Lambda is kept.
  * UniaxialLoadingBetaPDynamics (MattTools/Run/RunLoadingMPI.py:62-83):
    strain-rate control along z, betaE_zz = rate t - (<beta_zz> - initStrain).
  * ShearLoadingBetaPDynamics: sigma_xy = sigma_yx += stress0 + rate (t - initT)
    (uniform shear; no reference counterpart).
"""
import numpy as np

from Plasticity.amd import lib as L
from Plasticity.FieldDynamics._gpu_betap import GpuBetaPDynamics, GpuGlideOnlyDynamics


# RunPython's module constants, which its ExternalStrain applies whatever the
# state's elastic constants (Plasticity/RunPython.py:19-26)
RUNPYTHON_MU, RUNPYTHON_NU = 0.5, 0.3


def _load3(kind, initial, rate, initT):
    init = np.zeros(9)
    rt = np.zeros(9)
    init[:3] = np.asarray(initial, float)
    rt[:3] = np.asarray(rate, float)
    kw = dict(load_kind=L.LOAD_STRAIN if kind == 'strain' else L.LOAD_STRESS,
              load_init=tuple(init), load_rate=tuple(rt), load_t0=initT)
    if kind == 'strain':
        kw.update(load_mu=RUNPYTHON_MU, load_nu=RUNPYTHON_NU)
    return kw


class UpwindLoadingBetaPDynamics(GpuBetaPDynamics):
    def __init__(self, Dx=None, Lambda=0, coreEnergy=0, coreEnergyLog=0, rate=None, type='strain',
                 initial=np.array([0., 0., 0.]), initT=0.):
        GpuBetaPDynamics.__init__(self, Dx, Lambda, coreEnergy, coreEnergyLog,
                                  **_load3(type, initial, rate if rate is not None else (0., 0., 0.), initT))
        self.rate, self.type, self.initial, self.initT = rate, type, initial, initT


class NewGlideOnlyLoadDynamics(GpuGlideOnlyDynamics):
    def __init__(self, Dx=None, Lambda=0, coreEnergy=0, coreEnergyLog=0, rate=None, type='strain',
                 initial=np.array([0., 0., 0.]), initT=0.):
        GpuGlideOnlyDynamics.__init__(self, Dx, coreEnergy, coreEnergyLog,
                                      **_load3(type, initial, rate if rate is not None else (0., 0., 0.), initT))
        self.rate, self.type, self.initial, self.initT = rate, type, initial, initT


class UniaxialLoadingBetaPDynamics(GpuBetaPDynamics):
    def __init__(self, Dx=None, Lambda=0, coreEnergy=0, coreEnergyLog=0, rate=None, initStrain=None):
        GpuBetaPDynamics.__init__(self, Dx, Lambda, coreEnergy, coreEnergyLog, load_kind=L.LOAD_UNIAXIAL,
                                  load_init_strain=float(initStrain or 0.0), load_uniaxial_rate=float(rate or 0.0))
        self.rate, self.initStrain = rate, initStrain


class ShearLoadingBetaPDynamics(GpuBetaPDynamics):
    def __init__(self, Lambda=0, stress0=0.0, rate=0.0, initT=0.0, glide_only=False):
        init = np.zeros(9)
        rt = np.zeros(9)
        init[1] = init[3] = stress0
        rt[1] = rt[3] = rate
        GpuBetaPDynamics.__init__(self, None, Lambda, load_kind=L.LOAD_TENSOR, load_init=tuple(init),
                                  load_rate=tuple(rt), load_t0=initT)
        if glide_only:
            self.dynamics_kind = L.GLIDE_ONLY
