"""The reference's Python-path driver (Plasticity/RunPython.py) on the GPU.

simulation(dct) keeps RunPython.simulation's run names
(<method><dim>d<N>_s<seed>_<previous><postfix>), its dynamics per method
('mdp' Lambda = 1, 'gcd' Lambda = 0, 'lvp' glide only, a trailing 'x' for a
DisorderState), its mover (TVD-RK3, CFLsafeFactor 0.1), its record schedule
and its untransposed raw record file, with every step in the HIP library.
The loading dynamics classes take their load as data (the GPU stress solve
adds a uniform load as its k = 0 mode), with the reference's arguments.
"""
import os
import shutil as sh

import numpy as np

from Plasticity import PlasticitySystem
from Plasticity.Configure import Configuration, write_json
from Plasticity.FieldDynamics import CentralUpwindHJBetaPDynamics, CentralUpwindHJBetaPGlideOnlyDynamics
from Plasticity.FieldDynamics.VacancyDynamics import BetaP_VacancyDynamics
from Plasticity.FieldInitializers import FieldInitializer
from Plasticity.FieldMovers import FieldMover
from Plasticity.Observers import Observer
from Plasticity.amd import lib as L

mu, nu = 0.5, 0.3
lamb = 2. * mu * nu / (1. - 2. * nu)


def ExternalStrain(sigma, primaryStrain):
    """sigma_ii += lamb tr(e) + 2 mu e_i (RunPython.py:21-26), in place on a
    TensorField."""
    tr = primaryStrain[0] + primaryStrain[1] + primaryStrain[2]
    for i, c in enumerate((0, 4, 8)):
        sigma.data[c] += lamb * tr + 2. * mu * primaryStrain[i]
    return sigma


def ExternalStress(sigma, primaryStress):
    """sigma_ii += s_i (RunPython.py:28-32), in place on a TensorField."""
    for i, c in enumerate((0, 4, 8)):
        sigma.data[c] += primaryStress[i]
    return sigma


def _load_kwargs(rate, type, initial, initT):
    """GetSigma's ExternalStrain / ExternalStress of initial + rate (t - initT)
    (RunPython.py:44-49) as the GPU stepper's load."""
    if rate is None:
        return {}
    kind = L.LOAD_STRAIN if type == 'strain' else L.LOAD_STRESS
    kw = dict(load_kind=kind, load_init=tuple(np.asarray(initial, float).ravel()[:3]),
              load_rate=tuple(np.asarray(rate, float).ravel()[:3]), load_t0=float(initT))
    if kind == L.LOAD_STRAIN:
        # ExternalStrain's Hooke's law takes this module's mu, nu, not the state's
        kw.update(load_mu=mu, load_nu=nu)
    return kw


class UpwindLoadingBetaPDynamics(CentralUpwindHJBetaPDynamics.BetaPDynamics):
    """Glide + climb under an external strain or stress (RunPython.py:35-49).
    The reference passes (Dx, coreEnergy, coreEnergyLog) positionally into
    (Dx, Lambda, coreEnergy), so its Lambda is its coreEnergy argument (0);
    kept here."""

    def __init__(self, Dx=None, Lambda=0, coreEnergy=0, coreEnergyLog=0, rate=None, type='strain',
                 initial=np.array([0., 0., 0.]), initT=0.):
        CentralUpwindHJBetaPDynamics.BetaPDynamics.__init__(self, Dx, coreEnergy, coreEnergyLog, 0,
                                                            **_load_kwargs(rate, type, initial, initT))
        self.rate, self.type, self.initial, self.initT = rate, type, initial, initT


class NewGlideOnlyLoadDynamics(CentralUpwindHJBetaPGlideOnlyDynamics.BetaPDynamics):
    """Glide only ('lvp') under an external strain or stress (RunPython.py:52-66)."""

    def __init__(self, Dx=None, Lambda=0, coreEnergy=0, coreEnergyLog=0, rate=None, type='strain',
                 initial=np.array([0., 0., 0.]), initT=0.):
        CentralUpwindHJBetaPGlideOnlyDynamics.BetaPDynamics.__init__(self, Dx, coreEnergy, coreEnergyLog,
                                                                     **_load_kwargs(rate, type, initial, initT))
        self.rate, self.type, self.initial, self.initT = rate, type, initial, initT


class VacancyDynamicsExternalLoad(BetaP_VacancyDynamics):
    """The vacancy dynamics under an external strain or stress
    (RunPython.py:83-101); Lambda lands on coreEnergy as in the reference."""

    def __init__(self, alpha=1.0, gamma=1.0, beta=1.0, Dx=None, Lambda=0, coreEnergy=0, coreEnergyLog=0, rate=None,
                 type='strain', initial=np.array([0., 0., 0.]), initT=0.):
        BetaP_VacancyDynamics.__init__(self, Dx, coreEnergy, coreEnergyLog, 0, alpha=alpha, gamma=gamma, beta=beta,
                                       **_load_kwargs(rate, type, initial, initT))
        self.rate, self.type, self.initial, self.initT = rate, type, initial, initT


def local_get(store, dir, file):
    src = os.path.join(store, dir, file)
    if not os.path.isfile(src):
        return 1
    sh.copy(src, ".")
    return 0


def local_put(store, dir, file):
    os.makedirs(os.path.join(store, dir), exist_ok=True)
    sh.copy(file, os.path.join(store, dir))
    return 0


def convertStateToCUDA(state, filename):
    from Plasticity.RunCUDA import convertStateToCUDA as conv
    return conv(state, filename)


def record_increment(t):
    """The record spacing of RunPython.py:240-251."""
    if t <= 0.1 - 0.01:
        return 0.01
    if t <= 1.:
        return 0.05
    if t <= 5.:
        return 0.5
    return 2.5


def disorder_field(gridShape, seed, ndisorder=100):
    """The 'x' runs' disorder field (RunPython.py:196-212): ndisorder random
    cells set to 1e-4 rand, gaussian_filter(sigma=1).  The random numbers
    continue numpy's legacy global generator after
    GaussianRandomInitializer(seed) drew its nine 128^d noise fields, as in
    the reference."""
    import scipy.ndimage
    d = len(gridShape)
    rs = np.random.RandomState(0 if seed is None else seed)
    for _ in range(9):
        rs.normal(scale=128, size=(128,) * d)
    field = np.zeros(gridShape)
    if d == 2:
        rx = rs.randint(0, gridShape[0], ndisorder)
        ry = rs.randint(0, gridShape[1], ndisorder)
        field[rx, ry] = 1e-4 * rs.rand(ndisorder)
    else:
        rx = rs.randint(0, gridShape[0], ndisorder)
        ry = rs.randint(0, gridShape[1], ndisorder)
        rz = rs.randint(0, gridShape[2], ndisorder)
        field[rx, ry, rz] = 1e-4 * rs.rand(ndisorder)
    return scipy.ndimage.gaussian_filter(field, sigma=1)


def simulation(dct, get_file=local_get, put_file=local_put, end_time=30., verbose=True):
    """RunPython.simulation (RunPython.py:132-257) on the GPU.

    Keeps the reference's flow: the run's record file (raw, untransposed
    [t, field] records, RecordStateObserverRaw) gets the initial state, then
    the state after each record interval up to end_time (30 in the
    reference).  Differences: an existing archive of this run (local or
    fetched) is resumed from its last record, which the reference leaves
    unimplemented (its state is then undefined); the loading keys are read
    as conf["load_dir"], conf["load_start"] and conf.get("load_type",
    "strain") for every method (the reference reads "load_tye" for 'mdp',
    and "load_rate" for 'lvp'); the reference's commented-out archiving is
    left out too, so the records stay in the working directory.  Returns
    the name of the record file."""
    from Plasticity.PlasticityStates.DisorderState import DisorderState
    from Plasticity.TarFile import LoadTarState, RunArchive
    conf = Configuration(dct)
    homedir, N, dim = conf.homedir, conf.N, conf.dim
    previous, postfix, method, seed = conf.previous, conf.postfix, conf.method, conf.seed
    gridShape = (N,) * dim
    lengthscale = 0.2
    directory = method + str(len(gridShape)) + "d" + str(N)
    unique = directory + "_s" + str(seed)
    oldfile = unique + "_" + previous + ".tar"
    currstub = unique + "_" + previous + postfix
    currfile = currstub + ".tar"
    file_output = currstub + ".plas"

    start = 0.
    if not os.path.isfile(currfile):
        get_file(homedir, directory, currfile)
    if os.path.isfile(currfile):
        if method.endswith('x'):
            raise NotImplementedError("resuming a DisorderState run ('x') from its records")
        with RunArchive(currfile) as arch:
            arch.copy_to(".plas", file_output)
        start, state = FieldInitializer.LoadStateRawUntransposed(file_output, N, dim)
    else:
        if previous == "":
            state = FieldInitializer.GaussianRandomInitializer(gridShape, lengthscale, seed)
        else:
            get_file(homedir, directory, oldfile)
            tt, state = LoadTarState(oldfile)
        convertStateToCUDA(state, currstub + ".ics")
    write_json(conf, currstub + ".conf")

    loaded = "load" in conf
    lkw = dict(rate=conf.get("load_dir"), initial=conf.get("load_start", (0., 0., 0.)),
               type=conf.get("load_type", "strain")) if loaded else {}
    if method.startswith("mdp"):
        dynamics = (UpwindLoadingBetaPDynamics(Lambda=1, **lkw) if loaded
                    else CentralUpwindHJBetaPDynamics.BetaPDynamics(Lambda=1))
    elif method.startswith("gcd"):
        dynamics = (UpwindLoadingBetaPDynamics(Lambda=0, **lkw) if loaded
                    else CentralUpwindHJBetaPDynamics.BetaPDynamics(Lambda=0))
    elif method.startswith("lvp"):
        dynamics = (NewGlideOnlyLoadDynamics(**lkw) if loaded
                    else CentralUpwindHJBetaPGlideOnlyDynamics.BetaPDynamics())
    else:
        raise ValueError("method must start with mdp, gcd or lvp")
    mover = FieldMover.TVDRungeKutta_FieldMover(CFLsafeFactor=0.1)

    if method.endswith('x') and start == 0.:
        d = disorder_field(gridShape, seed)
        field = np.concatenate([d[None], state.GetOrderParameterField().numpy()])
        state = DisorderState(gridShape, field=field, inherit=state)

    obsState = Observer.RecallStateObserver()
    t = start
    recordState = Observer.RecordStateObserverRaw(filename=file_output, mode='wb' if start == 0. else 'ab')
    if start == 0.:
        recordState.Update(t, state)
    observers = [obsState] + ([Observer.VerboseTimestepObserver()] if verbose else [])
    system = PlasticitySystem.PlasticitySystem(gridShape, state, mover, dynamics, observers)
    while t <= end_time:
        preT = t
        t += record_increment(t)
        system.Run(startTime=preT, endTime=t)
        system.state = obsState.state
        recordState.Update(t, system.state)
    recordState.file.close()
    return file_output
