"""The reference's native path (Plasticity/RunCUDA.py) in process.

The reference writes the initial state to a .ics file and a C header of
#defines, compiles the CUDA executable against that header, runs it and
tars the .plas records it writes (RunCUDA.py:120-254).  Here the header is
interpreted instead of compiled: run_header performs the executable's
runTest (plasticity.cu:40-290) on the GPU with the executable's numerics
(the LLF scheme of Plasticity.amd, forward Euler as plasticity.h's
TVDstep = simpleTVD), reading and writing the same .ics and .plas files.
simulation() keeps RunCUDA.simulation's file and archive flow around it.
"""
import ast
import os
import re
import shutil as sh
import tarfile

import numpy as np

from Plasticity.Configure import Configuration, open_if_not, write_json
from Plasticity.FieldInitializers import FieldInitializer
from Plasticity.FieldInitializers.FieldInitializer import LoadStateRaw, _reverse_axes

# plasticity.h:14-15
MU, NU = 0.5, 0.3


def convertStateToCUDA(state, filename):
    """Write the nine components, each with all grid axes reversed, as raw
    float64 (RunCUDA.py:9-22)."""
    f = state.GetOrderParameterField()
    data = f.numpy() if hasattr(f, "numpy") else np.asarray(f)
    np.stack([_reverse_axes(np.asarray(data[c], dtype=np.float64)) for c in range(9)]).tofile(filename)


def convertCUDAToState(filename, N, dim, time=None):
    """(t, state) of a .plas record (the last, or the first with t >= time)."""
    return LoadStateRaw(filename, N, dim, time=time)


# ------------------------------------------------------------ headers
# (RunCUDA.py:24-100): a dict of #define names; a value of "" is a bare flag.
def write_header(dct, headername):
    headerfile = open_if_not(headername, "w")
    for key, val in dct.items():
        headerfile.write("#define " + key + " " + list_to_crepr(val) + "\n")
    if headerfile is not headername:
        headerfile.close()


def read_header(header):
    define_statement = re.compile(r"\#define\s([a-zA-Z0-9_\(\),]*)\s?(.*)$")
    dct = {}
    headerfile = open_if_not(header, "r")
    lines = headerfile.readlines()
    if headerfile is not header:
        headerfile.close()
    for line in lines:
        m = define_statement.match(line.strip())
        if m is None:
            print("Not a valid line: ", line.strip())
            continue
        var, val = m.groups()
        dct[var] = crepr_to_list(val)
    return dct


def list_to_crepr(arr):
    return repr(arr).replace("[", "{").replace("]", "}").replace("'", "\"")


def crepr_to_list(arr):
    return arr.replace("{", "[").replace("}", "]").replace("\"", "'")


def header_size(header, N, dim):
    header.update({"N": N})
    header.update({"CFLsafeFactor": 0.5})
    if dim == 3:
        header.update({"DIMENSION3": ""})


def header_load(header, direction, rate, start=0.0):
    header.update({"LOADING": ""})
    if isinstance(direction, np.ndarray):
        header.update({"LOAD_DEF": direction.tolist()})
    else:
        header.update({"LOAD_DEF": direction})
    header.update({"LOADING_RATE": rate})
    header.update({"LOAD_START": start})


def header_dynamics(header, dyn, **kwargs):
    if dyn == "gcd":
        header.update({"lambda": 0})
    if dyn == "mdp":
        header.update({"lambda": 1})
    if dyn == "lvp":
        header.update({"lambda": 0})
        header.update({"NEWGLIDEONLY": ""})
    if dyn == "slip":
        header.update({"lambda": 0})
        header.update(kwargs)
    if dyn == "vac":
        header.update({"VACANCIES": ""})


def header_dynamic_nucleation(header, dn=True):
    if dn is True:
        header.update({"DYNAMIC_NUCLEATION": 0})


def header_files(header, infile, outfile):
    header.update({"FILE_INPUT": infile})
    header.update({"FILE_OUTPUT": outfile})


def header_times(header, step, end):
    header.update({"TIME_STEP": step})
    header.update({"TIME_FINAL": end})


# ------------------------------------------------------------ the executable
# Compile-time options of the executable that this runner does not carry out.
UNSUPPORTED = ("VACANCIES", "DYNAMIC_NUCLEATION", "DEBUG_TIMESTEPS", "SINGLE_STEP_DEBUG")


def _value(v):
    """A header value as Python data: read_header leaves the C text
    ('0.5', '[[0, 0, 1], ...]', "'run.plas'"); write_header's input is data."""
    if isinstance(v, str):
        try:
            return ast.literal_eval(v)
        except (ValueError, SyntaxError):
            return v
    return v


def _cformat(fmt, args):
    """sprintf(fn, FILE_..., N, seed, lambda) (plasticity.cu:98,134): as many
    arguments as the name has conversions."""
    n = len(re.findall(r"%[-+ #0]*\d*(?:\.\d+)?[a-zA-Z]", fmt.replace("%%", "")))
    return fmt % tuple(args[:n]) if n else fmt


def default_time_increment(t):
    """Record spacing without TIME_STEP (plasticity.cu:248-259)."""
    if t <= 0.1:
        return 0.01
    if t <= 1.0:
        return 0.05
    if t <= 5.0:
        return 0.5
    return 1.0


def header_options(header):
    """(Plan arguments, parsed header) of the stepper a header compiles to:
    the LLF scheme with lambda (or NEWGLIDEONLY), CFLsafeFactor, mu = 0.5,
    nu = 0.3 and, with LOADING, the executable's load (loadSigma,
    plasticity_kernel.cu:1045-1066):
        sigma_ij += (2 mu L_ij + lambda_e tr(L) d_ij) (LOAD_START + LOADING_RATE t),
    lambda_e = 2 mu nu / (1 - 2 nu), a stress tensor linear in t."""
    from Plasticity.amd.lib import GLIDE_CLIMB, GLIDE_ONLY, LOAD_TENSOR
    h = {k: _value(v) for k, v in header.items()}
    for k in UNSUPPORTED:
        if k in h:
            raise NotImplementedError("the in-process executable does not implement %s" % k)
    glide = "NEWGLIDEONLY" in h
    lam = float(h["lambda"])
    kw = dict(ndim=3 if "DIMENSION3" in h else 2, n=int(h["N"]), dynamics=GLIDE_ONLY if glide else GLIDE_CLIMB,
              Lambda=0.0 if glide else lam, mu=MU, nu=NU, cfl=float(h["CFLsafeFactor"]), scheme="llf")
    if "LOADING" in h:
        L = np.asarray(h["LOAD_DEF"], dtype=float).reshape(3, 3)
        if not np.array_equal(L, L.T):
            raise NotImplementedError("LOAD_DEF must be symmetric (the stress is stored as six components)")
        lam_e = 2.0 * MU * NU / (1.0 - 2.0 * NU)
        S = 2.0 * MU * L + lam_e * np.trace(L) * np.eye(3)
        start, rate = float(h.get("LOAD_START", 0.0)), float(h["LOADING_RATE"])
        kw.update(load_kind=LOAD_TENSOR, load_init=(S * start).ravel(), load_rate=(S * rate).ravel(), load_t0=0.0)
    return kw, h


def header_plan(header, device=0, rk_order=1):
    """The header's stepper (header_options) as a Plan; rk_order 1 is the
    executable's TVDstep (simpleTVD), 3 its TVD3rd."""
    from Plasticity.amd.lib import Plan
    kw, h = header_options(header)
    return Plan(rk_order=rk_order, device=device, **kw), h


def run_header(header, state=None, seed=0, device=0, rk_order=1, directory=".", verbose=False):
    """The executable's runTest (plasticity.cu:40-290) for a header dict.

    The state comes from `state` or the .ics file FILE_INPUT; records
    [t, state] are appended to FILE_OUTPUT at t = 0 and at the end of every
    interval of TIME_STEP (or the default spacing) until TIME_FINAL.  Each
    interval ends with the step whose dt = t_end - t + ME carries t just past
    it (simpleTVD, plasticity_kernel.cu:1124-1125).  An existing non-empty
    FILE_OUTPUT is resumed from its last record (CONTINUE_RUN, plasticity.h:10,
    plasticity.cu:105-122).  Returns (t, state tensor on the GPU)."""
    plan, h = header_plan(header, device=device, rk_order=rk_order)
    N, dim = plan.n, plan.ndim
    args = (N, seed, float(h["lambda"]))
    out = os.path.join(directory, _cformat(str(h["FILE_OUTPUT"]), args))
    t = 0.0
    if os.path.isfile(out) and os.path.getsize(out) > 0:
        t, _, _ = plan.plas_resume(out)
        if verbose:
            print(" Restarting from t=%f" % t)
    else:
        if state is None:
            _, state = LoadStateRaw(os.path.join(directory, _cformat(str(h["FILE_INPUT"]), args)), N, dim,
                                    hastimes=False)
        # a PlasticityState, or the nine components as an array or tensor
        plan.set_state(state.GetOrderParameterField().data if hasattr(state, "GetOrderParameterField") else state)
        plan.set_time(0.0)
        plan.plas_append(out)  # the t = 0 record (plasticity.cu:232-233)
    end = float(h["TIME_FINAL"])
    timed = "LOADING" in h or "TIME_STEP" in h
    while t < end:
        inc = float(h["TIME_STEP"]) if timed else default_time_increment(t)
        plan.set_time(t, t + inc)
        chunk = 8
        while True:
            plan.step(chunk)
            t, dt, done, ns = plan.clock()
            if done:
                break
            chunk = min(2 * chunk, 256)
        if plan.status():
            raise FloatingPointError("a non-finite speed stopped the run at t=%g" % t)
        if verbose:
            print("%le +%le" % (t, dt))
        plan.plas_append(out)
    return t, plan.get_state()


# ------------------------------------------------------------ run archives
def header_fromfile(filename):
    from Plasticity.TarFile import read_json, tar_getfile
    with tarfile.open(filename) as tar:
        return read_json(tar_getfile(tar, ".json"))


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


def simulation(dct, get_file=local_get, put_file=local_put, delete_intermediates=True, resume=True, save_global=True,
               rk_order=1):
    """RunCUDA.simulation (RunCUDA.py:120-254) with the executable run in
    process: the same run names (<method><dim>d<N>_s<seed>_<previous><postfix>),
    .ics/.conf/.json/.h files, .plas records and the tar archive stored
    through put_file.  Differences: nothing is compiled, so no .exe is
    archived; a continuation fetches the previous archive before reading its
    configuration, which it reads as a Configuration
    (the reference reads attributes of the plain dict LoadTarJSON returns),
    and its LOAD_START is taken from the configuration's load_start or else
    from the previous archive's header, and stored as load_start in this
    run's configuration so that loading runs chain."""
    from Plasticity.TarFile import LoadTarJSON, LoadTarState, tar_extract, tar_open
    conf = Configuration(dct)
    homedir = conf.homedir
    N, dim = conf.N, conf.dim
    previous, postfix, method = conf.previous, conf.postfix, conf.method
    device, seed = conf.get("device", 0), conf.seed
    gridShape = (N,) * dim
    lengthscale = 0.2

    directory = method + str(len(gridShape)) + "d" + str(N)
    unique = directory + "_s" + str(seed)
    oldfile = unique + "_" + previous + ".tar"
    currstub = unique + "_" + previous + postfix
    currfile = currstub + ".tar"

    loading = postfix not in ("r", "d", "")
    if loading:
        if postfix in ("0", "1"):
            load_direction, load_rate = conf.load_direction, conf.load_rate
            time_start, load_start = 50.0, 0.0
        else:
            if not os.path.isfile(oldfile):
                get_file(homedir, directory, oldfile)
            old = Configuration(LoadTarJSON(oldfile))
            load_direction, load_rate = old.load_direction, old.load_rate
            time_start = conf.time_start
            prev_start = old.get("load_start")
            if prev_start is None:
                prev_start = float(_value(header_fromfile(oldfile).get("LOAD_START", 0.0)))
            load_start = time_start * load_rate + prev_start
        conf.load_start = load_start
        time_step, time_end = conf.time_step, conf.time_end
    else:
        time_step, time_end = 2.0, 50.0

    file_output = currstub + ".plas"
    file_input = currstub + ".ics"
    if not os.path.isfile(currfile) and resume:
        get_file(homedir, directory, currfile)
    if os.path.isfile(currfile) and resume:
        # resume from the archived records: the output exists, so the run
        # continues from its last record (CONTINUE_RUN)
        with tar_open(currfile) as tar:
            tar_extract(tar, ".plas")
        sh.move(os.path.join(currstub, file_output), file_output)
        sh.rmtree(currstub)
        file_input = file_output
    else:
        if previous == "":
            state = FieldInitializer.GaussianRandomInitializer(gridShape, lengthscale, seed)
        else:
            if not os.path.isfile(oldfile):
                get_file(homedir, directory, oldfile)
            tt, state = LoadTarState(oldfile, time=time_start)
        convertStateToCUDA(state, file_input)
    confname = currstub + ".conf"
    write_json(conf, confname)

    header = {"blank": "nothing"}
    header_size(header, N, dim)
    header_dynamics(header, method)
    header_dynamic_nucleation(header, postfix.find("d") >= 0)
    header_files(header, file_input, file_output)
    if loading:
        header_load(header, load_direction, load_rate, load_start)
    header_times(header, time_step, time_end)
    headername = currstub + ".h"
    jsonname = currstub + ".json"
    write_json(header, jsonname)
    write_header(header, headername)

    run_header(header, device=device, rk_order=rk_order)

    os.makedirs(currstub, exist_ok=True)
    for f in (jsonname, headername, confname, file_input, file_output):
        if os.path.isfile(f):
            sh.move(f, os.path.join(currstub, f))
    with tarfile.open(currfile, "w") as tar:
        tar.add(currstub)
    sh.rmtree(currstub)
    if save_global:
        put_file(homedir, directory, currfile)
    return currfile
