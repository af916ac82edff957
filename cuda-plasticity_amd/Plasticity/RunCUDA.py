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
import numbers
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


def read_raw_fields(filename, N, dim, ncomp):
    """The ncomp components of a raw .ics file (no time, every component with
    its grid axes reversed) in the NumPy layout."""
    raw = np.fromfile(filename, dtype=np.float64)
    if raw.size != ncomp * N ** dim:
        raise ValueError("%s holds %d values, not %d components of %d^%d" % (filename, raw.size, ncomp, N, dim))
    return np.stack([_reverse_axes(c) for c in raw.reshape((ncomp,) + (N,) * dim)])


def convertCUDAToState(filename, N, dim, time=None):
    """(t, state) of a .plas record (the last, or the first with t >= time)."""
    return LoadStateRaw(filename, N, dim, time=time)


# ------------------------------------------------------------ headers
# The executable is configured at compile time by a header of #define lines
# (plasticity.h and the #ifdef blocks of plasticity.cu); RunCUDA.simulation
# writes one from a dict (RunCUDA.py:24-100, 195-205).  Each value is a C
# literal: a number as Python prints it, a string in double quotes, a
# sequence in braces ({{0, 0, 1}, ...}), and a bare switch as the empty
# string "".  HEADER_KEYS lists the keys the executable reads and the kind
# of value each takes (the reader parses by the literal's own form, so
# headers with other keys read too).
FLAG = ""
HEADER_KEYS = {
    "N": "int",                   # grid side
    "DIMENSION3": "flag",         # 3D grid (else 2D)
    "CFLsafeFactor": "number",
    "lambda": "number",           # climb mobility switch: 0 'gcd', 1 'mdp'
    "NEWGLIDEONLY": "flag",       # 'lvp' glide-only velocity
    "VACANCIES": "flag",
    "DYNAMIC_NUCLEATION": "int",
    "LOADING": "flag",
    "LOAD_DEF": "matrix",         # 3x3 strain direction of the load
    "LOADING_RATE": "number",
    "LOAD_START": "number",
    "FILE_INPUT": "string",       # sprintf patterns of (N, seed, lambda)
    "FILE_OUTPUT": "string",
    "TIME_STEP": "number",        # record spacing
    "TIME_FINAL": "number",
    "vacancycost": "number",      # VACANCIES: sigma_ii -= vacancycost c
    "vacancydiffusion": "number",  # VACANCIES: c diffuses by exp(-vacancydiffusion dt k^2)
}
# the CFL factor RunCUDA compiles every run with (RunCUDA.py:54-58)
EXECUTABLE_CFL = 0.5
# the switches each dynamics method of RunCUDA compiles to; 'slip' also
# takes the slip system's own keys (the reference marks 'slip' and 'vac' as
# incomplete; 'vac' sets VACANCIES alone, which the executable here runs
# only beside a dynamics method that defines lambda)
METHOD_DEFINES = {
    "gcd": {"lambda": 0},
    "mdp": {"lambda": 1},
    "lvp": {"lambda": 0, "NEWGLIDEONLY": FLAG},
    "slip": {"lambda": 0},
    "vac": {"VACANCIES": FLAG},
}


def c_literal(value):
    """The C text of a header value."""
    if isinstance(value, str):
        return '"%s"' % value.replace("\\", "\\\\").replace('"', '\\"')
    if isinstance(value, numbers.Integral):
        return str(int(value))
    if isinstance(value, numbers.Real):
        return repr(float(value))
    if isinstance(value, np.ndarray):
        value = value.tolist()
    if isinstance(value, (list, tuple)):
        return "{" + ", ".join(c_literal(v) for v in value) + "}"
    raise TypeError("no C literal for %r" % (value,))


def parse_c_literal(text):
    """Python data of a header value's C text (the inverse of c_literal);
    text that is no literal (a C expression) is returned as it is."""
    text = text.strip()
    if text == "":
        return FLAG
    try:
        value, rest = _parse_literal(text, 0)
    except (ValueError, IndexError):
        return text
    return value if text[rest:].strip() == "" else text


def _parse_literal(s, i):
    while s[i] == " ":
        i += 1
    if s[i] == '"':
        out, i = [], i + 1
        while s[i] != '"':
            if s[i] == "\\":
                i += 1
            out.append(s[i])
            i += 1
        return "".join(out), i + 1
    if s[i] == "{":
        items, i = [], i + 1
        while True:
            while s[i] == " ":
                i += 1
            if s[i] == "}":
                return items, i + 1
            v, i = _parse_literal(s, i)
            items.append(v)
            while s[i] == " ":
                i += 1
            if s[i] == ",":
                i += 1
    j = i
    while j < len(s) and s[j] not in ",} ":
        j += 1
    tok = s[i:j]
    try:
        return int(tok), j
    except ValueError:
        return float(tok), j


def write_header(dct, headername):
    """One '#define KEY VALUE' line per entry, to a path or an open file."""
    text = "".join("#define %s %s\n" % (key, c_literal(val)) for key, val in dct.items())
    f = open_if_not(headername, "w")
    f.write(text)
    if f is not headername:
        f.close()


def read_header(header):
    """The dict of a header's #define lines (path or open file), values as
    Python data; other lines (comments, guards) are skipped."""
    f = open_if_not(header, "r")
    try:
        lines = f.read().splitlines()
    finally:
        if f is not header:
            f.close()
    dct = {}
    for line in lines:
        words = line.strip().split(None, 2)
        if len(words) >= 2 and words[0] == "#define":
            dct[words[1]] = parse_c_literal(words[2] if len(words) == 3 else "")
    return dct


def run_defines(N=None, dim=None, method=None, nucleation=False, files=None, times=None, load=None, **extra):
    """The #define set of one run, from the tables above: grid (N, dim),
    dynamics method, dynamic nucleation, (input, output) file patterns,
    (record spacing, final time) and load (direction, rate, start)."""
    d = {}
    if N is not None:
        d.update(N=int(N), CFLsafeFactor=EXECUTABLE_CFL)
        if dim == 3:
            d["DIMENSION3"] = FLAG
    if method is not None:
        d.update(METHOD_DEFINES.get(method, {}))
        if method == "slip":
            d.update(extra)
    if nucleation:
        d["DYNAMIC_NUCLEATION"] = 0
    if files is not None:
        d["FILE_INPUT"], d["FILE_OUTPUT"] = files
    if load is not None:
        direction, rate, start = load
        d.update(LOADING=FLAG, LOAD_DEF=np.asarray(direction).tolist(), LOADING_RATE=rate, LOAD_START=start)
    if times is not None:
        d["TIME_STEP"], d["TIME_FINAL"] = times
    return d


# the reference's header builders (RunCUDA.py:54-100), each adding its part
def header_size(header, N, dim):
    header.update(run_defines(N=N, dim=dim))


def header_load(header, direction, rate, start=0.0):
    header.update(run_defines(load=(direction, rate, start)))


def header_dynamics(header, dyn, **kwargs):
    header.update(run_defines(method=dyn, **kwargs))


def header_dynamic_nucleation(header, dn=True):
    header.update(run_defines(nucleation=dn is True))


def header_files(header, infile, outfile):
    header.update(run_defines(files=(infile, outfile)))


def header_times(header, step, end):
    header.update(run_defines(times=(step, end)))

# ------------------------------------------------------------ the executable
# Compile-time options of the executable that this runner does not carry out.
UNSUPPORTED = ("DEBUG_TIMESTEPS", "SINGLE_STEP_DEBUG")
# The executable's VACANCIES constants; the reference's headers leave them
# undefined, so they are read from the header when present and otherwise
# take the defaults of the Python vacancy dynamics (alpha = gamma = 1,
# FieldDynamics/VacancyDynamics.py:16-17).
VACANCY_DEFAULTS = {"vacancycost": 1.0, "vacancydiffusion": 1.0}


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
    if "lambda" not in h:
        # the reference's 'vac' method sets VACANCIES alone (RunCUDA.py:84-85)
        raise ValueError("the header defines no lambda: combine VACANCIES with a dynamics method (gcd, mdp, lvp)")
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


def nucleation_cap(plan, src):
    """maxNucleationTimestep (plasticity.cu:190-199): 1 / sqrt(|s|) for the
    entry s of largest magnitude over all nine components of the stress of
    src (calculateSigma, no load)."""
    from Plasticity.amd.lib import Plan
    p = Plan(plan.ndim, plan.n, mu=MU, nu=NU, device=plan.device)
    s = p.sigma(src)
    p.close()
    return 1.0 / float(s.abs().max()) ** 0.5


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
    plasticity.cu:105-122).

    DYNAMIC_NUCLEATION (plasticity.cu:178-201): the loaded field (the input,
    or on a resume the last record) becomes the source beta0dot, added to
    every flux as beta0dot / TIME_FINAL, the state starts from zero, and the
    time step is capped at 1 / sqrt(|sigma(beta0dot)|) of its largest stress
    entry; the t = 0 record holds the loaded field, as the executable writes
    its host copy before the first step.

    VACANCIES (plasticity_kernel.cu:176-196,284-290,676-775): ten
    components, the tensor and the vacancy concentration c (in the .ics
    input and every .plas record); the stress loses vacancycost c on the
    diagonal, c's rhs is the trace of the tensor rhs, and c diffuses by
    exp(-vacancydiffusion dt k^2) after every forward Euler step.  Returns
    (t, state tensor on the GPU; VACANCIES: the ten components)."""
    plan, h = header_plan(header, device=device, rk_order=rk_order)
    N, dim = plan.n, plan.ndim
    args = (N, seed, float(h["lambda"]))
    out = os.path.join(directory, _cformat(str(h["FILE_OUTPUT"]), args))
    nucleation = "DYNAMIC_NUCLEATION" in h
    vacancies = "VACANCIES" in h
    if vacancies:
        # ten components: the tensor and the vacancy concentration, whose rhs
        # is the tensor rhs's trace (the executable's forward Euler step)
        if nucleation:
            raise NotImplementedError("VACANCIES with DYNAMIC_NUCLEATION")
        vc = float(h.get("vacancycost", VACANCY_DEFAULTS["vacancycost"]))
        vd = float(h.get("vacancydiffusion", VACANCY_DEFAULTS["vacancydiffusion"]))
        plan.set_vacancy(alpha=vc, beta=1.0, gamma=vd, dt_bound=None)
    t = 0.0
    if os.path.isfile(out) and os.path.getsize(out) > 0:
        t, _, _ = plan.plas_resume(out)
        if verbose:
            print(" Restarting from t=%f" % t)
    else:
        if state is None and vacancies:
            state = read_raw_fields(os.path.join(directory, _cformat(str(h["FILE_INPUT"]), args)), N, dim, 10)
        elif state is None:
            _, state = LoadStateRaw(os.path.join(directory, _cformat(str(h["FILE_INPUT"]), args)), N, dim,
                                    hastimes=False)
        # a PlasticityState, or the nine (VACANCIES: ten) components as an array or tensor
        data = state.GetOrderParameterField().data if hasattr(state, "GetOrderParameterField") else state
        if vacancies:
            plan.set_state(data[:9])
            plan.vacancy_set(data[9])
        else:
            plan.set_state(data)
        plan.set_time(0.0)
        plan.plas_append(out)  # the t = 0 record (plasticity.cu:232-233)
    if nucleation:
        src = plan.get_state()
        plan.set_source(src, 1.0 / float(h["TIME_FINAL"]), nucleation_cap(plan, src))
        plan.set_state(plan.torch.zeros_like(src))
        plan.set_time(t)
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
    if vacancies:
        return t, plan.torch.cat([plan.get_state(), plan.vacancy_get()[None]])
    return t, plan.get_state()


# ------------------------------------------------------------ run archives
def header_fromfile(filename):
    """The header dict archived with a run (its .json member)."""
    from Plasticity.TarFile import RunArchive, read_json
    with RunArchive(filename) as arch:
        return read_json(arch.open(".json"))


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
    from Plasticity.TarFile import LoadTarJSON, LoadTarState, RunArchive
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
        with RunArchive(currfile) as arch:
            arch.copy_to(".plas", file_output)
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
