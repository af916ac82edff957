"""The header layer of the in-process CUDA executable (Plasticity/RunCUDA.py
write_header/read_header/header_* of RunCUDA.py:24-100, and what the
executable makes of a header: plasticity.h, plasticity.cu:98-259,
plasticity_kernel.cu loadSigma :1045-1066)."""
import io

import numpy as np
import pytest

from Plasticity import RunCUDA as R


def _header(dim=3, method="mdp", loading=True):
    h = {"blank": "nothing"}
    R.header_size(h, 32, dim)
    R.header_dynamics(h, method)
    R.header_files(h, "run.ics", "run_%d_%d_%f.plas")
    if loading:
        R.header_load(h, np.array([[0.0, 0.5, 0.0], [0.5, 0.0, 0.0], [0.0, 0.0, -0.2]]), 0.05, 0.01)
    R.header_times(h, 0.5, 10.0)
    return h


def test_header_text_round_trip():
    h = _header()
    buf = io.StringIO()
    R.write_header(h, buf)
    text = buf.getvalue()
    assert "#define N 32\n" in text and "#define DIMENSION3 \"\"\n" in text
    assert "#define LOAD_DEF {{0.0, 0.5, 0.0}, {0.5, 0.0, 0.0}, {0.0, 0.0, -0.2}}\n" in text
    back = R.read_header(io.StringIO(text))
    assert set(back) == set(h)
    for k, v in h.items():
        assert R._value(back[k]) == v, k


def test_file_names_take_the_executables_sprintf_arguments():
    assert R._cformat("run_%d_%d_%f.plas", (32, 3, 1.0)) == "run_32_3_1.000000.plas"
    assert R._cformat("in_%d_%d.ics", (32, 3, 1.0)) == "in_32_3.ics"
    assert R._cformat("plain.plas", (32, 3, 1.0)) == "plain.plas"


def test_default_record_spacing():
    assert [R.default_time_increment(t) for t in (0.0, 0.1, 0.2, 1.0, 3.0, 5.0, 7.0)] == \
        [0.01, 0.01, 0.05, 0.05, 0.5, 0.5, 1.0]


def test_header_options_dynamics_and_sizes():
    from Plasticity.amd.lib import GLIDE_CLIMB, GLIDE_ONLY
    kw, _ = R.header_options(_header(dim=2, method="lvp", loading=False))
    assert kw["ndim"] == 2 and kw["n"] == 32 and kw["dynamics"] == GLIDE_ONLY and kw["Lambda"] == 0.0
    assert kw["scheme"] == "llf" and kw["cfl"] == 0.5 and (kw["mu"], kw["nu"]) == (0.5, 0.3)
    assert "load_kind" not in kw
    kw, _ = R.header_options(_header(dim=3, method="mdp", loading=False))
    assert kw["ndim"] == 3 and kw["dynamics"] == GLIDE_CLIMB and kw["Lambda"] == 1.0
    kw, _ = R.header_options(_header(dim=3, method="gcd", loading=False))
    assert kw["Lambda"] == 0.0


def test_header_load_is_hookes_law_of_the_load_strain():
    """loadSigma adds the isotropic stress of the strain LOAD_DEF times
    LOAD_START + LOADING_RATE t: trace 3K tr(L) with the bulk modulus
    K = 2 mu (1 + nu) / (3 (1 - 2 nu)), deviator 2 mu dev(L)."""
    from Plasticity.amd.lib import LOAD_TENSOR
    kw, _ = R.header_options(_header())
    L = np.array([[0.0, 0.5, 0.0], [0.5, 0.0, 0.0], [0.0, 0.0, -0.2]])
    mu, nu = 0.5, 0.3
    K = 2 * mu * (1 + nu) / (3 * (1 - 2 * nu))
    assert kw["load_kind"] == LOAD_TENSOR and kw["load_t0"] == 0.0
    for arr, amt in ((kw["load_init"], 0.01), (kw["load_rate"], 0.05)):
        S = np.asarray(arr).reshape(3, 3)
        assert abs(np.trace(S) - 3 * K * np.trace(L) * amt) < 1e-15
        dev = S - np.trace(S) / 3 * np.eye(3)
        np.testing.assert_allclose(dev, 2 * mu * (L - np.trace(L) / 3 * np.eye(3)) * amt, rtol=0, atol=1e-16)


def test_unsupported_options_raise():
    h = _header()
    R.header_dynamic_nucleation(h)
    kw, hh = R.header_options(h)  # dynamic nucleation is carried out (run_header)
    assert "DYNAMIC_NUCLEATION" in hh
    h = _header(method="vac")
    with pytest.raises(ValueError, match="lambda"):
        R.header_options(h)
    h = _header(method="mdp")
    R.header_dynamics(h, "vac")
    assert "VACANCIES" in R.header_options(h)[1]
    h = _header()
    h["LOAD_DEF"] = [[0, 1, 0], [0, 0, 0], [0, 0, 0]]
    with pytest.raises(NotImplementedError):
        R.header_options(h)


def test_header_reads_reference_style_text():
    """A header as the reference's writer prints it (repr-based C text, keys
    in any order, a stray non-define line) reads back as data."""
    text = ('#define blank "nothing"\n#define N 64\n#define CFLsafeFactor 0.5\n#define DIMENSION3 ""\n'
            '#define lambda 1\n#define LOADING ""\n#define LOAD_DEF {{0, 0, 0}, {0, 0, 0}, {0, 0, 1}}\n'
            '#define LOADING_RATE 0.05\n#define LOAD_START 0.0\n#define FILE_INPUT "in_%d.ics"\n'
            '#define FILE_OUTPUT "out_%d_%d_%f.plas"\n#define TIME_STEP 2.0\n#define TIME_FINAL 50.0\n'
            '#define DYNAMIC_NUCLEATION 0\n// a comment\n')
    h = R.read_header(io.StringIO(text))
    assert h["N"] == 64 and h["lambda"] == 1 and h["DIMENSION3"] == "" and h["LOADING"] == ""
    assert h["LOAD_DEF"] == [[0, 0, 0], [0, 0, 0], [0, 0, 1]] and h["LOADING_RATE"] == 0.05
    assert h["FILE_OUTPUT"] == "out_%d_%d_%f.plas" and h["DYNAMIC_NUCLEATION"] == 0
    assert len(h) == 14
    # and writes back to the same lines
    buf = io.StringIO()
    R.write_header(h, buf)
    assert sorted(buf.getvalue().splitlines()) == sorted(text.splitlines()[:-1])


def test_c_literals():
    assert R.c_literal('a "q" \\ b') == '"a \\"q\\" \\\\ b"'
    assert R.parse_c_literal(R.c_literal('a "q" \\ b')) == 'a "q" \\ b'
    assert R.c_literal(np.float64(0.25)) == "0.25" and R.c_literal(np.int64(3)) == "3"
    assert R.c_literal(np.eye(2)) == "{{1.0, 0.0}, {0.0, 1.0}}"
    assert R.parse_c_literal("{ {1, -2.5e-3} ,{}}") == [[1, -0.0025], []]
    assert R.parse_c_literal("(N*2)") == "(N*2)"
    with pytest.raises(TypeError):
        R.c_literal(object())


def test_run_defines_tables():
    h = R.run_defines(N=16, dim=2, method="lvp", files=("a.ics", "b.plas"), times=(1.0, 5.0),
                      load=([[0, 0, 0], [0, 0, 0], [0, 0, 1]], 0.1, 0.2))
    assert "DIMENSION3" not in h and h["NEWGLIDEONLY"] == "" and h["lambda"] == 0
    assert (h["FILE_INPUT"], h["FILE_OUTPUT"], h["TIME_STEP"], h["TIME_FINAL"]) == ("a.ics", "b.plas", 1.0, 5.0)
    assert h["LOADING"] == "" and (h["LOADING_RATE"], h["LOAD_START"]) == (0.1, 0.2)
    assert set(h) <= set(R.HEADER_KEYS)
    assert R.run_defines(method="slip", SLIP_PLANE=3) == {"lambda": 0, "SLIP_PLANE": 3}
    assert R.run_defines(method="gcd", SLIP_PLANE=3) == {"lambda": 0}
