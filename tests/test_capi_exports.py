"""The C-ABI library loads and exports every symbol include/cdd.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols():
    hdr = open(os.path.join(ROOT, "include", "cdd.h")).read()
    names = set(re.findall(r"\b(cdd_[a-z_]+)\s*\(", hdr))
    assert "cdd_step" in names and "cdd_plan_create" in names
    from Plasticity.amd.lib import LIBPATH, EXPORTS
    L = ctypes.CDLL(LIBPATH)
    for n in sorted(names):
        assert hasattr(L, n), n
    assert names == set(EXPORTS)


def test_single_hip_runtime():
    """The binding maps exactly one HIP runtime (torch's): a second copy from
    /opt/rocm makes the later of the two fail to find the device."""
    from Plasticity.amd import lib
    rts = lib._hip_runtimes()
    assert len(rts) == 1, rts
