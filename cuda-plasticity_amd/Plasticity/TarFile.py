"""Run archives (Plasticity/TarFile.py): a tar of one run directory holding
the configuration (.conf, JSON with at least N and dim), the initial
condition (.ics) and the record file (.plas) of the reference's runs
(RunCUDA.py:120-254 writes them).  The .ics/.plas members are read with
FieldInitializer.LoadStateRaw straight from the archive."""
import os
import tarfile

from Plasticity.Configure import read_json, write_json, open_if_not  # noqa: F401
from Plasticity.FieldInitializers import FieldInitializer


def get_extension(filename):
    return os.path.splitext(filename)[1]


def tar_open(filename):
    return tarfile.open(filename)


def tar_getmem(tar, extension):
    if not isinstance(tar, tarfile.TarFile):
        tar = tarfile.open(tar)
    for mem in tar.getmembers():
        if mem.isfile() and get_extension(mem.name) == extension:
            return mem
    return None


def tar_getfile(tar, extension):
    mem = tar_getmem(tar, extension)
    if mem is None:
        raise KeyError("no %s member in the archive" % extension)
    return tar.extractfile(mem)


def tar_extract(tar, extension, path="."):
    return tar.extract(tar_getmem(tar, extension), path=path)


def LoadTarState(file, time=None, device=None):
    """(t, state) of the .plas record at `time` (default: the last)."""
    with tarfile.open(file) as tar:
        dct = read_json(tar_getfile(tar, ".conf"))
        return FieldInitializer.LoadStateRaw(tar_getfile(tar, ".plas"), int(dct["N"]), int(dct["dim"]), time=time,
                                             device=device)


def LoadTarJSON(file):
    with tarfile.open(file) as tar:
        return read_json(tar_getfile(tar, ".conf"))


def LoadTarICS(file, device=None):
    """(0, state) of the initial condition."""
    with tarfile.open(file) as tar:
        dct = read_json(tar_getfile(tar, ".conf"))
        return FieldInitializer.LoadStateRaw(tar_getfile(tar, ".ics"), int(dct["N"]), int(dct["dim"]),
                                             hastimes=False, device=device)
