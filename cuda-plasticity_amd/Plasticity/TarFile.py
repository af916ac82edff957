"""Run archives: a tar of one run directory (what RunCUDA.simulation and
RunPython.simulation store after a run, Plasticity/RunCUDA.py:120-254).

Member roles are told apart by file extension, one member per role:
  .conf  the run configuration, JSON (at least N and dim)
  .ics   the initial state, raw float64 without times (LoadStateRaw layout)
  .plas  the record file, [t, state] records
  .json  the executable's header as JSON (RunCUDA)
  .h     the same header as #define text (RunCUDA)
The module functions keep the reference's names (Plasticity/TarFile.py)
and work on a RunArchive, which indexes the members once.  Member names are
checked before anything is written to disk: an absolute name or one with a
'..' component is refused (Python 3.10's tarfile applies no extraction
filter), and extraction copies the member's bytes to a path the caller
names instead of recreating the archive's directory tree.
"""
import os
import shutil
import tarfile

from Plasticity.Configure import read_json, write_json, open_if_not  # noqa: F401
from Plasticity.FieldInitializers import FieldInitializer

ROLES = (".conf", ".ics", ".plas", ".json", ".h")


def get_extension(filename):
    return os.path.splitext(filename)[1]


def _safe_name(name):
    parts = name.replace("\\", "/").split("/")
    if name.startswith(("/", "\\")) or ".." in parts or (parts and parts[0].endswith(":")):
        raise ValueError("unsafe member name in run archive: %r" % name)
    return name


class RunArchive:
    """An open run archive with its regular-file members indexed by role
    (extension; the first member of each role counts)."""

    def __init__(self, source):
        self._own = not isinstance(source, tarfile.TarFile)
        self.tar = tarfile.open(source) if self._own else source
        self.by_role = {}
        for info in self.tar.getmembers():
            if info.isfile():
                self.by_role.setdefault(get_extension(info.name), info)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        if self._own:
            self.tar.close()

    def member(self, role):
        try:
            return self.by_role[role]
        except KeyError:
            raise KeyError("run archive has no %s member" % role) from None

    def open(self, role):
        """File object reading the member of that role."""
        return self.tar.extractfile(self.member(role))

    def copy_to(self, role, dest):
        """Write the member's bytes to the file `dest`; returns dest."""
        info = self.member(role)
        _safe_name(info.name)
        with self.open(role) as src, open(dest, "wb") as out:
            shutil.copyfileobj(src, out)
        return dest

    def configuration(self):
        return read_json(self.open(".conf"))

    def state(self, role, time=None, device=None):
        conf = self.configuration()
        return FieldInitializer.LoadStateRaw(self.open(role), int(conf["N"]), int(conf["dim"]), time=time,
                                             hastimes=(role == ".plas"), device=device)


def tar_open(filename):
    return tarfile.open(filename)


def tar_getmem(tar, extension):
    """The member of that role, or None."""
    return RunArchive(tar).by_role.get(extension)


def tar_getfile(tar, extension):
    return RunArchive(tar).open(extension)


def tar_extract(tar, extension, path="."):
    """Extract the member of that role under `path`, at its name inside the
    archive (the reference's tar.extract), after checking that name."""
    arch = RunArchive(tar)
    info = arch.member(extension)
    dest = os.path.join(path, _safe_name(info.name))
    os.makedirs(os.path.dirname(dest) or ".", exist_ok=True)
    return arch.copy_to(extension, dest)


def LoadTarState(file, time=None, device=None):
    """(t, state) of the .plas record at `time` (default: the last)."""
    with RunArchive(file) as arch:
        return arch.state(".plas", time=time, device=device)


def LoadTarJSON(file):
    with RunArchive(file) as arch:
        return arch.configuration()


def LoadTarICS(file, device=None):
    """(0, state) of the initial condition."""
    with RunArchive(file) as arch:
        return arch.state(".ics", device=device)
