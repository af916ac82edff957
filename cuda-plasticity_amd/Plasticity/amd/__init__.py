"""GPU runtime of the stepper: the ctypes binding of libcdd.so (``lib``
module) and the per-configuration plan cache (``plans``)."""
from . import lib  # noqa: F401  (the module; the CDLL handle is lib.lib)
from .lib import Config, Plan, CDDError  # noqa: F401
from . import plans  # noqa: F401
