"""MI355X (gfx950) C-IPC contact-and-Newton step: host bindings over include/cipc.h."""
from . import scenes  # noqa: F401
from ._lib import LIB_PATH, load  # noqa: F401
from .sim import CipcError, Simulator  # noqa: F401
from .batch import BatchStepper  # noqa: F401
