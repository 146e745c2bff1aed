"""``pbat.gpu`` counterpart (VBD only)."""
from . import diff  # noqa: F401
from . import vbd  # noqa: F401
