"""``pbat.sim`` counterpart (VBD only)."""
from . import vbd  # noqa: F401
