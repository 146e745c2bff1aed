"""``pbat.gpu`` counterpart (VBD only)."""
from . import vbd  # noqa: F401
