"""MI355X-native Vertex Block Descent, a drop-in for PBAT's ``pbat.gpu.vbd`` path.

Layout mirrors the reference's Python surface (``from pbatoolkit import pbat``):
``sim.vbd.Data`` builds the problem, ``gpu.vbd.Integrator`` steps it with the
HIP kernels of ``libpbat_vbd.so``, ``graph`` and ``fem`` hold the colouring
enums and Lame helper. ``pbat`` is an alias of this package so reference code
written as ``pbat.sim.vbd.Data()`` runs unchanged.
"""
import sys as _sys

from . import fem, graph, meshes  # noqa: F401
from . import sim, gpu  # noqa: F401

pbat = _sys.modules[__name__]

__version__ = "0.1.0"
