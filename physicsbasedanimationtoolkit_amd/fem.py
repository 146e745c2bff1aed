"""FEM helpers used when setting up VBD materials."""
from __future__ import annotations

import numpy as np


def lame_coefficients(Y, nu):
    """Lame coefficients from Young's modulus and Poisson's ratio
    (python/pbatoolkit/py/fem.py:6-17; physics/HyperElasticity.cpp:6-11).

    Returns ``(mu, lambda)`` with mu = Y / (2 (1 + nu)) and
    lambda = Y nu / ((1 + nu)(1 - 2 nu)).
    """
    Y = np.asarray(Y, dtype=np.float64) if not np.isscalar(Y) else float(Y)
    nu = np.asarray(nu, dtype=np.float64) if not np.isscalar(nu) else float(nu)
    mue = Y / (2 * (1 + nu))
    lambdae = (Y * nu) / ((1 + nu) * (1 - 2 * nu))
    return mue, lambdae
