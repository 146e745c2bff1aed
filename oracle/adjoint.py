"""FP64 CPU adjoint of a VBD rollout -- TEST INFRASTRUCTURE ONLY.

The implicit-function-theorem gradient the MI355X adjoint computes, evaluated in
FP64 with an assembled sparse Hessian and a direct solve (SuperLU from scipy; the
reference's own sparse solvers are Eigen's, which this image lacks).

Per step t, x_{t+1} = argmin E_t(x) over the free (non-Dirichlet) vertices, with
    E_t(x) = 1/2 (x - xt~)^T M (x - xt~) + h^2 sum_e w_e Psi_e(x),
    xt~ = x_t + h v_t + h^2 a,     v_{t+1} = (x_{t+1} - x_t) / h
(sim/vbd/Integrator.cpp:22-45, 138-166). In the forward's scaling
H = M / h^2 + sum_e w_e d2Psi_e/dx2 at x_{t+1}. Going backwards from the loss
gradients (xbar, vbar) of x_T, v_T:
    solve H y = xbar + vbar / h                       (free dofs)
    d L / d log Y_e -= y_e . (w_e dPsi_e/dx)          (mu, lambda proportional to Y)
    xt~bar = M y / h^2
    xbar <- xt~bar - vbar / h,   vbar <- h xt~bar     (free dofs; Dirichlet rows 0)
after which (xbar, vbar) are the gradients with respect to (x_0, v_0).
"""
from __future__ import annotations

import ctypes

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import oracle as O


def _elem(o: O.OracleVbd, x: np.ndarray, grad=True, hess=False, psi=False):
    L = O.lib()
    L.ora_element_derivatives.argtypes = [ctypes.c_void_p, O.ctypes.POINTER(ctypes.c_double)] + \
        [ctypes.c_void_p] * 3
    s = o._state()
    xv = np.ascontiguousarray(x, dtype=np.float64)  # nV x 3
    ge = np.zeros(12 * o.nT) if grad else None
    He = np.zeros(144 * o.nT) if hess else None
    ps = np.zeros(o.nT) if psi else None
    ptr = (lambda a: a.ctypes.data if a is not None else None)
    L.ora_element_derivatives(ctypes.byref(s), O._dp(xv), ptr(ge), ptr(He), ptr(ps))
    return (ge.reshape(o.nT, 4, 3) if grad else None,
            He.reshape(o.nT, 12, 12) if hess else None,
            ps)


def young_modulus(o: O.OracleVbd) -> np.ndarray:
    mu, lam = o.lame[:, 0], o.lame[:, 1]
    return mu * (3 * lam + 2 * mu) / (lam + mu)


def set_log_young_modulus(o: O.OracleVbd, log_y: np.ndarray) -> None:
    """Rescale each tet's (mu, lambda) so its Young's modulus is exp(log_y), keeping nu."""
    scale = np.exp(np.asarray(log_y, dtype=np.float64)) / young_modulus(o)
    o.lame *= scale[:, None]


def free_mask(o: O.OracleVbd) -> np.ndarray:
    m = np.ones(o.nV, dtype=bool)
    m[o.dbc] = False
    return m


def hessian(o: O.OracleVbd, x: np.ndarray, h: float) -> sp.csc_matrix:
    """M / h^2 + sum_e w_e d2Psi_e/dx2 at x, restricted to free dofs."""
    _, He, _ = _elem(o, x, grad=False, hess=True)
    E = o.E  # nT x 4
    dofs = (3 * E[:, :, None] + np.arange(3)[None, None, :]).reshape(o.nT, 12)
    rows = np.repeat(dofs, 12, axis=1).reshape(-1)
    cols = np.tile(dofs, (1, 12)).reshape(-1)
    n = 3 * o.nV
    K = sp.coo_matrix((He.reshape(-1), (rows, cols)), shape=(n, n)).tocsr()
    K = K + sp.diags(np.repeat(o.m, 3) / (h * h))
    free = np.repeat(free_mask(o), 3)
    idx = np.flatnonzero(free)
    return K[idx][:, idx].tocsc(), idx


def rollout(o: O.OracleVbd, dt: float, iterations: int, steps: int):
    """Step the oracle; return positions x_0..x_T (each nV x 3)."""
    xs = [o.x.copy()]
    for _ in range(steps):
        o.step(dt, iterations, 1)
        xs.append(o.x.copy())
    return xs


def ift_gradient(o: O.OracleVbd, xs, dt: float, dL_dxT: np.ndarray, dL_dvT: np.ndarray | None = None):
    """Gradients of L with respect to (per-tet log Y, v_0, x_0) by the IFT adjoint.

    xs: positions x_0..x_T of the rollout (nV x 3 each). dL_dxT, dL_dvT: nV x 3.
    Returns (dlogY [nT], dv0 [nV x 3], dx0 [nV x 3])."""
    h = float(dt)
    free = free_mask(o)
    xbar = np.array(dL_dxT, dtype=np.float64)
    vbar = np.zeros_like(xbar) if dL_dvT is None else np.array(dL_dvT, dtype=np.float64)
    xbar[~free] = 0.0
    vbar[~free] = 0.0
    dlogy = np.zeros(o.nT)
    for t in reversed(range(len(xs) - 1)):
        x1 = xs[t + 1]
        H, idx = hessian(o, x1, h)
        rhs = (xbar + vbar / h).reshape(-1)[idx]
        y = np.zeros(3 * o.nV)
        y[idx] = spla.splu(H).solve(rhs)
        y = y.reshape(o.nV, 3)
        ge, _, _ = _elem(o, x1)
        dlogy -= np.einsum("ekd,ekd->e", y[o.E], ge)
        xtil_bar = (o.m / (h * h))[:, None] * y
        xbar = xtil_bar - vbar / h
        vbar = h * xtil_bar
        xbar[~free] = 0.0
        vbar[~free] = 0.0
    return dlogy, vbar, xbar


def hessian_vector(o: O.OracleVbd, x: np.ndarray, h: float, p: np.ndarray) -> np.ndarray:
    """(M / h^2 + sum_e w_e d2Psi_e/dx2) p at x over all dofs (nV x 3)."""
    _, He, _ = _elem(o, x, grad=False, hess=True)
    pe = p[o.E].reshape(o.nT, 12)
    out = np.zeros((o.nV, 3))
    np.add.at(out, o.E, np.einsum("eij,ej->ei", He, pe).reshape(o.nT, 4, 3))
    return out + (o.m / (h * h))[:, None] * p
