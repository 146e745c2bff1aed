"""Synthetic tetrahedral meshes for the VBD workloads.

Cube grids split into 6 positively oriented tetrahedra per cell (the Kuhn
split along each cell's main diagonal), optional seeded jitter of interior
vertices, per-tet material draws and bottom-face Dirichlet sets. Arrays use
the reference's shapes: positions ``3 x nV`` (float64) and elements ``4 x nT``
(int64), as ``pbat.sim.vbd.Data.with_volume_mesh`` takes them
(bindings/pypbat/sim/vbd/Data.cpp:33-45).

All randomness comes from ``numpy.random.default_rng(seed)`` so every caller
(tests, the benchmark, the oracle) rebuilds bit-identical inputs.
"""
from __future__ import annotations

import itertools

import numpy as np

# Six axis orderings; the tet of ordering (a, b, c) walks from the cell's
# min corner to its max corner along +e_a, +e_b, +e_c.
_KUHN_PATHS = list(itertools.permutations(range(3)))


def _perm_sign(p) -> int:
    s = 1
    p = list(p)
    for i in range(len(p)):
        for j in range(i + 1, len(p)):
            if p[i] > p[j]:
                s = -s
    return s


def cube_tet_grid(nx: int, ny: int | None = None, nz: int | None = None, *, size=1.0,
                  jitter: float = 0.0, seed: int = 0):
    """Unit-cube grid of ``nx x ny x nz`` cells, 6 tets per cell.

    Vertex ``(i, j, k)`` (x, y, z cell coordinates) has index
    ``i + (nx+1) * (j + (ny+1) * k)`` and sits at ``(i/nx, j/ny, k/nz) * size``.
    Interior vertices move by ``U(-jitter, jitter)`` cell widths per axis.
    Returns ``X (3, nV) float64`` and ``E (4, nT) int64``; every tet has
    positive signed volume.
    """
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    sx, sy, sz = nx + 1, ny + 1, nz + 1
    i, j, k = np.meshgrid(np.arange(sx), np.arange(sy), np.arange(sz), indexing="ij")
    # flatten with x fastest
    I = i.transpose(2, 1, 0).reshape(-1)
    J = j.transpose(2, 1, 0).reshape(-1)
    K = k.transpose(2, 1, 0).reshape(-1)
    h = np.array([size / nx, size / ny, size / nz])
    X = np.stack([I * h[0], J * h[1], K * h[2]]).astype(np.float64)
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter, jitter, size=X.shape) * h[:, None]
        interior = (I > 0) & (I < nx) & (J > 0) & (J < ny) & (K > 0) & (K < nz)
        X[:, interior] += d[:, interior]

    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci = ci.transpose(2, 1, 0).reshape(-1)
    cj = cj.transpose(2, 1, 0).reshape(-1)
    ck = ck.transpose(2, 1, 0).reshape(-1)

    def vid(di, dj, dk):
        return (ci + di) + sx * ((cj + dj) + sy * (ck + dk))

    tets = []
    for p in _KUHN_PATHS:
        corner = [0, 0, 0]
        verts = [vid(*corner)]
        for axis in p:
            corner[axis] = 1
            verts.append(vid(*corner))
        if _perm_sign(p) < 0:
            verts[1], verts[2] = verts[2], verts[1]
        tets.append(np.stack(verts))
    # tets of one cell are consecutive: element index = 6 * cell + path
    E = np.stack(tets, axis=1).reshape(4, -1).astype(np.int64)
    return X, E


def signed_volumes(X: np.ndarray, E: np.ndarray) -> np.ndarray:
    x0 = X[:, E[0]]
    D = np.stack([X[:, E[1]] - x0, X[:, E[2]] - x0, X[:, E[3]] - x0], axis=1)  # (3, 3, nT)
    return np.linalg.det(np.moveaxis(D, 2, 0)) / 6.0


def boundary_triangles(E: np.ndarray) -> np.ndarray:
    """Boundary faces of a tet mesh with positively oriented tets, ``3 x nF``, each
    ordered so that cross(b - a, c - a) points out of its tet."""
    a, b, c, d = E
    faces = np.concatenate([np.stack([a, c, b]), np.stack([a, b, d]), np.stack([a, d, c]), np.stack([b, c, d])],
                           axis=1)
    key = np.sort(faces, axis=0)
    _, inv, counts = np.unique(key, axis=1, return_inverse=True, return_counts=True)
    return np.ascontiguousarray(faces[:, counts[inv.reshape(-1)] == 1])


def bottom_vertices(X: np.ndarray, axis: int = 1) -> np.ndarray:
    """Indices of the vertices on the minimum-coordinate face along ``axis``."""
    lo = X[axis].min()
    return np.flatnonzero(X[axis] <= lo + 1e-12).astype(np.int64)


def log_uniform(rng, lo: float, hi: float, n: int) -> np.ndarray:
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=n))


def lame_coefficients(Y, nu):
    """mu = Y / (2 (1 + nu)), lambda = Y nu / ((1 + nu)(1 - 2 nu))
    (physics/HyperElasticity.cpp:6-11; python/pbatoolkit/py/fem.py:6-17)."""
    Y = np.asarray(Y, dtype=np.float64)
    nu = np.asarray(nu, dtype=np.float64)
    mu = Y / (2.0 * (1.0 + nu))
    lam = (Y * nu) / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return mu, lam


class Scene:
    """One VBD workload: mesh, material, loading and solver settings."""

    def __init__(self, X, E, mue, lambdae, rhoe, dbc, aext, dt, iterations, steps, name=""):
        self.X, self.E = X, E
        self.mue, self.lambdae, self.rhoe = mue, lambdae, rhoe
        self.dbc = dbc
        self.aext = aext
        self.dt, self.iterations, self.steps = dt, iterations, steps
        self.name = name

    @property
    def n_vertices(self) -> int:
        return self.X.shape[1]

    @property
    def n_tets(self) -> int:
        return self.E.shape[1]


def gravity(nV: int, g: float = 9.81) -> np.ndarray:
    a = np.zeros((3, nV))
    a[1] = -g
    return a


def plain_cpu_scene(n: int = 10) -> Scene:
    """10^3 grid, 6 tets/cell, no jitter; SNH Y=1e6, nu=0.45, rho=1000; gravity -y,
    bottom face fixed; h=1/60, 10 iterations, 100 steps."""
    X, E = cube_tet_grid(n)
    mu, lam = lame_coefficients(1e6, 0.45)
    nT = E.shape[1]
    return Scene(X, E, np.full(nT, mu), np.full(nT, lam), np.full(nT, 1e3), bottom_vertices(X),
                 gravity(X.shape[1]), 1.0 / 60.0, 10, 100, name=f"plain_{n}^3")


def plain_gpu_scene(n: int = 64, seed: int = 0) -> Scene:
    """n^3 grid (64^3: 274,625 verts, 1,572,864 tets), interior jitter +-0.1 cell,
    per-tet Y log-uniform in [1e5, 1e7], nu=0.45, rho=1000; gravity -y, bottom
    fixed; h=1/60, 20 iterations, 50 steps."""
    X, E = cube_tet_grid(n, jitter=0.1, seed=seed)
    rng = np.random.default_rng(seed + 1)
    nT = E.shape[1]
    Y = log_uniform(rng, 1e5, 1e7, nT)
    mu, lam = lame_coefficients(Y, 0.45)
    return Scene(X, E, mu, lam, np.full(nT, 1e3), bottom_vertices(X), gravity(X.shape[1]),
                 1.0 / 60.0, 20, 50, name=f"plain_jittered_{n}^3")


def batched_scenes(count: int = 256, lo: int = 4, hi: int = 24, seed: int = 0, *,
                   per_tet_log_y: bool = False, dt: float = 1.0 / 60.0, iterations: int = 10,
                   steps: int = 100):
    """``count`` jittered cube meshes with grid dims drawn per axis from U{lo..hi};
    per-mesh Y log-uniform in [1e5, 1e7] and nu uniform in [0.3, 0.48] (or per-tet
    log Y when ``per_tet_log_y``), rho=1000, gravity -y, bottom faces fixed."""
    rng = np.random.default_rng(seed)
    scenes = []
    for b in range(count):
        dims = rng.integers(lo, hi + 1, size=3)
        X, E = cube_tet_grid(int(dims[0]), int(dims[1]), int(dims[2]), jitter=0.1,
                             seed=seed * 100003 + b)
        nT = E.shape[1]
        if per_tet_log_y:
            Y = log_uniform(rng, 1e5, 1e7, nT)
            nu = np.full(nT, 0.45)
        else:
            Y = np.full(nT, log_uniform(rng, 1e5, 1e7, 1)[0])
            nu = np.full(nT, rng.uniform(0.3, 0.48))
        mu, lam = lame_coefficients(Y, nu)
        scenes.append(Scene(X, E, mu, lam, np.full(nT, 1e3), bottom_vertices(X),
                            gravity(X.shape[1]), dt, iterations, steps, name=f"batch{b}"))
    return scenes


class DiffScene(Scene):
    """A differentiable-rollout workload: a Scene plus hidden per-tet log Y (the
    parameters the target rollout used), the initial guess of log Y and v0."""

    def __init__(self, scene: Scene, log_y_hidden, log_y_guess, v0):
        super().__init__(scene.X, scene.E, scene.mue, scene.lambdae, scene.rhoe, scene.dbc, scene.aext,
                         scene.dt, scene.iterations, scene.steps, scene.name)
        self.log_y_hidden = log_y_hidden
        self.log_y_guess = log_y_guess
        self.v0 = v0


def _random_unit_vectors(rng, n: int) -> np.ndarray:
    u = rng.normal(size=(3, n))
    return u / np.linalg.norm(u, axis=0, keepdims=True)


def differentiable_scene(n: int = 32, seed: int = 0, *, nx=None, ny=None, nz=None, dt: float = 1.0 / 60.0,
                         iterations: int = 20, steps: int = 20) -> DiffScene:
    """n^3 jittered grid, hidden per-tet log Y log-uniform in [1e5, 1e7] (nu = 0.45,
    rho = 1000), bottom face fixed, gravity; v0 = 0.5 x a random unit vector per
    vertex (0 on Dirichlet vertices); the material the scene is built with is the
    guess Y = 1e6 everywhere. 32^3: 35,937 vertices, 196,608 tets."""
    nx = n if nx is None else nx
    X, E = cube_tet_grid(nx, ny, nz, jitter=0.1, seed=seed)
    rng = np.random.default_rng(seed + 7919)
    nT = E.shape[1]
    log_y_hidden = np.log(log_uniform(rng, 1e5, 1e7, nT))
    mu, lam = lame_coefficients(np.full(nT, 1e6), 0.45)
    dbc = bottom_vertices(X)
    v0 = 0.5 * _random_unit_vectors(rng, X.shape[1])
    v0[:, dbc] = 0.0
    s = Scene(X, E, mu, lam, np.full(nT, 1e3), dbc, gravity(X.shape[1]), dt, iterations, steps,
              name=f"diff_{nx}x{ny or nx}x{nz or nx}")
    return DiffScene(s, log_y_hidden, np.full(nT, np.log(1e6)), v0)


def batched_differentiable_scenes(count: int = 64, lo: int = 8, hi: int = 20, seed: int = 0, *,
                                  dt: float = 1.0 / 120.0, iterations: int = 10, steps: int = 200):
    """``count`` differentiable scenes with grid dims U{lo..hi} per axis (h = 1/120,
    10 iterations, 200 steps by default)."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(count):
        d = rng.integers(lo, hi + 1, size=3)
        out.append(differentiable_scene(int(d[0]), seed=seed * 100003 + b, nx=int(d[0]), ny=int(d[1]),
                                        nz=int(d[2]), dt=dt, iterations=iterations, steps=steps))
    return out
