"""FP64 CPU oracle for the VBD path -- TEST INFRASTRUCTURE ONLY.

Python front end of ``libvbd_oracle.so`` (``vbd_oracle.c``), the checker the
MI355X product is compared against. Only ``tests/``, ``__graft_entry__.smoke``
and the CPU-baseline leg of ``bench.py`` may import this module; the product
package never does.

``OracleVbd`` restates ``pbat::sim::vbd::Data::Construct``
(sim/vbd/Data.cpp:179-244) and ``pbat::sim::vbd::Integrator``
(sim/vbd/Integrator.cpp:22-200). Arrays use the reference's shapes at the
interface (3 x nV positions, 4 x nT elements) and are stored vertex-major
(nV x 3, C order) internally, which is the memory layout of an Eigen 3 x nV
column-major matrix.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libvbd_oracle.so")

# sim/vbd/Enums.h:12-18
INIT_POSITION = 0
INIT_INERTIA = 1
INIT_KINETIC_ENERGY_MINIMUM = 2
INIT_ADAPTIVE_VBD = 3
INIT_ADAPTIVE_PBAT = 4
# graph/Enums.h
SELECT_LEAST_USED = 0
SELECT_FIRST_AVAILABLE = 1
ORDER_NATURAL = 0
ORDER_SMALLEST_DEGREE = 1
ORDER_LARGEST_DEGREE = 2

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle shared library with gcc (see oracle/Makefile)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "all", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int64)
        i64 = ctypes.c_int64
        L.ora_lame.argtypes = [ctypes.c_double, ctypes.c_double, dp, dp]
        L.ora_snh_eval.argtypes = [dp, ctypes.c_double, ctypes.c_double]
        L.ora_snh_eval.restype = ctypes.c_double
        L.ora_snh_grad.argtypes = [dp, ctypes.c_double, ctypes.c_double, dp]
        L.ora_snh_hessian.argtypes = [dp, ctypes.c_double, ctypes.c_double, dp]
        for f in ("ora_shape_function_gradients", "ora_quadrature_weights"):
            getattr(L, f).argtypes = [i64, i64, dp, ip, dp]
            getattr(L, f).restype = ctypes.c_int
        L.ora_lumped_mass.argtypes = [i64, i64, dp, ip, dp, dp]
        L.ora_lumped_mass.restype = ctypes.c_int
        L.ora_vertex_tet_adjacency.argtypes = [i64, i64, ip, ip, ip, ip]
        L.ora_greedy_color.argtypes = [i64, i64, ip, ctypes.c_int, ctypes.c_int, ip]
        L.ora_greedy_color.restype = i64
        L.ora_map_to_adjacency.argtypes = [i64, ip, ip, ip]
        L.ora_map_to_adjacency.restype = i64
        L.ora_remove_vertices.argtypes = [i64, ip, ip, i64, i64, ip]
        L.ora_remove_vertices.restype = i64
        L.ora_step.argtypes = [ctypes.c_void_p, ctypes.c_double, i64, i64]
        L.ora_step.restype = ctypes.c_int
        L.ora_vertex_elastic_block.argtypes = [ctypes.c_void_p, i64, dp, dp]
        L.ora_objective.argtypes = [ctypes.c_void_p, dp, dp, ctypes.c_double]
        L.ora_objective.restype = ctypes.c_double
        L.ora_objective_gradient.argtypes = [ctypes.c_void_p, dp, dp, ctypes.c_double, dp]
        L.ora_energy.argtypes = [ctypes.c_void_p, dp]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a: np.ndarray):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


class _State(ctypes.Structure):
    _fields_ = [
        ("nV", ctypes.c_int64), ("nT", ctypes.c_int64),
        ("x", ctypes.c_void_p), ("v", ctypes.c_void_p), ("aext", ctypes.c_void_p),
        ("m", ctypes.c_void_p), ("xt", ctypes.c_void_p), ("vt", ctypes.c_void_p),
        ("xtilde", ctypes.c_void_p),
        ("E", ctypes.c_void_p), ("GP", ctypes.c_void_p), ("wg", ctypes.c_void_p),
        ("lame", ctypes.c_void_p),
        ("GVGp", ctypes.c_void_p), ("GVGe", ctypes.c_void_p), ("GVGilocal", ctypes.c_void_p),
        ("nPartitions", ctypes.c_int64), ("Pptr", ctypes.c_void_p), ("Padj", ctypes.c_void_p),
        ("strategy", ctypes.c_int), ("kD", ctypes.c_double), ("detHZero", ctypes.c_double),
        ("nthreads", ctypes.c_int), ("gpu_velocity_semantics", ctypes.c_int),
        ("chebyshev", ctypes.c_int), ("rho", ctypes.c_double), ("xkm1", ctypes.c_void_p),
        ("xkm2", ctypes.c_void_p),
        ("fc", ctypes.c_void_p), ("F", ctypes.c_void_p), ("XVA", ctypes.c_void_p), ("FA", ctypes.c_void_p),
        ("muC", ctypes.c_double), ("muF", ctypes.c_double), ("epsv", ctypes.c_double), ("xb", ctypes.c_void_p),
    ]


def chebyshev_omega(k: int, rho2: float, omega: float) -> float:
    lib().ora_chebyshev_omega.restype = ctypes.c_double
    lib().ora_chebyshev_omega.argtypes = [ctypes.c_int64, ctypes.c_double, ctypes.c_double]
    return lib().ora_chebyshev_omega(int(k), float(rho2), float(omega))


def point_triangle_sqdist(p, a, b, c) -> float:
    """Squared point-triangle distance (geometry/DistanceQueries.h:217-222)."""
    L = lib()
    L.ora_point_triangle_sqdist.restype = ctypes.c_double
    L.ora_point_triangle_sqdist.argtypes = [ctypes.POINTER(ctypes.c_double)] * 4
    args = [np.ascontiguousarray(np.asarray(q, np.float64)) for q in (p, a, b, c)]
    return L.ora_point_triangle_sqdist(*[_dp(q) for q in args])


def point_triangle_sqdists(p, tris) -> np.ndarray:
    """point_triangle_sqdist from p to each triangle of tris (n x 3 vertices x 3)."""
    L = lib()
    dp = ctypes.POINTER(ctypes.c_double)
    L.ora_point_triangle_sqdist_many.restype = None
    L.ora_point_triangle_sqdist_many.argtypes = [dp, dp, ctypes.c_int64, dp]
    q = np.ascontiguousarray(np.asarray(p, np.float64))
    t = np.ascontiguousarray(np.asarray(tris, np.float64).reshape(-1, 9))
    out = np.zeros(t.shape[0])
    L.ora_point_triangle_sqdist_many(_dp(q), _dp(t), t.shape[0], _dp(out))
    return out


def vertex_triangle_contact(xtv, xv, xtf, xf, dt, muC, muF, epsv):
    """(E, g, H) of AccumulateVertexTriangleContact (sim/vbd/Kernels.h:215-302); xtf, xf
    are 3x3 with one triangle vertex per column."""
    L = lib()
    dp = ctypes.POINTER(ctypes.c_double)
    L.ora_vertex_triangle_contact.restype = ctypes.c_double
    L.ora_vertex_triangle_contact.argtypes = [dp, dp, dp, dp] + [ctypes.c_double] * 4 + [dp, dp]
    a = [np.ascontiguousarray(np.asarray(q, np.float64)) for q in (xtv, xv)]
    m = [np.ascontiguousarray(np.asarray(q, np.float64).T) for q in (xtf, xf)]  # column k = vertex k
    g, H = np.zeros(3), np.zeros(9)
    E = L.ora_vertex_triangle_contact(_dp(a[0]), _dp(a[1]), _dp(m[0]), _dp(m[1]), float(dt), float(muC), float(muF),
                                      float(epsv), _dp(g), _dp(H))
    return E, g, H.reshape(3, 3).T.copy()


# ---- constitutive model ---------------------------------------------------

def lame(Y: float, nu: float):
    mu, la = ctypes.c_double(), ctypes.c_double()
    lib().ora_lame(Y, nu, ctypes.byref(mu), ctypes.byref(la))
    return mu.value, la.value


def snh_eval(F: np.ndarray, mu: float, lam: float) -> float:
    """F is 3x3; flattened column-major as the reference does."""
    f = np.ascontiguousarray(np.asarray(F, dtype=np.float64).T.reshape(9))
    return lib().ora_snh_eval(_dp(f), mu, lam)


def snh_grad(F, mu, lam) -> np.ndarray:
    f = np.ascontiguousarray(np.asarray(F, dtype=np.float64).T.reshape(9))
    g = np.zeros(9)
    lib().ora_snh_grad(_dp(f), mu, lam, _dp(g))
    return g


def snh_hessian(F, mu, lam) -> np.ndarray:
    f = np.ascontiguousarray(np.asarray(F, dtype=np.float64).T.reshape(9))
    H = np.zeros(81)
    lib().ora_snh_hessian(_dp(f), mu, lam, _dp(H))
    return H.reshape(9, 9).T.copy()


# ---- precompute pieces ----------------------------------------------------

class SingularJacobian(ValueError):
    pass


def _prep(X, E):
    Xv = np.ascontiguousarray(np.asarray(X, dtype=np.float64).T)  # nV x 3
    Ev = np.ascontiguousarray(np.asarray(E, dtype=np.int64).T)    # nT x 4
    return Xv, Ev


def shape_function_gradients(X, E) -> np.ndarray:
    """Returns GP as the reference's 4 x 3nT matrix."""
    Xv, Ev = _prep(X, E)
    nV, nT = Xv.shape[0], Ev.shape[0]
    GP = np.zeros(12 * nT)
    if lib().ora_shape_function_gradients(nV, nT, _dp(Xv), _ip(Ev), _dp(GP)) != 0:
        raise SingularJacobian("Inverted or singular jacobian")
    return GP.reshape(3 * nT, 4).T.copy()


def quadrature_weights(X, E) -> np.ndarray:
    Xv, Ev = _prep(X, E)
    wg = np.zeros(Ev.shape[0])
    if lib().ora_quadrature_weights(Xv.shape[0], Ev.shape[0], _dp(Xv), _ip(Ev), _dp(wg)) != 0:
        raise SingularJacobian("Inverted or singular jacobian")
    return wg


def lumped_mass(X, E, rhoe) -> np.ndarray:
    Xv, Ev = _prep(X, E)
    rho = np.ascontiguousarray(np.asarray(rhoe, dtype=np.float64))
    m = np.zeros(Xv.shape[0])
    if lib().ora_lumped_mass(Xv.shape[0], Ev.shape[0], _dp(Xv), _ip(Ev), _dp(rho), _dp(m)) != 0:
        raise SingularJacobian("Inverted or singular jacobian")
    return m


def vertex_tet_adjacency(E, nV):
    Ev = np.ascontiguousarray(np.asarray(E, dtype=np.int64).T)
    nT = Ev.shape[0]
    p = np.zeros(nV + 1, dtype=np.int64)
    e = np.zeros(4 * nT, dtype=np.int64)
    il = np.zeros(4 * nT, dtype=np.int64)
    lib().ora_vertex_tet_adjacency(nV, nT, _ip(Ev), _ip(p), _ip(e), _ip(il))
    return p, e, il


def greedy_color(E, nV, ordering=ORDER_LARGEST_DEGREE, selection=SELECT_LEAST_USED):
    Ev = np.ascontiguousarray(np.asarray(E, dtype=np.int64).T)
    C = np.zeros(nV, dtype=np.int64)
    nc = lib().ora_greedy_color(nV, Ev.shape[0], _ip(Ev), int(ordering), int(selection), _ip(C))
    if nc < 0:
        raise RuntimeError(f"greedy colouring failed ({nc})")
    return C


def map_to_adjacency(colors):
    c = np.ascontiguousarray(np.asarray(colors, dtype=np.int64))
    nc = int(c.max()) + 1 if c.size else 0
    Pptr = np.zeros(nc + 1, dtype=np.int64)
    Padj = np.zeros(c.size, dtype=np.int64)
    lib().ora_map_to_adjacency(c.size, _ip(c), _ip(Pptr), _ip(Padj))
    return Pptr, Padj


def remove_vertices(Pptr, Padj, nV, remove):
    Pptr = np.ascontiguousarray(Pptr, dtype=np.int64).copy()
    Padj = np.ascontiguousarray(Padj, dtype=np.int64).copy()
    r = np.ascontiguousarray(np.asarray(remove, dtype=np.int64))
    n = lib().ora_remove_vertices(Pptr.size - 1, _ip(Pptr), _ip(Padj), nV, r.size, _ip(r))
    return Pptr, Padj[:n].copy()


# ---- the integrator ---------------------------------------------------------

class OracleVbd:
    """``Data().with_*(...).construct()`` followed by ``Integrator(data)``.

    Restates sim/vbd/Data.cpp:179-244 (defaults: aext = (0,0,-9.81),
    Y = 1e6 / nu = 0.45 Lame, rho = 1e3, AdaptivePbat initialisation,
    detHZero = 1e-7, kD = 0; sim/vbd/Data.h:205-246) and
    sim/vbd/Integrator.cpp:22-200.
    """

    def __init__(self, X, E, *, v=None, aext=None, rhoe=None, mue=None, lambdae=None,
                 dbc=None, ordering=ORDER_LARGEST_DEGREE, selection=SELECT_LEAST_USED,
                 strategy=INIT_ADAPTIVE_PBAT, kD=0.0, detHZero=1e-7, nthreads=1,
                 gpu_velocity_semantics=False, colors=None, chebyshev_rho=None, contact=None):
        Xv, Ev = _prep(X, E)
        nV, nT = Xv.shape[0], Ev.shape[0]
        self.nV, self.nT = nV, nT
        self.X = Xv.copy()
        self.E = Ev
        self.x = Xv.copy()
        self.xt = Xv.copy()
        self.v = np.zeros((nV, 3)) if v is None else np.ascontiguousarray(np.asarray(v, np.float64).T).copy()
        if aext is None:
            self.aext = np.tile(np.array([0.0, 0.0, -9.81]), (nV, 1))
        else:
            self.aext = np.ascontiguousarray(np.asarray(aext, np.float64).T).copy()
        self.xtilde = np.zeros((nV, 3))
        self.vt = np.zeros((nV, 3))
        if mue is None or lambdae is None:
            mu, la = lame(1e6, 0.45)
            mue, lambdae = np.full(nT, mu), np.full(nT, la)
        self.lame = np.ascontiguousarray(np.stack([np.asarray(mue, np.float64),
                                                   np.asarray(lambdae, np.float64)], axis=1))
        self.rhoe = np.full(nT, 1e3) if rhoe is None else np.asarray(rhoe, np.float64).copy()
        self.m = lumped_mass(X, E, self.rhoe)
        self.GP = np.ascontiguousarray(shape_function_gradients(X, E).T.reshape(-1))  # 12 per tet
        self.wg = quadrature_weights(X, E)
        self.GVGp, self.GVGe, self.GVGilocal = vertex_tet_adjacency(E, nV)
        if colors is None:
            colors = greedy_color(E, nV, ordering, selection)
        self.colors = np.asarray(colors, dtype=np.int64)
        Pptr, Padj = map_to_adjacency(self.colors)
        self.dbc = np.sort(np.asarray([] if dbc is None else dbc, dtype=np.int64))
        self.v[self.dbc] = 0.0
        self.aext[self.dbc] = 0.0
        self.Pptr, self.Padj = remove_vertices(Pptr, Padj, nV, self.dbc)
        self.strategy = int(strategy)
        self.kD = float(kD)
        self.detHZero = float(detHZero)
        self.nthreads = int(nthreads)
        self.gpu_velocity_semantics = bool(gpu_velocity_semantics)
        # ChebyshevIntegrator (sim/vbd/ChebyshevIntegrator.cpp:10-13): previous iterates start at x
        self.chebyshev = chebyshev_rho is not None
        self.rho = 1.0 if chebyshev_rho is None else float(chebyshev_rho)
        self.xkm1 = self.x.copy()
        self.xkm2 = self.x.copy()
        self.contact = contact is not None
        if self.contact:
            self._init_contact(**contact)

    # ---- vertex-triangle contact (gpu/impl/vbd/Integrator.cu:82-103,163-301;
    # gpu/impl/contact/VertexTriangleMixedCcdDcd.cu) --------------------------------
    def _init_contact(self, V, F, B, XVA, FA, muC=1e6, muF=0.3, epsv=1e-3, frequency=1):
        """Collision vertices V (into X), triangles F (3 x nF), body of every vertex B,
        vertex areas XVA (nV), triangle areas FA (nF) (sim/vbd/Data.cpp:34-60), the contact
        parameters (sim/vbd/Data.h:225-229). The active set starts empty, the search
        radii at FLT_MAX and the contact lists at -1
        (gpu/impl/contact/VertexTriangleMixedCcdDcd.cu:22-48, gpu/impl/vbd/Integrator.cu:60)."""
        self.cV = np.asarray(V, dtype=np.int64).reshape(-1)
        self.cF = np.ascontiguousarray(np.asarray(F, dtype=np.int64).T)  # nF x 3
        self.B = np.asarray(B, dtype=np.int64).reshape(-1)
        self.XVA = np.ascontiguousarray(np.asarray(XVA, np.float64))
        self.FA = np.ascontiguousarray(np.asarray(FA, np.float64))
        self.muC, self.muF, self.epsv = float(muC), float(muF), float(epsv)
        self.frequency = int(frequency)
        self.fc = np.full((self.nV, 8), -1, dtype=np.int64)
        self.cactive = np.zeros(self.cV.size, dtype=bool)
        self.in_av = np.zeros(self.cV.size, dtype=bool)
        self.dupper = np.full(self.cV.size, float(np.finfo(np.float32).max))
        self.xb = np.zeros((self.nV, 3))

    def initialize_active_set(self, x0, x1):
        """VertexTriangleMixedCcdDcd::InitializeActiveSet (:53-142): a collision vertex
        becomes active when the box of its segment x0 -> x1 overlaps the box of a swept
        triangle of another body (boxes closed); its search radius becomes the squared
        diagonal of the two boxes' union, the largest over such triangles (the reference
        keeps the last one its BVH traversal meets). Vertices stay active from before."""
        p0, p1 = x0[self.cV], x1[self.cV]
        lo_v, hi_v = np.minimum(p0, p1), np.maximum(p0, p1)
        tri = np.concatenate([x0[self.cF], x1[self.cF]], axis=1)  # nF x 6 x 3
        lo_f, hi_f = tri.min(axis=1), tri.max(axis=1)
        other = self.B[self.cV][:, None] != self.B[self.cF[:, 0]][None, :]
        overlap = other & np.all((lo_v[:, None, :] <= hi_f[None]) & (lo_f[None] <= hi_v[:, None, :]), axis=2)
        diag = np.maximum(hi_v[:, None, :], hi_f[None]) - np.minimum(lo_v[:, None, :], lo_f[None])
        d2 = np.where(overlap, (diag * diag).sum(axis=2), -np.inf).max(axis=1)
        hit = overlap.any(axis=1)
        self.dupper[hit] = d2[hit]
        self.cactive |= hit
        self.in_av = self.cactive.copy()

    def nearest_triangles(self, x, v):
        """Bvh::NearestNeighbours (gpu/impl/geometry/Bvh.cuh:355-466) as used by
        VertexTriangleMixedCcdDcd (:117-164) for collision vertex v: squared distances to
        the triangles of other bodies, d* the smallest, eps = FLT_EPSILON. With
        d* < r - eps (r the vertex's search radius) the triangles within d* + eps, with
        d* within r +- eps those within r + eps, else none; in triangle order, at most 8
        (the reference's order and tie-breaking follow its BVH traversal)."""
        i = self.cV[v]
        fs = np.nonzero(self.B[self.cF[:, 0]] != self.B[i])[0]
        if fs.size == 0:
            return []
        d = point_triangle_sqdists(x[i], x[self.cF[fs]])
        eps = float(np.finfo(np.float32).eps)
        ds, r = d.min(), self.dupper[v]
        if ds < r - eps:
            thr = ds + eps
        elif ds <= r + eps:
            thr = r + eps
        else:
            return []
        return [int(f) for f in fs[d <= thr][:8]]

    def update_active_set(self, x):
        """Integrator::UpdateActiveSet (gpu/impl/vbd/Integrator.cu:250-271): each vertex
        of this step's active list gets its nearest triangles as contacts; the lists of
        other vertices are left as they are."""
        for v in np.nonzero(self.in_av)[0]:
            S = self.nearest_triangles(x, v)
            self.fc[self.cV[v]] = -1
            self.fc[self.cV[v], :len(S)] = S

    def finalize_active_set(self, x):
        """VertexTriangleMixedCcdDcd::FinalizeActiveSet (:185-215): a vertex of this step's
        active list stays active when it lies below the plane of the last of its nearest
        triangles (signed distance < 0); vertices without any keep their flag."""
        for v in np.nonzero(self.in_av)[0]:
            S = self.nearest_triangles(x, v)
            if S:
                a, b, c = x[self.cF[S[-1]]]
                self.cactive[v] = float(np.dot(x[self.cV[v]] - a, np.cross(b - a, c - a))) < 0.0

    def step_contact(self, dt: float, iterations: int, substeps: int = 1) -> None:
        """Integrator::Step with contact (gpu/impl/vbd/Integrator.cu:82-103): the active
        set from the full step's sweep x_t -> x_t + dt v + dt^2 a, contacts refreshed every
        `frequency` substeps at the substep's initial iterate, VBD sweeps with contact
        terms (each colour reading positions from before its pass), the active set
        finalised at the end."""
        L = lib()
        for f in ("ora_substep_begin", "ora_sweep", "ora_substep_end"):
            getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_double]
            getattr(L, f).restype = None
        sdt = float(dt) / substeps
        self.initialize_active_set(self.x.copy(), self.x + dt * self.v + dt * dt * self.aext)
        for k in range(substeps):
            s = self._state()
            L.ora_substep_begin(ctypes.byref(s), sdt)
            if k % self.frequency == 0:
                self.update_active_set(self.x)
            for _ in range(iterations):
                L.ora_sweep(ctypes.byref(s), sdt)
            L.ora_substep_end(ctypes.byref(s), sdt)
        self.finalize_active_set(self.x)

    def _state(self) -> _State:
        s = _State()
        s.nV, s.nT = self.nV, self.nT
        for name in ("x", "v", "aext", "m", "xt", "vt", "xtilde", "GP", "wg", "lame"):
            a = getattr(self, name)
            assert a.dtype == np.float64 and a.flags.c_contiguous, name
            setattr(s, name, a.ctypes.data)
        for name in ("E", "GVGp", "GVGe", "GVGilocal", "Pptr", "Padj"):
            a = getattr(self, name)
            assert a.dtype == np.int64 and a.flags.c_contiguous, name
            setattr(s, name, a.ctypes.data)
        s.nPartitions = self.Pptr.size - 1
        s.strategy, s.kD, s.detHZero = self.strategy, self.kD, self.detHZero
        s.nthreads = self.nthreads
        s.gpu_velocity_semantics = int(self.gpu_velocity_semantics)
        s.chebyshev, s.rho = int(self.chebyshev), self.rho
        s.xkm1, s.xkm2 = self.xkm1.ctypes.data, self.xkm2.ctypes.data
        if self.contact:
            for name in ("fc", "XVA", "FA", "xb"):
                assert getattr(self, name).flags.c_contiguous, name
            s.fc, s.F = self.fc.ctypes.data, self.cF.ctypes.data
            s.XVA, s.FA, s.xb = self.XVA.ctypes.data, self.FA.ctypes.data, self.xb.ctypes.data
            s.muC, s.muF, s.epsv = self.muC, self.muF, self.epsv
        return s

    # sim/vbd/Integrator.cpp:22-45
    def step(self, dt: float, iterations: int, substeps: int = 1) -> None:
        s = self._state()
        rc = lib().ora_step(ctypes.byref(s), float(dt), int(iterations), int(substeps))
        if rc != 0:
            raise ValueError(f"oracle step failed ({rc})")

    def step_anderson(self, dt: float, iterations: int, substeps: int = 1, window: int = 5,
                      rcond: float = 1e-10) -> None:
        """AndersonIntegrator::Solve (sim/vbd/AndersonIntegrator.cpp:24-56) around the
        same substep pieces as step(): after each VBD iteration k >= 1, with
        G_k = x, F_k = G_k - x_{k-1} and the last min(k, m) differences DG, DF,
        alpha = argmin |DF alpha - F_k| and x = G_k - DG alpha. The least-squares
        solve is numpy's SVD-based lstsq with singular values below rcond x the
        largest dropped (the reference uses Eigen's complete orthogonal
        decomposition with threshold 1e-10)."""
        L = lib()
        for f in ("ora_substep_begin", "ora_sweep", "ora_substep_end"):
            getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_double]
            getattr(L, f).restype = None
        sdt = float(dt) / substeps
        m = int(window)
        n = 3 * self.nV
        for _ in range(substeps):
            s = self._state()
            L.ora_substep_begin(ctypes.byref(s), sdt)
            DGK = np.zeros((n, m))
            DFK = np.zeros((n, m))
            xkm1 = self.x.reshape(-1).copy()
            if iterations > 0:
                L.ora_sweep(ctypes.byref(s), sdt)
                Gkm1 = self.x.reshape(-1).copy()
                Fkm1 = Gkm1 - xkm1
            for k in range(1, iterations):
                xkm1 = self.x.reshape(-1).copy()
                L.ora_sweep(ctypes.byref(s), sdt)
                Gk = self.x.reshape(-1).copy()
                Fk = Gk - xkm1
                col = (k - 1) % m
                DGK[:, col] = Gk - Gkm1
                DFK[:, col] = Fk - Fkm1
                Gkm1, Fkm1 = Gk, Fk
                mk = min(m, k)
                alpha = np.linalg.lstsq(DFK[:, :mk], Fk, rcond=rcond)[0]
                self.x.reshape(-1)[:] = Gk - DGK[:, :mk] @ alpha
            L.ora_substep_end(ctypes.byref(s), sdt)

    def step_trust_region(self, dt: float, iterations: int, substeps: int = 1, eta: float = 0.2,
                          tau: float = 2.0, curved: bool = True, log=None, forced=None) -> None:
        """TrustRegionIntegrator::Solve (gpu/impl/vbd/TrustRegionIntegrator.cu:36-46) in FP64
        around the same substep pieces as step(); f is ObjectiveFunction (:265-364, no
        contact term: single-body scenes have no active contacts).

        Linear path (:48-165): iterations 0 and 1 are plain VBD iterations; from k = 2 a
        quadratic model m(t) through the last three (t, f) pairs, time shifted so the
        current iterate sits at t = 0 (ConstructModel, :452-468), is minimised
        (ModelOptimalStep, :400-430) over the line x(t) = x_k + t (x_VBD - x_k) with t
        clamped to [1, R / |x_VBD - x_k|]; the step (:491-506) is kept when
        (f_k - f(x(t))) / (f_k - m(t)) > eta, else rolled back (:508-522) to x_VBD.
        Curved path (:167-263): its constraint solve returns 0 (:700-703), so every
        clamped step sits at the lower bound and each iteration is the fallback VBD
        iteration (:240-251); positions are those of plain VBD.

        log, if a list, receives one (k, t, accepted, ratio) tuple per iteration k >= 2
        (ratio: the actual over the model's decrease, None when no step was tried).
        forced, if given, is a sequence of (t, accepted) pairs, one per iteration
        k >= 2 of each substep, that replaces the path's own choices (the log still
        records the oracle's own t and decision at that state)."""
        L = lib()
        for f in ("ora_substep_begin", "ora_sweep", "ora_substep_end"):
            getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_double]
            getattr(L, f).restype = None
        sdt = float(dt) / substeps
        flt_min = float(np.finfo(np.float32).tiny)
        flt_max = float(np.finfo(np.float32).max)
        small = float(np.float32(1e-6))  # the linear path's 1e-6f tolerance

        def objective():
            return self.objective(self.x.T, self.xtilde.T, sdt)

        n_forced = 0
        for _ in range(substeps):
            s = self._state()
            L.ora_substep_begin(ctypes.byref(s), sdt)
            fk, tk = objective(), -1.0
            fkm1 = fkm2 = tkm1 = tkm2 = 0.0
            R2 = 0.0 if not curved else float(np.finfo(np.float64).eps)
            for k in range(iterations):
                if k < 2 or curved:
                    if k >= 2:
                        # curved: SolveCurvedTrustRegionConstraint returns 0 and
                        # clamp(t*, FLT_MIN, 0) is at the lower bound either way
                        R2 = max(R2 / (tau * tau), float(np.finfo(np.float64).eps))
                        if log is not None:
                            log.append((k, 0.0, False, None))
                    fkm2, fkm1, tkm2, tkm1 = fkm1, fk, tkm1, tk
                    xkm1 = self.x.copy()
                    L.ora_sweep(ctypes.byref(s), sdt)
                    fk = objective()
                    tk = (tkm1 + 1.0) if k >= 2 else float(k)
                    continue
                # ConstructModel (linear): shift time so t_k = 0, fit a t^2 + b t + c
                tkm2, tkm1, tk = tkm2 - tk, tkm1 - tk, 0.0
                Q = np.array([[tkm2 * tkm2, tkm2, 1.0], [tkm1 * tkm1, tkm1, 1.0], [0.0, 0.0, 1.0]])
                aQ = np.linalg.solve(Q, np.array([fkm2, fkm1, fk]))
                # UpdateIterates
                fkm2, fkm1, tkm2, tkm1 = fkm1, fk, tkm1, tk
                xkm1 = self.x.copy()
                L.ora_sweep(ctypes.byref(s), sdt)
                dx = self.x - xkm1
                dx2 = float((dx * dx).sum())
                if R2 < dx2 + small:
                    R2 = tau * tau * dx2
                lower = 1.0
                upper = np.sqrt(R2 / dx2) if dx2 > 0 else float("nan")
                # ModelOptimalStep
                if abs(aQ[0]) > flt_min:
                    tstar = -aQ[1] / (2.0 * aQ[0]) if aQ[0] > flt_min else flt_max
                elif abs(aQ[1]) > flt_min:
                    tstar = -flt_max if aQ[1] > 0 else flt_max
                else:
                    tstar = 0.0
                t = lower if tstar < lower else (upper if upper < tstar else tstar)
                x_vbd = self.x.copy()

                def try_step(tt):
                    # TakeLinearStep, then the actual over the model's decrease
                    self.x[:] = xkm1 + tt * dx
                    f_try = objective()
                    m_try = aQ[0] * tt * tt + aQ[1] * tt + aQ[2]
                    return f_try, float((fkm1 - f_try) / (fkm1 - m_try))

                attempt = not (abs(t - lower) < small)
                accepted, ratio = False, None
                if attempt:
                    fk, ratio = try_step(t)
                    accepted = ratio > eta
                own = (t, accepted, ratio)
                if forced is not None:
                    t, accepted = forced[n_forced]
                    accepted = bool(accepted)
                    n_forced += 1
                    self.x[:] = x_vbd
                    attempt = not (abs(t - lower) < small)
                    if attempt:
                        fk, _ = try_step(t)
                if accepted:
                    if (upper - t) < small:
                        R2 *= tau * tau
                    tk = t
                else:
                    R2 /= tau * tau
                    if attempt:
                        self.x[:] = xkm1 + (self.x - xkm1) / t
                    fk = objective()
                    tk = tkm1 + 1.0
                if log is not None:
                    log.append((k, float(own[0]), bool(own[1]), own[2]))
            L.ora_substep_end(ctypes.byref(s), sdt)

    def objective(self, x, xtilde, dt) -> float:
        s = self._state()
        xv = np.ascontiguousarray(np.asarray(x, np.float64).T)
        xl = np.ascontiguousarray(np.asarray(xtilde, np.float64).T)
        return lib().ora_objective(ctypes.byref(s), _dp(xv), _dp(xl), float(dt))

    def objective_gradient(self, x, xtilde, dt) -> np.ndarray:
        """Gradient in the reference's reshaped (3*nV,) vertex-major order."""
        s = self._state()
        xv = np.ascontiguousarray(np.asarray(x, np.float64).T)
        xl = np.ascontiguousarray(np.asarray(xtilde, np.float64).T)
        g = np.zeros(3 * self.nV)
        lib().ora_objective_gradient(ctypes.byref(s), _dp(xv), _dp(xl), float(dt), _dp(g))
        return g

    def vertex_elastic_block(self, i: int):
        s = self._state()
        H = np.zeros(9)
        g = np.zeros(3)
        lib().ora_vertex_elastic_block(ctypes.byref(s), int(i), _dp(H), _dp(g))
        return H.reshape(3, 3).T.copy(), g

    def energy(self) -> np.ndarray:
        """(kinetic, elastic, external potential) in FP64."""
        s = self._state()
        out = np.zeros(3)
        lib().ora_energy(ctypes.byref(s), _dp(out))
        return out

    # reference-shaped views (3 x nV)
    @property
    def positions(self) -> np.ndarray:
        return self.x.T

    @property
    def velocities(self) -> np.ndarray:
        return self.v.T
