"""``pbat.sim.vbd`` counterpart: the VBD problem description and its builder.

Mirrors ``pbat::sim::vbd::Data`` (sim/vbd/Data.h:14-247) and its Python
binding (bindings/pypbat/sim/vbd/Data.cpp:16-290): same builder method names,
argument names, defaults, fields and validation errors. ``construct`` runs the
native precompute of libpbat_vbd.so (``pbat_vbd_data_construct``), the
counterpart of ``Data::Construct`` (sim/vbd/Data.cpp:179-308).
"""
from __future__ import annotations

import ctypes
import enum

import numpy as np

from .. import _native
from ..graph import GreedyColorOrderingStrategy, GreedyColorSelectionStrategy


class InitializationStrategy(enum.IntEnum):
    """sim/vbd/Enums.h:12-18"""
    Position = 0
    Inertia = 1
    KineticEnergyMinimum = 2
    AdaptiveVbd = 3
    AdaptivePbat = 4


class AccelerationStrategy(enum.IntEnum):
    """sim/vbd/Enums.h:21-28 (Python names of bindings/pypbat/sim/vbd/Data.cpp:26-32)"""
    Base = 0
    Chebyshev = 1
    Anderson = 2
    Nesterov = 3
    Broyden = 4
    TrustRegion = 5


Position = InitializationStrategy.Position
Inertia = InitializationStrategy.Inertia
KineticEnergyMinimum = InitializationStrategy.KineticEnergyMinimum
AdaptiveVbd = InitializationStrategy.AdaptiveVbd
AdaptivePbat = InitializationStrategy.AdaptivePbat


def _f64_matrix(a, rows: int | None, name: str) -> np.ndarray:
    a = np.array(a, dtype=np.float64, copy=True)
    if a.ndim == 1 and rows is None:
        return a
    if a.ndim != 2 or (rows is not None and a.shape[0] != rows):
        raise ValueError(f"{name} must be a {rows}x|n| matrix, got shape {a.shape}")
    return a


def _i64(a, name: str) -> np.ndarray:
    a = np.array(a, copy=True)
    if a.size and not np.issubdtype(a.dtype, np.integer):
        raise ValueError(f"{name} must hold integers")
    return a.astype(np.int64)


class Data:
    """VBD problem data with a fluent builder (``with_*`` methods return self)."""

    def __init__(self):
        empty3 = np.zeros((3, 0))
        self.X = empty3.copy()
        self.E = np.zeros((4, 0), dtype=np.int64)
        self.B = np.zeros(0, dtype=np.int64)
        self.V = np.zeros(0, dtype=np.int64)
        self.F = np.zeros((3, 0), dtype=np.int64)
        self.XVA = np.zeros(0)
        self.FA = np.zeros(0)
        self.x = empty3.copy()
        self.v = empty3.copy()
        self.aext = empty3.copy()
        self.m = np.zeros(0)
        self.xt = empty3.copy()
        self.xtilde = empty3.copy()
        self.vt = empty3.copy()
        self.wg = np.zeros(0)
        self.GP = np.zeros((4, 0))
        self.rhoe = np.zeros(0)
        self.lame = np.zeros((2, 0))
        self.GVGp = np.zeros(0, dtype=np.int64)
        self.GVGe = np.zeros(0, dtype=np.int64)
        self.GVGilocal = np.zeros(0, dtype=np.int64)
        self.muD = 1.0
        self.dbc = np.zeros(0, dtype=np.int64)
        self.vertex_coloring_ordering = GreedyColorOrderingStrategy.LargestDegree
        self.vertex_coloring_selection = GreedyColorSelectionStrategy.LeastUsed
        self.colors = np.zeros(0, dtype=np.int64)
        self.Pptr = np.zeros(0, dtype=np.int64)
        self.Padj = np.zeros(0, dtype=np.int64)
        self.strategy = InitializationStrategy.AdaptivePbat
        self.kD = 0.0
        self.muC = 1e6
        self.muF = 0.3
        self.epsv = 1e-3
        self.active_set_update_frequency = 1
        self.detH_zero = 1e-7
        self.accelerator = AccelerationStrategy.Base
        self.rho = 1.0
        self.manderson = 5
        self.nesterov_L = 1.0
        self.nesterov_start = 3
        self.eta = 0.2
        self.tau = 2.0
        self.curved = True

    # ---- builder (sim/vbd/Data.cpp:21-177) ------------------------------------
    def with_volume_mesh(self, X, T):
        """Sets 3x|#nodes| positions X and 4x|#elements| tetrahedra T."""
        self.X = _f64_matrix(X, 3, "X")
        T = _i64(T, "T")
        if T.ndim != 2 or T.shape[0] != 4:
            raise ValueError(f"T must be a 4x|#elements| matrix, got shape {T.shape}")
        self.E = T
        if self.B.size == 0:
            self.B = np.ones(self.X.shape[1], dtype=np.int64)
        return self

    def with_surface_mesh(self, V, F):
        """Sets the collision mesh: vertex indices V and 3x|#triangles| F
        (contact data; plain VBD does not read it)."""
        self.V = _i64(V, "V").reshape(-1)
        self.F = _i64(F, "F")
        if self.F.ndim != 2 or self.F.shape[0] != 3:
            raise ValueError("F must be a 3x|#triangles| matrix")
        X = self.X
        AB = X[:, self.F[1]] - X[:, self.F[0]]
        AC = X[:, self.F[2]] - X[:, self.F[0]]
        dblarea = np.linalg.norm(np.cross(AB.T, AC.T), axis=1)
        self.XVA = np.zeros(X.shape[1])
        for k in range(3):
            np.add.at(self.XVA, self.F[k], dblarea / 6.0)
        self.FA = dblarea / 2.0
        return self

    def with_bodies(self, B):
        self.B = _i64(B, "B").reshape(-1)
        return self

    def with_velocity(self, v):
        self.v = _f64_matrix(v, 3, "v")
        return self

    def with_acceleration(self, a):
        self.aext = _f64_matrix(a, 3, "a")
        return self

    def with_material(self, rhoe, mue, lambdae):
        self.rhoe = np.array(rhoe, dtype=np.float64).reshape(-1)
        mue = np.array(mue, dtype=np.float64).reshape(-1)
        lambdae = np.array(lambdae, dtype=np.float64).reshape(-1)
        if mue.size != lambdae.size:
            raise ValueError("mue and lambdae must have the same size")
        self.lame = np.stack([mue, lambdae])
        return self

    def with_dirichlet_vertices(self, dbc, muD: float = 1.0, input_sorted: bool = True):
        """Dirichlet vertices never move (they leave the vertex partitions)."""
        self.dbc = _i64(dbc, "dbc").reshape(-1)
        self.muD = float(muD)
        if not input_sorted:
            self.dbc = np.sort(self.dbc)
        return self

    def with_vertex_coloring_strategy(self, ordering, selection):
        self.vertex_coloring_ordering = GreedyColorOrderingStrategy(int(ordering))
        self.vertex_coloring_selection = GreedyColorSelectionStrategy(int(selection))
        return self

    def with_initialization_strategy(self, strategy):
        self.strategy = InitializationStrategy(int(strategy))
        return self

    def with_rayleigh_damping(self, kD: float):
        self.kD = float(kD)
        return self

    def with_contact_parameters(self, muC: float, muF: float, epsv: float):
        self.muC, self.muF, self.epsv = float(muC), float(muF), float(epsv)
        return self

    def with_active_set_update_frequency(self, frequency: int):
        self.active_set_update_frequency = int(frequency)
        return self

    def with_hessian_determinant_zero(self, zero: float):
        self.detH_zero = float(zero)
        return self

    def with_chebyshev_acceleration(self, rho: float):
        self.rho = float(rho)
        self.accelerator = AccelerationStrategy.Chebyshev
        return self

    def with_anderson_acceleration(self, window_size: int):
        self.manderson = int(window_size)
        self.accelerator = AccelerationStrategy.Anderson
        return self

    def with_broyden_acceleration(self, window_size: int):
        self.manderson = int(window_size)
        self.accelerator = AccelerationStrategy.Broyden
        return self

    def with_nesterov_acceleration(self, L: float, start: int = 3):
        self.nesterov_L, self.nesterov_start = float(L), int(start)
        self.accelerator = AccelerationStrategy.Nesterov
        return self

    def with_trust_region_acceleration(self, eta: float, tau: float, curved: bool = True):
        self.eta, self.tau, self.curved = float(eta), float(tau), bool(curved)
        self.accelerator = AccelerationStrategy.TrustRegion
        return self

    # ---- construct (sim/vbd/Data.cpp:179-308) ---------------------------------
    def construct(self, validate: bool = True):
        X, E = self.X, self.E
        nV, nT = X.shape[1], E.shape[1]
        if nV == 0 or nT == 0:
            raise ValueError("with_volume_mesh must be called with a non-empty mesh before construct")
        self.x = X.copy()
        if self.xt.size == 0:
            self.xt = self.x.copy()
        if self.v.size == 0:
            self.v = np.zeros((3, nV))
        if self.aext.size == 0:
            self.aext = np.tile(np.array([[0.0], [0.0], [-9.81]]), (1, nV))
        self.xtilde = np.zeros((3, nV))
        self.vt = np.zeros((3, nV))
        L = _native.lib()
        Xc = np.ascontiguousarray(X.T)
        Ec = np.ascontiguousarray(E.T)
        have_material = self.lame.size != 0
        if have_material and self.lame.shape[1] != nT:
            raise ValueError(f"mue/lambdae must have |#elements|={nT} entries")
        if self.rhoe.size != 0 and self.rhoe.size != nT:
            raise ValueError(f"rhoe must have |#elements|={nT} entries")
        mu = np.ascontiguousarray(self.lame[0]) if have_material else None
        la = np.ascontiguousarray(self.lame[1]) if have_material else None
        rho = np.ascontiguousarray(self.rhoe) if self.rhoe.size else None
        dbc = np.ascontiguousarray(self.dbc, dtype=np.int64)
        m = np.zeros(nV)
        GP = np.zeros(12 * nT)
        wg = np.zeros(nT)
        lame = np.zeros(2 * nT)
        GVGp = np.zeros(nV + 1, dtype=np.int64)
        GVGe = np.zeros(4 * nT, dtype=np.int64)
        GVGil = np.zeros(4 * nT, dtype=np.int64)
        colors = np.zeros(nV, dtype=np.int64)
        Pptr = np.zeros(_native.MAX_COLORS + 1, dtype=np.int64)
        Padj = np.zeros(nV, dtype=np.int64)
        nP = ctypes.c_int64()
        nAdj = ctypes.c_int64()
        f64, i64 = ctypes.c_double, ctypes.c_int64
        opt = (lambda a: _native.ptr(a, f64) if a is not None else None)
        rc = L.pbat_vbd_data_construct(
            nV, nT, _native.ptr(Xc, f64), _native.ptr(Ec, i64), opt(rho), opt(mu), opt(la),
            _native.ptr(dbc, i64) if dbc.size else None, dbc.size,
            int(self.vertex_coloring_ordering), int(self.vertex_coloring_selection),
            _native.ptr(m, f64), _native.ptr(GP, f64), _native.ptr(wg, f64), _native.ptr(lame, f64),
            _native.ptr(GVGp, i64), _native.ptr(GVGe, i64), _native.ptr(GVGil, i64), _native.ptr(colors, i64),
            _native.ptr(Pptr, i64), _native.ptr(Padj, i64), ctypes.byref(nP), ctypes.byref(nAdj))
        _native.check(rc)
        self.m = m
        self.GP = GP.reshape(3 * nT, 4).T.copy()
        self.wg = wg
        self.lame = lame.reshape(nT, 2).T.copy()
        if self.rhoe.size == 0:
            self.rhoe = np.full(nT, 1e3)
        self.GVGp, self.GVGe, self.GVGilocal = GVGp, GVGe, GVGil
        self.colors = colors
        self.Pptr = Pptr[: nP.value + 1].copy()
        self.Padj = Padj[: nAdj.value].copy()
        # Dirichlet vertices carry no velocity or external force (sim/vbd/Data.cpp:236-238)
        if self.dbc.size:
            self.v[:, self.dbc] = 0.0
            self.aext[:, self.dbc] = 0.0
        if validate:
            self._validate()
        return self

    def _validate(self):
        """sim/vbd/Data.cpp:245-305"""
        nV = self.x.shape[1]
        ok = (self.xt.shape[1] == nV and self.v.shape[1] == nV and self.aext.shape[1] == nV
              and self.m.size == nV and self.B.size == nV and self.x.shape[0] == 3
              and self.xt.shape[0] == 3 and self.v.shape[0] == 3 and self.aext.shape[0] == 3)
        if not ok:
            raise ValueError(
                f"x, v, aext, m and B must have same #columns={nV} as x, and 3 rows (except m and B)")
        acc = self.accelerator
        if acc == AccelerationStrategy.Chebyshev and not (0.0 < self.rho < 1.0):
            raise ValueError("Expected 0 < rho < 1")
        if acc in (AccelerationStrategy.Anderson, AccelerationStrategy.Broyden) and self.manderson < 1:
            raise ValueError("Expected m > 0")
        if acc == AccelerationStrategy.Nesterov:
            if self.nesterov_L <= 0:
                raise ValueError("Expected L > 0")
            if self.nesterov_start < 0:
                raise ValueError("Expected start >= 0")
        if acc == AccelerationStrategy.TrustRegion:
            if self.eta < 0:
                raise ValueError("Expected eta >= 0")
            if self.tau <= 1:
                raise ValueError("Expected tau > 1")


class Integrator:
    """``pbat.sim.vbd.Integrator`` (bindings/pypbat/sim/vbd/Integrator.cpp:30-90) over
    the HIP path: the reference's CPU VBD integrator (sim/vbd/Integrator.h:17-69,
    Integrator.cpp:22-136), stepped by the same device kernels as
    ``pbat.gpu.vbd.Integrator`` but with the CPU integrator's semantics:

    - the integrator the CPU binding's factory builds from ``data.accelerator``:
      Chebyshev and Anderson accelerated, anything else plain VBD (the binding has
      no trust-region integrator); Nesterov and Broyden (CPU-only in the reference)
      are not built here and raise ``ValueError``;
    - ``vt = v`` before the initial guess (Integrator.cpp:31-32), so the
      AdaptiveVbd guess is ``x_t + h v`` (its acceleration estimate is 0);
    - no contact: the CPU integrator reads no collision mesh;
    - ``trace_next_step`` writes the reference's ``{t}.{s}.f.mtx``,
      ``{t}.{s}.grad.mtx`` and ``{t}.{s}.x.mtx`` (Integrator.cpp:202-235), with the
      objective and its gradient reduced in FP64 on the device.

    The state lives on the device in FP32 (the reference CPU path is FP64); ``x``
    and ``v`` read it back as float64 arrays, and ``data`` returns the Data with
    ``x`` and ``v`` refreshed from the device. Assigning ``data`` rebuilds the
    integrator from it; editing the arrays of a returned ``data`` in place does not
    reach the device (assign ``x``/``v`` instead)."""

    def __init__(self, data: Data):
        self._trace = None
        self._build(data)

    def _build(self, data: Data):
        from ..gpu.vbd import Integrator as _DeviceIntegrator

        if data.accelerator in (AccelerationStrategy.Nesterov, AccelerationStrategy.Broyden):
            raise ValueError(f"the {data.accelerator.name}-accelerated VBD integrator (CPU-only in the reference) "
                             "is not available here; Base, Chebyshev and Anderson are")
        accel = data.accelerator if data.accelerator in (AccelerationStrategy.Chebyshev,
                                                         AccelerationStrategy.Anderson) else AccelerationStrategy.Base
        dev = _DeviceIntegrator.__new__(_DeviceIntegrator)
        dev._create(data, contact=False, accelerator=accel)
        self._dev = dev
        self._data = data
        self._strategy = InitializationStrategy(int(data.strategy))
        self._apply_strategy()

    def _apply_strategy(self):
        # sim/vbd/Integrator.cpp:31-32 sets vt = v before InitialPositionsForSolve, whose
        # AdaptiveVbd estimate (vt - vtm1) / h . a then vanishes: x0 = xt + h vt, the
        # Inertia guess (sim/vbd/Kernels.h:45-94)
        s = self._strategy
        self._dev.strategy = InitializationStrategy.Inertia if s == InitializationStrategy.AdaptiveVbd else s

    def step(self, dt: float, iterations: int, substeps: int = 1) -> None:
        """Integrate the VBD simulation 1 time step (sim/vbd/Integrator.cpp:22-45)."""
        if self._trace is not None:
            path, t = self._trace
            self._trace = None  # mTraceIterates = false after the step (Integrator.cpp:44)
            self._dev.trace_objective_step(dt, iterations, substeps, t, path)
        else:
            self._dev.step(dt, iterations, substeps)

    def trace_next_step(self, path: str = ".", t: int = -1) -> None:
        """Mark the next time step for optimization tracing (sim/vbd/Integrator.cpp:47-52)."""
        self._trace = (str(path), int(t))

    # ---- the reference's debugging methods (sim/vbd/Integrator.h:41-49) ----------
    def objective_function(self, xk, xtilde, dt: float) -> float:
        """1/2 |xk - xtilde|_M^2 + dt^2 sum_e w_e Psi_e (sim/vbd/Integrator.cpp:138-166)."""
        return self._dev.objective(dt, xk, xtilde)[0]

    def objective_function_gradient(self, xk, xtilde, dt: float) -> np.ndarray:
        """Its gradient, vertex-major (3 |#nodes|,) (sim/vbd/Integrator.cpp:168-200)."""
        return self._dev.objective(dt, xk, xtilde, gradient=True)[2]

    # ---- properties of the binding ----------------------------------------------
    @property
    def x(self) -> np.ndarray:
        """3x|#nodes| nodal positions"""
        return self._dev.x.astype(np.float64)

    @x.setter
    def x(self, value):
        self._dev.x = value

    @property
    def v(self) -> np.ndarray:
        """3x|#nodes| nodal velocities"""
        return self._dev.v.astype(np.float64)

    @v.setter
    def v(self, value):
        self._dev.v = value

    @property
    def strategy(self) -> InitializationStrategy:
        """Initialization strategy of the step's optimisation"""
        return self._strategy

    @strategy.setter
    def strategy(self, value):
        self._strategy = InitializationStrategy(int(value))
        self._data.strategy = self._strategy
        self._apply_strategy()

    @property
    def kD(self) -> float:
        """Rayleigh damping coefficient"""
        return float(self._data.kD)

    @kD.setter
    def kD(self, value):
        self._data.kD = float(value)
        self._dev.kD = float(value)

    @property
    def detH_residual(self) -> float:
        """Determinant of the residual Hessian for pseudo singularity check"""
        return float(self._data.detH_zero)

    @detH_residual.setter
    def detH_residual(self, value):
        self._data.detH_zero = float(value)
        self._dev.detH_residual = float(value)

    @property
    def data(self) -> Data:
        """The problem data, with x and v read back from the device"""
        self._data.x = self.x
        self._data.v = self.v
        return self._data

    @data.setter
    def data(self, value: Data):
        self._build(value)
