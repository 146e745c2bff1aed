"""``pbat.gpu.vbd`` counterpart: the VBD integrator on an MI355X.

Mirrors ``pbat.gpu.vbd.Integrator`` (bindings/pypbat/gpu/vbd/Integrator.cpp:19-105)
over ``pbat::gpu::vbd::Integrator`` (gpu/vbd/Integrator.h:33-150): same
constructor, ``step``/``traced_step`` signatures and defaults, and the same
properties (``x``, ``v`` read-write; ``a``, ``detH_residual``, ``kD``,
``strategy``, ``gpu_block_size``, ``scene_bounding_box`` write-only). All work
runs in hand-written HIP kernels of libpbat_vbd.so, called through its C ABI.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .. import _native
from ..sim.vbd import AccelerationStrategy, Data, InitializationStrategy


def _vertex_matrix_f32(a, n: int, name: str) -> np.ndarray:
    a = np.asarray(a)
    if a.ndim != 2 or a.shape[0] != 3:
        raise ValueError(f"Expected input dimensions 3x{n}, but got {'x'.join(map(str, a.shape))} for {name}")
    # column-major 3 x n == row-major n x 3
    return np.ascontiguousarray(a.T, dtype=np.float32)


def _write_only(name):
    def getter(self):
        raise AttributeError(f"'{name}' is write-only")
    return getter


# gpu/vbd/Integrator.cu:15-33: the integrators the device factory builds
_ACCELERATION = {AccelerationStrategy.Base: _native.ACCELERATION_NONE,
                 AccelerationStrategy.Chebyshev: _native.ACCELERATION_CHEBYSHEV,
                 AccelerationStrategy.Anderson: _native.ACCELERATION_ANDERSON,
                 AccelerationStrategy.TrustRegion: _native.ACCELERATION_TRUST_REGION}


def _marshal(data: Data, contact: bool = True, accelerator=None):
    """pbat_vbd_data view of a constructed Data, plus the arrays it points into.
    contact=False leaves the collision mesh out; accelerator overrides data.accelerator."""
    if not isinstance(data, Data):
        raise TypeError("expected a pbat.sim.vbd.Data")
    accelerator = data.accelerator if accelerator is None else accelerator
    if accelerator not in _ACCELERATION:
        raise ValueError(
            f"acceleration strategy {AccelerationStrategy(accelerator).name} is not available on this device path; "
            "only the base, Chebyshev-, Anderson- and trust-region-accelerated VBD integrators are implemented")
    if data.m.size == 0:
        raise ValueError("data must be constructed (call data.construct()) before creating an Integrator")
    nV, nT = data.x.shape[1], data.E.shape[1]
    keep = {
        "x": np.ascontiguousarray(data.x.T, dtype=np.float64),
        "v": np.ascontiguousarray(data.v.T, dtype=np.float64),
        "aext": np.ascontiguousarray(data.aext.T, dtype=np.float64),
        "m": np.ascontiguousarray(data.m, dtype=np.float64),
        "E": np.ascontiguousarray(data.E.T, dtype=np.int64),
        "GP": np.ascontiguousarray(data.GP.T, dtype=np.float64),
        "wg": np.ascontiguousarray(data.wg, dtype=np.float64),
        "lame": np.ascontiguousarray(data.lame.T, dtype=np.float64),
        "GVGp": np.ascontiguousarray(data.GVGp, dtype=np.int64),
        "GVGe": np.ascontiguousarray(data.GVGe, dtype=np.int64),
        "GVGilocal": np.ascontiguousarray(data.GVGilocal, dtype=np.int64),
        "Pptr": np.ascontiguousarray(data.Pptr, dtype=np.int64),
        "Padj": np.ascontiguousarray(data.Padj, dtype=np.int64),
    }
    for k in ("x", "v", "aext"):
        if keep[k].shape != (nV, 3):
            raise ValueError(f"data.{k} must be 3x{nV}")
    if keep["Padj"].size == 0:
        keep["Padj"] = np.zeros(1, dtype=np.int64)  # non-null pointer; no entries are read
    d = _native.Data()
    d.nV, d.nT = nV, nT
    f64, i64 = ctypes.c_double, ctypes.c_int64
    for k in ("x", "v", "aext", "m", "GP", "wg", "lame"):
        setattr(d, k, _native.ptr(keep[k], f64))
    for k in ("E", "GVGp", "GVGe", "GVGilocal", "Pptr", "Padj"):
        setattr(d, k, _native.ptr(keep[k], i64))
    d.nPartitions = max(keep["Pptr"].size - 1, 0)
    d.strategy = int(data.strategy)
    d.kD = float(data.kD)
    d.detHZero = float(data.detH_zero)
    d.acceleration = _ACCELERATION[accelerator]
    d.rho = float(data.rho)
    d.window = int(data.manderson)
    d.eta, d.tau, d.curved = float(data.eta), float(data.tau), int(bool(data.curved))
    # collision mesh and contact model (sim/vbd/Data.h:165-178, 225-229)
    if contact and data.V.size and data.F.size:
        keep["V"] = np.ascontiguousarray(data.V, dtype=np.int64).reshape(-1)
        keep["F"] = np.ascontiguousarray(np.asarray(data.F, dtype=np.int64).T)
        keep["B"] = np.ascontiguousarray(data.B, dtype=np.int64).reshape(-1)
        keep["XVA"] = np.ascontiguousarray(data.XVA, dtype=np.float64)
        keep["FA"] = np.ascontiguousarray(data.FA, dtype=np.float64)
        if keep["B"].size != nV or keep["XVA"].size != nV or keep["FA"].size != keep["F"].shape[0]:
            raise ValueError("B and XVA must have |#vertices| entries and FA one per collision triangle")
        d.nCollisionVertices, d.V = keep["V"].size, _native.ptr(keep["V"], i64)
        d.nCollisionTriangles, d.F = keep["F"].shape[0], _native.ptr(keep["F"], i64)
        d.B, d.XVA, d.FA = _native.ptr(keep["B"], i64), _native.ptr(keep["XVA"], f64), _native.ptr(keep["FA"], f64)
    d.muC, d.muF, d.epsv = float(data.muC), float(data.muF), float(data.epsv)
    d.activeSetUpdateFrequency = int(data.active_set_update_frequency)
    return d, keep


class Integrator:
    """VBD time integrator running on the current HIP device."""

    def __init__(self, data: Data):
        self._create(data)

    def _create(self, data: Data, contact: bool = True, accelerator=None):
        self._lib = _native.lib()
        d, keep = _marshal(data, contact, accelerator)
        self._nV = d.nV
        self._nT = d.nT
        self._nCV = d.nCollisionVertices
        handle = ctypes.c_void_p()
        _native.check(self._lib.pbat_vbd_integrator_create(ctypes.byref(d), ctypes.byref(handle)))
        del keep
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.pbat_vbd_integrator_destroy(h)
            self._h = None

    # ---- stepping -------------------------------------------------------------
    def step(self, dt: float = 0.01, iterations: int = 20, substeps: int = 1) -> None:
        """Integrate 1 time step (gpu/vbd/Integrator.cu:58-61)."""
        _native.check(self._lib.pbat_vbd_integrator_step(self._h, float(dt), int(iterations), int(substeps)))

    def traced_step(self, dt: float = 0.01, iterations: int = 20, substeps: int = 1, t: int = 0,
                    dir: str = ".") -> None:
        """Integrate 1 time step and write {variable}.t.{t}.s.{s}[.k.{k}].mtx files into dir."""
        os.makedirs(dir, exist_ok=True)
        _native.check(self._lib.pbat_vbd_integrator_traced_step(
            self._h, float(dt), int(iterations), int(substeps), int(t), str(dir).encode()))

    # ---- state ----------------------------------------------------------------
    def _get(self, fn) -> np.ndarray:
        out = np.empty((self._nV, 3), dtype=np.float32)
        _native.check(fn(self._h, _native.ptr(out, ctypes.c_float), self._nV))
        return out.T

    def _set(self, fn, value, name):
        a = _vertex_matrix_f32(value, self._nV, name)
        _native.check(fn(self._h, _native.ptr(a, ctypes.c_float), a.shape[0]))

    @property
    def x(self) -> np.ndarray:
        """|#dims|x|#vertices| vertex positions"""
        return self._get(self._lib.pbat_vbd_integrator_get_positions)

    @x.setter
    def x(self, value):
        self._set(self._lib.pbat_vbd_integrator_set_positions, value, "positions")

    @property
    def v(self) -> np.ndarray:
        """|#dims|x|#vertices| vertex velocities"""
        return self._get(self._lib.pbat_vbd_integrator_get_velocities)

    @v.setter
    def v(self, value):
        self._set(self._lib.pbat_vbd_integrator_set_velocities, value, "velocities")

    def _set_a(self, value):
        self._set(self._lib.pbat_vbd_integrator_set_external_acceleration, value, "external acceleration")

    a = property(_write_only("a"), _set_a, doc="|#dims|x|#vertices| vertex external accelerations")

    def _set_detH(self, value):
        _native.check(self._lib.pbat_vbd_integrator_set_detH_zero(self._h, float(value)))

    detH_residual = property(_write_only("detH_residual"), _set_detH,
                             doc="Numerical zero used in the Hessian determinant check")

    def _set_kD(self, value):
        _native.check(self._lib.pbat_vbd_integrator_set_rayleigh_damping(self._h, float(value)))

    kD = property(_write_only("kD"), _set_kD, doc="Uniform Rayleigh damping coefficient on the mesh")

    def _set_strategy(self, value):
        _native.check(self._lib.pbat_vbd_integrator_set_initialization_strategy(
            self._h, int(InitializationStrategy(int(value)))))

    strategy = property(_write_only("strategy"), _set_strategy,
                        doc="VBD's time step minimization initialization strategy")

    def _set_block(self, value):
        _native.check(self._lib.pbat_vbd_integrator_set_block_size(self._h, int(value)))

    gpu_block_size = property(_write_only("gpu_block_size"), _set_block,
                              doc="Number of threads per GPU workgroup of the sweep kernel")

    def _set_box(self, box):
        lo = np.ascontiguousarray(np.asarray(box[0], dtype=np.float32).reshape(3))
        hi = np.ascontiguousarray(np.asarray(box[1], dtype=np.float32).reshape(3))
        _native.check(self._lib.pbat_vbd_integrator_set_scene_bounding_box(
            self._h, _native.ptr(lo, ctypes.c_float), _native.ptr(hi, ctypes.c_float)))

    scene_bounding_box = property(_write_only("scene_bounding_box"), _set_box,
                                  doc="Tuple of (min,max) scene bounding box extremities")

    # ---- extensions (no reference counterpart) ----------------------------------
    def _set_lanes(self, value):
        _native.check(self._lib.pbat_vbd_integrator_set_lanes_per_vertex(self._h, int(value)))

    lanes_per_vertex = property(_write_only("lanes_per_vertex"), _set_lanes,
                                doc="Lanes cooperating on one vertex in the sweep kernel (0 = automatic)")

    def set_sweep_config(self, lanes: int = 0, incidences_per_lane: int = 0, xcd_remap: bool = True,
                         nontemporal: bool = False, fixed_stride: bool | None = None) -> None:
        """Sweep-kernel shape: lanes per vertex (0 = auto), incidences prefetched per
        lane (0 = auto, else 1..4), XCD-aware workgroup order, non-temporal stream
        loads, fixed-stride (ELL) incidence rows (None = keep the current layout)."""
        if fixed_stride is None:
            fixed_stride = self.sweep_config()["fixed_stride"]
        c = _native.SweepConfig(int(lanes), int(incidences_per_lane), int(bool(xcd_remap)), int(bool(nontemporal)),
                                int(bool(fixed_stride)))
        _native.check(self._lib.pbat_vbd_integrator_set_sweep_config(self._h, ctypes.byref(c)))

    def sweep_config(self) -> dict:
        """Sweep shape in effect (automatic choices resolved)."""
        c = _native.SweepConfig()
        _native.check(self._lib.pbat_vbd_integrator_get_sweep_config(self._h, ctypes.byref(c)))
        return {"lanes_per_vertex": c.lanes_per_vertex, "incidences_per_lane": c.incidences_per_lane,
                "xcd_remap": bool(c.xcd_remap), "nontemporal": bool(c.nontemporal),
                "fixed_stride": bool(c.fixed_stride)}

    def autotune(self, dt: float = 0.01, sweeps: int = 10):
        """Pick the fastest sweep shape by timing full sweeps; state is unchanged.
        Returns rows (lanes, incidences per lane, xcd remap, non-temporal, fixed stride, us per sweep)."""
        t = np.zeros(192)
        n = ctypes.c_int32()
        _native.check(self._lib.pbat_vbd_integrator_autotune(self._h, float(dt), int(sweeps),
                                                             _native.ptr(t, ctypes.c_double), ctypes.byref(n)))
        return [tuple(r) for r in t[: 6 * min(n.value, 32)].reshape(-1, 6)]

    def _set_graph(self, value):
        _native.check(self._lib.pbat_vbd_integrator_set_graph_capture(self._h, int(bool(value))))

    graph_capture = property(_write_only("graph_capture"), _set_graph,
                             doc="Replay each step's launches from a captured hipGraph")

    def energy(self) -> np.ndarray:
        """(kinetic, elastic, external potential) energy terms reduced in FP64 on the device."""
        out = np.zeros(3)
        _native.check(self._lib.pbat_vbd_integrator_energy(self._h, _native.ptr(out, ctypes.c_double)))
        return out

    def counts(self):
        """(#colours, #updated vertices, #vertex-tet incidences of updated vertices)"""
        p, u, i = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _native.check(self._lib.pbat_vbd_integrator_counts(self._h, ctypes.byref(p), ctypes.byref(u),
                                                           ctypes.byref(i)))
        return p.value, u.value, i.value

    def contacts(self):
        """(active, fc) after the last step: the active flag of each collision vertex (in
        the order of data.V) and up to 8 contacting triangles per vertex (|#vertices| x 8,
        -1 padded). Raises ValueError when the integrator runs no contact."""
        active = np.zeros(max(self._nCV, 1), dtype=np.int32)
        fc = np.zeros((self._nV, 8), dtype=np.int32)
        _native.check(self._lib.pbat_vbd_integrator_contacts(self._h, _native.ptr(active, ctypes.c_int32),
                                                             _native.ptr(fc, ctypes.c_int32)))
        return active[: self._nCV].astype(bool), fc

    def trust_region_log(self):
        """[(t, accepted)] of the linear trust-region path in the last step: one entry per
        iteration k >= 2 of each substep (the reference logs the same through its
        profiler, gpu/impl/vbd/TrustRegionIntegrator.cu:111,133,146)."""
        n = ctypes.c_int64()
        _native.check(self._lib.pbat_vbd_integrator_trust_region_log(self._h, None, None, 0, ctypes.byref(n)))
        t = np.zeros(max(n.value, 1))
        acc = np.zeros(max(n.value, 1), dtype=np.int32)
        _native.check(self._lib.pbat_vbd_integrator_trust_region_log(
            self._h, _native.ptr(t, ctypes.c_double), _native.ptr(acc, ctypes.c_int32), n.value, ctypes.byref(n)))
        return [(float(t[k]), bool(acc[k])) for k in range(n.value)]

    def objective(self, dt: float, x=None, xtilde=None, gradient: bool = False):
        """FP64 objective of the step (sim::vbd::Integrator::ObjectiveFunction /
        ObjectiveFunctionGradient, sim/vbd/Integrator.cpp:138-200) reduced on the device:
        f = 1/2 |x - x~|_M^2 + dt^2 sum_e w_e Psi_e at x and xtilde (3 x nV; None = the
        current positions / the inertial target of the last substep). Returns
        (f, squared gradient norm over the non-Dirichlet vertices) and, with
        gradient=True, also the gradient as a vertex-major (3 nV,) vector."""
        xa = None if x is None else _vertex_matrix_f32(x, self._nV, "x")
        la = None if xtilde is None else _vertex_matrix_f32(xtilde, self._nV, "xtilde")
        f, r2 = ctypes.c_double(), ctypes.c_double()
        g = np.zeros(3 * self._nV) if gradient else None
        fp = ctypes.c_float
        _native.check(self._lib.pbat_vbd_integrator_objective(
            self._h, None if xa is None else _native.ptr(xa, fp), None if la is None else _native.ptr(la, fp),
            float(dt), ctypes.byref(f), None if g is None else _native.ptr(g, ctypes.c_double), ctypes.byref(r2)))
        return (f.value, r2.value, g) if gradient else (f.value, r2.value)

    def trace_objective_step(self, dt: float, iterations: int, substeps: int = 1, t: int = -1,
                             path: str = ".") -> None:
        """One step whose iterates are traced as sim::vbd::Integrator::TraceNextStep does
        (sim/vbd/Integrator.cpp:202-235): {path}/{t}.{s}.f.mtx, .grad.mtx and .x.mtx per substep."""
        os.makedirs(path, exist_ok=True)
        _native.check(self._lib.pbat_vbd_integrator_trace_objective_step(
            self._h, float(dt), int(iterations), int(substeps), int(t), str(path).encode()))

    def fp64_solves(self, reset: bool = False) -> int:
        """Vertex block solves of the sweep that took the FP64 path (ill-conditioned
        3x3 blocks) since creation or the last reset."""
        n = ctypes.c_int64()
        _native.check(self._lib.pbat_vbd_integrator_fp64_solves(self._h, ctypes.byref(n), int(bool(reset))))
        return n.value

    def synchronize(self) -> None:
        _native.check(self._lib.pbat_vbd_integrator_synchronize(self._h))

    def set_timing(self, enabled: bool) -> None:
        _native.check(self._lib.pbat_vbd_integrator_set_timing(self._h, int(bool(enabled))))

    def sweep_timing(self, reset: bool = False):
        """(summed HIP-event milliseconds of sweep launches, number of launches)"""
        ms, n = ctypes.c_double(), ctypes.c_int64()
        _native.check(self._lib.pbat_vbd_integrator_sweep_timing(self._h, ctypes.byref(ms), ctypes.byref(n),
                                                                 int(bool(reset))))
        return ms.value, n.value

    # ---- differentiable rollouts (no reference counterpart) ---------------------
    @property
    def element_count(self) -> int:
        return self._nT

    def set_log_young_modulus(self, log_y) -> None:
        """Per-tet Young's modulus exp(log_y) (|#elements| values), Poisson ratios kept."""
        a = np.ascontiguousarray(np.asarray(log_y, dtype=np.float32).reshape(-1))
        _native.check(self._lib.pbat_vbd_integrator_set_log_young_modulus(
            self._h, _native.ptr(a, ctypes.c_float), a.size))

    def log_young_modulus(self) -> np.ndarray:
        out = np.empty(self._nT, dtype=np.float32)
        _native.check(self._lib.pbat_vbd_integrator_get_log_young_modulus(
            self._h, _native.ptr(out, ctypes.c_float), self._nT))
        return out

    def rollout_forward(self, dt: float = 0.01, iterations: int = 20, steps: int = 1,
                        checkpoint_every: int = 10) -> None:
        """Step `steps` times (1 substep each), keeping the state every
        `checkpoint_every` steps for rollout_backward."""
        _native.check(self._lib.pbat_vbd_rollout_forward(self._h, float(dt), int(iterations), int(steps),
                                                         int(checkpoint_every)))

    def rollout_generation(self):
        """(generation, held): identifies the rollout rollout_backward would
        differentiate; the generation changes whenever a rollout is started or
        discarded (a setter of the state, moduli or solver settings)."""
        g, v = ctypes.c_int64(), ctypes.c_int32()
        _native.check(self._lib.pbat_vbd_rollout_generation(self._h, ctypes.byref(g), ctypes.byref(v)))
        return g.value, bool(v.value)

    @staticmethod
    def adjoint_config(solver: str = "pcg", max_iterations: int = 200, tolerance: float = 1e-6,
                       check_every: int = 5, history: int = 6) -> _native.AdjointConfig:
        """history: the block Jacobi and two-level solvers start each solve of a
        backward pass from the H-norm projection of its solution onto the span
        of the last `history` solutions (0..8; 0 starts from the last one)."""
        kinds = {"pcg": _native.ADJOINT_PCG, "gauss_seidel": _native.ADJOINT_GAUSS_SEIDEL,
                 "pcg_jacobi": _native.ADJOINT_PCG_JACOBI, "pcg_two_level": _native.ADJOINT_PCG_TWO_LEVEL}
        if solver not in kinds:
            raise ValueError(f"solver must be one of {sorted(kinds)}")
        return _native.AdjointConfig(kinds[solver], int(max_iterations), float(tolerance), int(check_every),
                                     int(history))

    def rollout_backward(self, dL_dx, dL_dv=None, **solver):
        """Gradients of a loss L(x_T, v_T) of the last rollout_forward.

        dL_dx, dL_dv: 3 x nV. Returns dict(log_young_modulus=[nT], v0=3 x nV,
        x0=3 x nV, stats=dict). The end-of-rollout state is kept."""
        gx = _vertex_matrix_f32(dL_dx, self._nV, "dL/dx")
        gv = None if dL_dv is None else _vertex_matrix_f32(dL_dv, self._nV, "dL/dv")
        dlogy = np.zeros(self._nT, dtype=np.float32)
        dv0 = np.zeros((self._nV, 3), dtype=np.float32)
        dx0 = np.zeros((self._nV, 3), dtype=np.float32)
        cfg = self.adjoint_config(**solver)
        st = _native.AdjointStats()
        f = ctypes.c_float
        _native.check(self._lib.pbat_vbd_rollout_backward(
            self._h, _native.ptr(gx, f), None if gv is None else _native.ptr(gv, f), _native.ptr(dlogy, f),
            _native.ptr(dv0, f), _native.ptr(dx0, f), ctypes.byref(cfg), ctypes.byref(st)))
        return {"log_young_modulus": dlogy, "v0": dv0.T, "x0": dx0.T,
                "stats": {"solves": st.solves, "total_iterations": st.total_iterations,
                          "max_iterations": st.max_iterations,
                          "max_relative_residual": st.max_relative_residual}}

    def hessian_vector(self, x, dt: float, p) -> np.ndarray:
        """(M / dt^2 + d2U/dx2(x)) p on free vertices (3 x nV)."""
        xa = _vertex_matrix_f32(x, self._nV, "x")
        pa = _vertex_matrix_f32(p, self._nV, "p")
        out = np.zeros((self._nV, 3), dtype=np.float32)
        f = ctypes.c_float
        _native.check(self._lib.pbat_vbd_hessian_vector(self._h, _native.ptr(xa, f), float(dt), _native.ptr(pa, f),
                                                        _native.ptr(out, f)))
        return out.T

    def adjoint_solve(self, x, dt: float, b, **solver):
        """Solve (M / dt^2 + d2U/dx2(x)) y = b on free vertices; returns (y, stats)."""
        xa = _vertex_matrix_f32(x, self._nV, "x")
        ba = _vertex_matrix_f32(b, self._nV, "b")
        y = np.zeros((self._nV, 3), dtype=np.float32)
        cfg = self.adjoint_config(**solver)
        st = _native.AdjointStats()
        f = ctypes.c_float
        _native.check(self._lib.pbat_vbd_adjoint_solve(self._h, _native.ptr(xa, f), float(dt), _native.ptr(ba, f),
                                                       _native.ptr(y, f), ctypes.byref(cfg), ctypes.byref(st)))
        return y.T, {"iterations": st.total_iterations, "relative_residual": st.max_relative_residual}


class BatchIntegrator(Integrator):
    """One integrator for many independent meshes (no reference counterpart).

    The meshes are stepped together: one launch per colour covers the whole
    batch, and each mesh is swept in the same order as its own ``Integrator``
    would sweep it. ``x``/``v`` are the meshes' ``3 x nV_b`` matrices
    concatenated in batch order; ``mesh_positions()`` splits them.
    """

    def __init__(self, datas):
        self._lib = _native.lib()
        datas = list(datas)
        if not datas:
            raise ValueError("BatchIntegrator needs at least one Data")
        views = [_marshal(d) for d in datas]
        self._nT = int(sum(v[0].nT for v in views))
        self._nCV = 0
        arr = (_native.Data * len(views))(*[v[0] for v in views])
        handle = ctypes.c_void_p()
        _native.check(self._lib.pbat_vbd_batch_create(arr, len(views), ctypes.byref(handle)))
        del views
        self._h = handle
        count = ctypes.c_int32()
        _native.check(self._lib.pbat_vbd_batch_vertex_offsets(self._h, None, ctypes.byref(count)))
        offs = np.zeros(count.value + 1, dtype=np.int64)
        _native.check(self._lib.pbat_vbd_batch_vertex_offsets(self._h, _native.ptr(offs, ctypes.c_int64),
                                                              ctypes.byref(count)))
        self.offsets = offs
        self._nV = int(offs[-1])

    @property
    def mesh_count(self) -> int:
        return self.offsets.size - 1

    def mesh_positions(self):
        x = self.x
        return [x[:, self.offsets[b]:self.offsets[b + 1]] for b in range(self.mesh_count)]

    def mesh_velocities(self):
        v = self.v
        return [v[:, self.offsets[b]:self.offsets[b + 1]] for b in range(self.mesh_count)]
