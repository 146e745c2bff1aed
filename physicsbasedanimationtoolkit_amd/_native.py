"""ctypes binding of ``libpbat_vbd.so``, the C ABI declared in ``include/pbat_vbd.h``.

The library holds the HIP kernels for gfx950 and the host code around them.
There is no fallback: if the library is missing or cannot be loaded, every
entry point of this package raises ``ImportError``.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "libpbat_vbd.so")

OK = 0
ERROR_INVALID_ARGUMENT = -1
ERROR_HIP = -2
ERROR_OUT_OF_MEMORY = -3
ERROR_SINGULAR_ELEMENT = -4
ERROR_PALETTE = -5
ERROR_NONFINITE = -6

MAX_COLORS = 128

ACCELERATION_NONE = 0
ACCELERATION_CHEBYSHEV = 1
ACCELERATION_ANDERSON = 2
ACCELERATION_TRUST_REGION = 3

_lib = None

c_i64 = ctypes.c_int64
c_i32 = ctypes.c_int32
c_f32 = ctypes.c_float
c_f64 = ctypes.c_double
p_f64 = ctypes.POINTER(ctypes.c_double)
p_f32 = ctypes.POINTER(ctypes.c_float)
p_i64 = ctypes.POINTER(ctypes.c_int64)


class Data(ctypes.Structure):
    """``pbat_vbd_data`` of include/pbat_vbd.h."""

    _fields_ = [
        ("nV", c_i64), ("nT", c_i64),
        ("x", p_f64), ("v", p_f64), ("aext", p_f64), ("m", p_f64),
        ("E", p_i64), ("GP", p_f64), ("wg", p_f64), ("lame", p_f64),
        ("GVGp", p_i64), ("GVGe", p_i64), ("GVGilocal", p_i64),
        ("nPartitions", c_i64), ("Pptr", p_i64), ("Padj", p_i64),
        ("strategy", c_i32), ("kD", c_f64), ("detHZero", c_f64),
        ("acceleration", c_i32), ("rho", c_f64), ("window", c_i32),
        ("eta", c_f64), ("tau", c_f64), ("curved", c_i32),
        ("nCollisionVertices", c_i64), ("V", p_i64), ("nCollisionTriangles", c_i64), ("F", p_i64),
        ("B", p_i64), ("XVA", p_f64), ("FA", p_f64), ("muC", c_f64), ("muF", c_f64), ("epsv", c_f64),
        ("activeSetUpdateFrequency", c_i32),
    ]


class SweepConfig(ctypes.Structure):
    """``pbat_vbd_sweep_config`` of include/pbat_vbd.h."""

    _fields_ = [
        ("lanes_per_vertex", c_i32), ("incidences_per_lane", c_i32), ("xcd_remap", c_i32),
        ("nontemporal", c_i32), ("fixed_stride", c_i32),
    ]


ADJOINT_PCG = 0
ADJOINT_GAUSS_SEIDEL = 1
ADJOINT_PCG_JACOBI = 2
ADJOINT_PCG_TWO_LEVEL = 3


class AdjointConfig(ctypes.Structure):
    """``pbat_vbd_adjoint_config`` of include/pbat_vbd.h."""

    _fields_ = [("solver", c_i32), ("max_iterations", c_i32), ("tolerance", c_f32), ("check_every", c_i32),
                ("history", c_i32)]


class AdjointStats(ctypes.Structure):
    """``pbat_vbd_adjoint_stats`` of include/pbat_vbd.h."""

    _fields_ = [("solves", c_i32), ("total_iterations", c_i32), ("max_iterations", c_i32),
                ("max_relative_residual", c_f32)]


# (name, restype, argtypes) of every exported entry point
SIGNATURES = [
    ("pbat_vbd_last_error", ctypes.c_char_p, []),
    ("pbat_vbd_version", ctypes.c_char_p, []),
    ("pbat_vbd_data_construct", ctypes.c_int,
     [c_i64, c_i64, p_f64, p_i64, p_f64, p_f64, p_f64, p_i64, c_i64, ctypes.c_int, ctypes.c_int,
      p_f64, p_f64, p_f64, p_f64, p_i64, p_i64, p_i64, p_i64, p_i64, p_i64, p_i64, p_i64]),
    ("pbat_vbd_integrator_create", ctypes.c_int, [ctypes.POINTER(Data), ctypes.POINTER(ctypes.c_void_p)]),
    ("pbat_vbd_integrator_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("pbat_vbd_batch_create", ctypes.c_int, [ctypes.POINTER(Data), c_i32, ctypes.POINTER(ctypes.c_void_p)]),
    ("pbat_vbd_batch_vertex_offsets", ctypes.c_int, [ctypes.c_void_p, p_i64, ctypes.POINTER(c_i32)]),
    ("pbat_vbd_integrator_step", ctypes.c_int, [ctypes.c_void_p, c_f32, c_i32, c_i32]),
    ("pbat_vbd_integrator_traced_step", ctypes.c_int,
     [ctypes.c_void_p, c_f32, c_i32, c_i32, c_i32, ctypes.c_char_p]),
    ("pbat_vbd_integrator_set_positions", ctypes.c_int, [ctypes.c_void_p, p_f32, c_i64]),
    ("pbat_vbd_integrator_set_velocities", ctypes.c_int, [ctypes.c_void_p, p_f32, c_i64]),
    ("pbat_vbd_integrator_set_external_acceleration", ctypes.c_int, [ctypes.c_void_p, p_f32, c_i64]),
    ("pbat_vbd_integrator_get_positions", ctypes.c_int, [ctypes.c_void_p, p_f32, c_i64]),
    ("pbat_vbd_integrator_get_velocities", ctypes.c_int, [ctypes.c_void_p, p_f32, c_i64]),
    ("pbat_vbd_integrator_set_detH_zero", ctypes.c_int, [ctypes.c_void_p, c_f32]),
    ("pbat_vbd_integrator_set_rayleigh_damping", ctypes.c_int, [ctypes.c_void_p, c_f32]),
    ("pbat_vbd_integrator_set_initialization_strategy", ctypes.c_int, [ctypes.c_void_p, c_i32]),
    ("pbat_vbd_integrator_set_block_size", ctypes.c_int, [ctypes.c_void_p, c_i32]),
    ("pbat_vbd_integrator_set_scene_bounding_box", ctypes.c_int, [ctypes.c_void_p, p_f32, p_f32]),
    ("pbat_vbd_integrator_set_lanes_per_vertex", ctypes.c_int, [ctypes.c_void_p, c_i32]),
    ("pbat_vbd_integrator_set_sweep_config", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SweepConfig)]),
    ("pbat_vbd_integrator_get_sweep_config", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SweepConfig)]),
    ("pbat_vbd_integrator_autotune", ctypes.c_int,
     [ctypes.c_void_p, c_f32, c_i32, p_f64, ctypes.POINTER(c_i32)]),
    ("pbat_vbd_integrator_set_graph_capture", ctypes.c_int, [ctypes.c_void_p, c_i32]),
    ("pbat_vbd_integrator_energy", ctypes.c_int, [ctypes.c_void_p, p_f64]),
    ("pbat_vbd_integrator_synchronize", ctypes.c_int, [ctypes.c_void_p]),
    ("pbat_vbd_integrator_objective", ctypes.c_int, [ctypes.c_void_p, p_f32, p_f32, c_f32, p_f64, p_f64, p_f64]),
    ("pbat_vbd_integrator_objective_noalloc", ctypes.c_int, [ctypes.c_void_p, c_f32, p_f64]),
    ("pbat_vbd_integrator_trace_objective_step", ctypes.c_int,
     [ctypes.c_void_p, c_f32, c_i32, c_i32, c_i32, ctypes.c_char_p]),
    ("pbat_vbd_integrator_fp64_solves", ctypes.c_int, [ctypes.c_void_p, p_i64, c_i32]),
    ("pbat_vbd_device_memory", ctypes.c_int, [p_i64, p_i64, c_i32]),
    ("pbat_vbd_integrator_counts", ctypes.c_int, [ctypes.c_void_p, p_i64, p_i64, p_i64]),
    ("pbat_vbd_integrator_contacts", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(c_i32), ctypes.POINTER(c_i32)]),
    ("pbat_vbd_integrator_trust_region_log", ctypes.c_int,
     [ctypes.c_void_p, p_f64, ctypes.POINTER(c_i32), c_i64, p_i64]),
    ("pbat_vbd_integrator_set_timing", ctypes.c_int, [ctypes.c_void_p, c_i32]),
    ("pbat_vbd_integrator_sweep_timing", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(c_f64), p_i64, c_i32]),
    ("pbat_vbd_integrator_set_log_young_modulus", ctypes.c_int, [ctypes.c_void_p, p_f32, c_i64]),
    ("pbat_vbd_integrator_get_log_young_modulus", ctypes.c_int, [ctypes.c_void_p, p_f32, c_i64]),
    ("pbat_vbd_rollout_forward", ctypes.c_int, [ctypes.c_void_p, c_f32, c_i32, c_i32, c_i32]),
    ("pbat_vbd_rollout_backward", ctypes.c_int,
     [ctypes.c_void_p, p_f32, p_f32, p_f32, p_f32, p_f32, ctypes.POINTER(AdjointConfig),
      ctypes.POINTER(AdjointStats)]),
    ("pbat_vbd_rollout_generation", ctypes.c_int, [ctypes.c_void_p, p_i64, ctypes.POINTER(c_i32)]),
    ("pbat_vbd_hessian_vector", ctypes.c_int, [ctypes.c_void_p, p_f32, c_f32, p_f32, p_f32]),
    ("pbat_vbd_adjoint_solve", ctypes.c_int,
     [ctypes.c_void_p, p_f32, c_f32, p_f32, p_f32, ctypes.POINTER(AdjointConfig), ctypes.POINTER(AdjointStats)]),
]


def lib():
    """Load libpbat_vbd.so (once) and declare its signatures."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY_PATH):
            raise ImportError(
                f"{LIBRARY_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `make -C physicsbasedanimationtoolkit_amd/csrc`")
        L = ctypes.CDLL(LIBRARY_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class HipError(RuntimeError):
    pass


def check(rc: int) -> None:
    """Raise the Python counterpart of a C ABI status code."""
    if rc == OK:
        return
    msg = lib().pbat_vbd_last_error().decode(errors="replace")
    if rc in (ERROR_INVALID_ARGUMENT, ERROR_SINGULAR_ELEMENT):
        raise ValueError(msg)
    if rc == ERROR_OUT_OF_MEMORY:
        raise MemoryError(msg)
    if rc == ERROR_HIP:
        raise HipError(msg)
    raise RuntimeError(f"{msg} (status {rc})")


def device_memory(reset_peak: bool = False):
    """(bytes held now, peak bytes) of the library's device arrays in this process."""
    cur, peak = ctypes.c_int64(), ctypes.c_int64()
    check(lib().pbat_vbd_device_memory(ctypes.byref(cur), ctypes.byref(peak), int(bool(reset_peak))))
    return cur.value, peak.value


def ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))
