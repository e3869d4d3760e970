"""ctypes binding of libxadv.so (include/xadv.h).

The compiled HIP library is the only compute path.  If it is missing this
module raises on first use: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxadv.so")

XADV_OK, XADV_E_ARG, XADV_E_HIP, XADV_E_STATE = 0, 1, 2, 3
NMETRIC = 24
NCHECK = 16

# Exported symbols, in the order of include/xadv.h.
EXPORTS = (
    "xadv_batch_create", "xadv_batch_destroy", "xadv_instance_load", "xadv_instance_init", "xadv_set_active",
    "xadv_set_inlet_velocity", "xadv_step", "xadv_readout", "xadv_stats", "xadv_fields", "xadv_metrics",
    "xadv_checks", "xadv_get_stream", "xadv_geometry_mask", "xadv_heat_map", "xadv_instance_load_device",
    "xadv_last_error", "xadv_version",
)


class BatchDesc(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int), ("fluid", ctypes.c_int), ("thermal", ctypes.c_int), ("general", ctypes.c_int),
        ("slots", ctypes.c_int), ("cap_nx", ctypes.c_int), ("cap_ny", ctypes.c_int), ("cap_nz", ctypes.c_int),
        ("block", ctypes.c_int), ("xcd_remap", ctypes.c_int), ("graph_chunk", ctypes.c_int),
        ("cells_per_thread", ctypes.c_int), ("fluid_storage", ctypes.c_int), ("alternate_sweep", ctypes.c_int),
        ("thermal_q", ctypes.c_int), ("slab_planes", ctypes.c_int), ("thermal_prefetch", ctypes.c_int),
    ]


STORAGE_AUTO, STORAGE_DENSE, STORAGE_COMPACT = 0, 1, 2

HEAT_MAX_PRIMS, HEAT_MAX_RADIUS = 16, 63


class GeometryDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("buffer", ctypes.c_int),
        ("base", ctypes.c_int), ("top", ctypes.c_int), ("kind", ctypes.c_int),
        ("feature_size", ctypes.c_double), ("spacing", ctypes.c_double), ("period", ctypes.c_double),
        ("threshold", ctypes.c_double), ("manifold", ctypes.c_int),
    ]


class HeatPrim(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int), ("x0", ctypes.c_int), ("x1", ctypes.c_int), ("y0", ctypes.c_int),
        ("y1", ctypes.c_int), ("cx", ctypes.c_int), ("cy", ctypes.c_int), ("r2", ctypes.c_int),
        ("value", ctypes.c_float), ("gx", ctypes.c_double), ("gy", ctypes.c_double), ("den", ctypes.c_double),
        ("amplitude", ctypes.c_double),
    ]


class HeatDesc(ctypes.Structure):
    _fields_ = [
        ("mx", ctypes.c_int), ("my", ctypes.c_int), ("base", ctypes.c_float), ("nprims", ctypes.c_int),
        ("prims", ctypes.POINTER(HeatPrim)), ("radius", ctypes.c_int), ("weights", ctypes.POINTER(ctypes.c_double)),
        ("normalise", ctypes.c_int),
    ]


class InstanceDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("mask", ctypes.c_void_p), ("heat", ctypes.c_void_p),
        ("tau", ctypes.c_double), ("tau_th_fluid", ctypes.c_double), ("tau_th_solid", ctypes.c_double),
        ("heat_power", ctypes.c_double), ("T_in", ctypes.c_double), ("rho_out", ctypes.c_double),
        ("u_in", ctypes.c_double),
        ("periodic", ctypes.c_int * 3), ("force_x", ctypes.c_double), ("u_wall", ctypes.c_double),
        ("dirichlet_x", ctypes.c_int), ("dirichlet_T", ctypes.c_double), ("frozen_velocity", ctypes.c_int),
    ]


class XadvError(RuntimeError):
    pass


_lib = None


def load():
    """Load libxadv.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m xadvcool_amd.build` "
                          "(or __graft_entry__.build()); there is no fallback path")
    lib = ctypes.CDLL(LIB_PATH)
    vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.xadv_batch_create.argtypes = [ctypes.POINTER(BatchDesc), ctypes.POINTER(vp)]
    lib.xadv_batch_destroy.argtypes = [vp]
    lib.xadv_instance_load.argtypes = [vp, ip, ctypes.POINTER(InstanceDesc)]
    lib.xadv_instance_init.argtypes = [vp, ip, vp, vp, vp]
    lib.xadv_set_active.argtypes = [vp, ip, ip]
    lib.xadv_set_inlet_velocity.argtypes = [vp, ip, dp]
    lib.xadv_step.argtypes = [vp, ip, ip]
    lib.xadv_readout.argtypes = [vp, ip]
    lib.xadv_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    lib.xadv_fields.argtypes = [vp, ip, vp, vp, vp]
    lib.xadv_metrics.argtypes = [vp, ip, ip, ip, ip, ctypes.POINTER(ctypes.c_double)]
    lib.xadv_get_stream.argtypes = [vp, ctypes.POINTER(vp)]
    lib.xadv_checks.argtypes = [vp, ip, dp, dp, ctypes.POINTER(ctypes.c_double)]
    lib.xadv_geometry_mask.argtypes = [vp, ctypes.POINTER(GeometryDesc), vp]
    lib.xadv_heat_map.argtypes = [vp, ctypes.POINTER(HeatDesc), ip, ip, ip, ip, vp]
    lib.xadv_instance_load_device.argtypes = [vp, ip, ctypes.POINTER(InstanceDesc)]
    lib.xadv_last_error.restype = ctypes.c_char_p
    lib.xadv_version.restype = ctypes.c_char_p
    for name in EXPORTS:
        if name not in ("xadv_last_error", "xadv_version"):
            getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != XADV_OK:
        msg = load().xadv_last_error().decode()
        if rc == XADV_E_ARG:
            raise ValueError(f"{what}: {msg}")
        raise XadvError(f"{what} failed (code {rc}): {msg}")
