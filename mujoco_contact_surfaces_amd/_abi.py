"""ctypes binding of include/hcs.h (libhcs.so, built in-tree).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libhcs.so")

HCS_OK = 0
HCS_ERR_ARG = -1
HCS_ERR_HIP = -2
HCS_ERR_OOM = -3
HCS_ERR_CAPACITY = -4
HCS_ERR_STATE = -5
HCS_ERR_LIMIT = -6


class Props(C.Structure):
    _fields_ = [("hydroelastic_modulus", C.c_double), ("dissipation", C.c_double),
                ("resolution_hint", C.c_double), ("static_friction", C.c_double),
                ("dynamic_friction", C.c_double)]


class PairIn(C.Structure):
    _fields_ = [("geom_a", C.c_int32), ("geom_b", C.c_int32), ("reserved", C.c_int32 * 2),
                ("X_WA", C.c_double * 12), ("X_WB", C.c_double * 12),
                ("V_WA", C.c_double * 6), ("V_WB", C.c_double * 6)]


class PairOut(C.Structure):
    _fields_ = [("F_A", C.c_double * 3), ("tau_A", C.c_double * 3), ("F_B", C.c_double * 3),
                ("tau_B", C.c_double * 3), ("n_candidates", C.c_int64), ("n_polygons", C.c_int64),
                ("n_faces", C.c_int64), ("status", C.c_int64)]


class StepOpts(C.Structure):
    _fields_ = [("representation", C.c_int32), ("apply_forces", C.c_int32), ("dt", C.c_double),
                ("keep_polygons", C.c_int32), ("reserved", C.c_int32)]


class Faces(C.Structure):
    _fields_ = [("capacity", C.c_int64), ("count", C.c_int64), ("rec", C.POINTER(C.c_double)),
                ("key", C.POINTER(C.c_int64))]


class TaxelGrid(C.Structure):
    _fields_ = [("geom", C.c_int32), ("sampling_resolution", C.c_int32), ("window", C.c_int32),
                ("reserved", C.c_int32), ("sigma", C.c_double), ("X_W_sensor", C.c_double * 12),
                ("half_size", C.c_double * 3), ("resolution", C.c_double)]


class TaxelSensor(C.Structure):
    _fields_ = [("geom", C.c_int32), ("ntaxels", C.c_int32), ("method", C.c_int32), ("sample_method", C.c_int32),
                ("visualize", C.c_int32), ("reserved", C.c_int32), ("include_margin", C.c_double),
                ("sample_resolution", C.c_double), ("X_W_sensor", C.c_double * 12),
                ("taxels", C.POINTER(C.c_double))]


class CurvedSensor(C.Structure):
    _fields_ = [("geom", C.c_int32), ("ntaxels", C.c_int32), ("nsamples", C.c_int32), ("reserved", C.c_int32),
                ("include_margin", C.c_double), ("X_W_sensor", C.c_double * 12), ("points", C.POINTER(C.c_double)),
                ("normals", C.POINTER(C.c_double)), ("taxel_ptr", C.POINTER(C.c_int32)),
                ("sample_idx", C.POINTER(C.c_int32)), ("weights", C.POINTER(C.c_double))]


class MjData(C.Structure):
    _fields_ = [("geom_xpos", C.POINTER(C.c_double)), ("geom_xmat", C.POINTER(C.c_double)),
                ("geom_bodyid", C.POINTER(C.c_int32)), ("user", C.c_void_p),
                ("object_velocity", C.CFUNCTYPE(None, C.c_void_p, C.c_int32, C.POINTER(C.c_double))),
                ("apply_ft", C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                         C.POINTER(C.c_double), C.c_int32))]


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)

# name -> (restype, argtypes); exactly the entry points of include/hcs.h
SIGNATURES = {
    "hcs_create": (C.c_int, [C.POINTER(_P), C.c_int]),
    "hcs_destroy": (C.c_int, [_P]),
    "hcs_last_error": (C.c_char_p, [_P]),
    "hcs_synchronize": (C.c_int, [_P]),
    "hcs_stream": (_P, [_P]),
    "hcs_add_soft_mesh": (C.c_int, [_P, _D, C.c_int32, _I32, C.c_int32, _D, C.POINTER(Props), C.c_int32, _I32]),
    "hcs_add_rigid_mesh": (C.c_int, [_P, _D, C.c_int32, _I32, C.c_int32, C.POINTER(Props), C.c_int32, _I32]),
    "hcs_add_rigid_plane": (C.c_int, [_P, C.POINTER(Props), C.c_int32, _I32]),
    "hcs_replace_soft_mesh": (C.c_int, [_P, C.c_int32, _D, C.c_int32, _I32, C.c_int32, _D]),
    "hcs_replace_rigid_mesh": (C.c_int, [_P, C.c_int32, _D, C.c_int32, _I32, C.c_int32]),
    "hcs_mesh_info": (C.c_int, [_P, C.c_int32, _I64]),
    "hcs_mesh_tet_records": (C.c_int, [_P, C.c_int32, _D]),
    "hcs_contact_step": (C.c_int, [_P, C.c_int32, C.POINTER(PairIn), C.POINTER(StepOpts), C.POINTER(PairOut),
                                   C.POINTER(Faces)]),
    "hcs_contact_step_noalloc": (C.c_int, [_P, C.c_int32, C.POINTER(PairIn), C.POINTER(StepOpts),
                                           C.POINTER(PairOut)]),
    "hcs_candidates": (C.c_int, [_P, C.POINTER(PairIn), _I64, C.c_int64, _I64]),
    "hcs_taxel_dims": (C.c_int, [C.POINTER(TaxelGrid), _I32, _I32]),
    "hcs_taxels": (C.c_int, [_P, C.c_int32, C.POINTER(TaxelGrid), _D]),
    "hcs_taxel_sensor_update": (C.c_int, [_P, C.POINTER(TaxelSensor), _D, _I64]),
    "hcs_curved_sensor_update": (C.c_int, [_P, C.POINTER(CurvedSensor), _D]),
    "hcs_batch_create": (C.c_int, [_P, C.c_int32, _I32, C.POINTER(StepOpts), C.POINTER(_P)]),
    "hcs_batch_destroy": (C.c_int, [_P]),
    "hcs_batch_set_state": (C.c_int, [_P, _D]),
    "hcs_batch_step": (C.c_int, [_P]),
    "hcs_batch_get": (C.c_int, [_P, C.POINTER(PairOut)]),
    "hcs_batch_tune": (C.c_int, [_P, C.c_int32, _D]),
    "hcs_batch_config": (C.c_int, [_P, _I32]),
    "hcs_batch_set_config": (C.c_int, [_P, _I32]),
    "hcs_batch_kernel_times": (C.c_int, [_P, C.c_int32, _D]),
    "hcs_batch_time_steps": (C.c_int, [_P, C.c_int32, _D]),
    "hcs_batch_counts": (C.c_int, [_P, _I64]),
    "hcs_batch_set_sensors": (C.c_int, [_P, C.c_int32, _I32, C.POINTER(TaxelGrid)]),
    "hcs_batch_taxel_count": (C.c_int, [_P, _I64]),
    "hcs_batch_get_taxels": (C.c_int, [_P, _D]),
    "hcs_batch_item_stats": (C.c_int, [_P, _I64, C.c_int64, _I64]),
    "hcs_contact_jacobian": (C.c_int, [_P, C.c_int32, C.POINTER(PairIn), C.POINTER(StepOpts), _D]),
    "hcs_contact_backward": (C.c_int, [_P, C.c_int32, C.POINTER(PairIn), C.POINTER(StepOpts), _D, _D]),
    "hcs_taxel_jacobian": (C.c_int, [_P, C.POINTER(PairIn), C.POINTER(StepOpts), C.POINTER(TaxelGrid), _D]),
    "hcs_gen_soft_size": (C.c_int, [C.c_int32, C.c_int32, _I32, _I32]),
    "hcs_gen_soft": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double,
                               _D, _I32, _D]),
    "hcs_gen_rigid_size": (C.c_int, [C.c_int32, C.c_int32, _I32, _I32]),
    "hcs_gen_rigid": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, _D, _I32]),
    "hcs_gen_soft_from_surface": (C.c_int, [_D, C.c_int32, _I32, C.c_int32, C.c_double, _D, _I32, _D]),
    "hcs_gen_geom_size": (C.c_int, [C.c_int32, C.c_int32, _D, C.c_double, _I32, _I32]),
    "hcs_gen_geom": (C.c_int, [C.c_int32, C.c_int32, _D, C.c_double, C.c_double, _D, _I32, _D]),
    # include/hcs_mujoco.h
    "hcs_mj_create": (C.c_int, [_P, C.c_int32, C.POINTER(StepOpts), C.POINTER(_P)]),
    "hcs_mj_destroy": (C.c_int, [_P]),
    "hcs_mj_bind_geom": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "hcs_mj_collision": (C.c_int, [_P, C.POINTER(MjData), C.c_int32, C.c_int32, _I32, _I32]),
    "hcs_mj_passive": (C.c_int, [_P, C.POINTER(MjData), C.POINTER(PairOut), _I32, C.c_int32, _I32, C.POINTER(Faces)]),
}

_lib = None


def lib():
    """The loaded libhcs.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libhcs.so not built (run python -m mujoco_contact_surfaces_amd.build): " + LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def dptr(a):
    return a.ctypes.data_as(_D)


def i32ptr(a):
    return a.ctypes.data_as(_I32)


def i64ptr(a):
    return a.ctypes.data_as(_I64)


def as_f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def as_i32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.int32)
    if shape is not None:
        a = a.reshape(shape)
    return a
