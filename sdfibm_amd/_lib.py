"""ctypes binding of ``libsdfibm_hip.so`` (the C ABI in include/sdfibm_hip.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback for the coupling step.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsdfibm_hip.so")

# Every symbol include/sdfibm_hip.h declares.
EXPORTS = (
    "sdfibm_create", "sdfibm_destroy", "sdfibm_set_solids", "sdfibm_set_velocity",
    "sdfibm_interact", "sdfibm_get_fields", "sdfibm_field_pointers", "sdfibm_candidates",
    "sdfibm_last_timing", "sdfibm_set_phase_timing", "sdfibm_set_stream", "sdfibm_last_error", "sdfibm_status_string",
    "sdfibm_abi_version", "sdfibm_update_velocity", "sdfibm_fix_internal", "sdfibm_get_velocity",
    "sdfibm_set_dynamics", "sdfibm_evolve", "sdfibm_get_solids", "sdfibm_set_contact", "sdfibm_run_steps", "sdfibm_pair_counts",
    "sdfibm_surrogate_train", "sdfibm_surrogate_samples", "sdfibm_surrogate_eval", "sdfibm_surrogate_grad",
    "sdfibm_set_surrogate", "sdfibm_surrogate_count", "sdfibm_buffer_capacity", "sdfibm_set_csg_shapes",
    "sdfibm_set_velocity_cells_noalloc", "sdfibm_set_tuning", "sdfibm_autotune",
    "sdfibm_interact_steps", "sdfibm_create_polymesh",
)

SDFIBM_OK = 0
NOALLOC_MAX_CELLS = 1048576  # SDFIBM_NOALLOC_MAX_CELLS
STATUS = {0: "OK", -1: "ERR_ARG", -2: "ERR_HIP", -3: "ERR_NOMEM", -4: "ERR_SHAPE", -5: "ERR_TILE",
          -6: "ERR_STATE", -7: "ERR_MESH", -8: "ERR_CAPACITY", -9: "ERR_NODEVICE"}
MESH_RECTILINEAR, MESH_STRUCTURED, MESH_POLY = 0, 1, 2
FP32, FP64 = 0, 1


class MeshDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("xs", C.c_void_p), ("ys", C.c_void_p), ("zs", C.c_void_p),
                ("points", C.c_void_p), ("cell_centres", C.c_void_p), ("cell_volumes", C.c_void_p),
                ("face_centres", C.c_void_p * 3), ("face_areas", C.c_void_p * 3)]


class PolyMeshDesc(C.Structure):
    _fields_ = [("n_points", C.c_int64), ("n_faces", C.c_int64), ("n_internal_faces", C.c_int64),
                ("n_cells", C.c_int64), ("points", C.c_void_p), ("face_offsets", C.c_void_p),
                ("face_points", C.c_void_p), ("owner", C.c_void_p), ("neighbour", C.c_void_p),
                ("cell_centres", C.c_void_p), ("cell_volumes", C.c_void_p), ("face_centres", C.c_void_p),
                ("face_areas", C.c_void_p)]


class Config(C.Structure):
    _fields_ = [("two_d", C.c_int32), ("precision", C.c_int32), ("rho_f", C.c_double),
                ("reserved", C.c_int64 * 4)]


class SurrogateTrainConfig(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n_train", C.c_int64), ("n_val", C.c_int64), ("batch", C.c_int64),
                ("epochs", C.c_int32), ("reserved", C.c_int32), ("lr", C.c_double)]


class SdfibmError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -m sdfibm_amd.build` "
                          "(the coupling step has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    P, i64, i32, d = C.c_void_p, C.c_int64, C.c_int32, C.c_double
    sig = {
        "sdfibm_create": (i32, [P, P, P]),
        "sdfibm_destroy": (i32, [P]),
        "sdfibm_set_solids": (i32, [P, P, i64]),
        "sdfibm_set_velocity": (i32, [P, P]),
        "sdfibm_interact": (i32, [P, d, d, P]),
        "sdfibm_get_fields": (i32, [P, P, P, P, P]),
        "sdfibm_field_pointers": (i32, [P, P, P, P, P, P]),
        "sdfibm_candidates": (i32, [P, i64, P, P, P, i64]),
        "sdfibm_last_timing": (i32, [P, P, i32]),
        "sdfibm_set_phase_timing": (i32, [P, i32]),
        "sdfibm_set_stream": (i32, [P, P]),
        "sdfibm_last_error": (C.c_char_p, [P]),
        "sdfibm_status_string": (C.c_char_p, [i32]),
        "sdfibm_abi_version": (i32, []),
        "sdfibm_update_velocity": (i32, [P, d]),
        "sdfibm_fix_internal": (i32, [P]),
        "sdfibm_get_velocity": (i32, [P, P]),
        "sdfibm_set_dynamics": (i32, [P, P, i64]),
        "sdfibm_evolve": (i32, [P, d, d, P]),
        "sdfibm_get_solids": (i32, [P, P, i64]),
        "sdfibm_set_contact": (i32, [P, d, d]),
        "sdfibm_run_steps": (i32, [P, d, d, i64, P, P]),
        "sdfibm_pair_counts": (i32, [P, P]),
        "sdfibm_surrogate_train": (i32, [P, P, P]),
        "sdfibm_surrogate_samples": (i32, [C.c_uint64, i64, P, P]),
        "sdfibm_surrogate_eval": (i32, [P, P, i64, P]),
        "sdfibm_surrogate_grad": (i32, [P, P, P, i64, P, P]),
        "sdfibm_set_surrogate": (i32, [P, P, d, d, d]),
        "sdfibm_surrogate_count": (i32, [P, P]),
        "sdfibm_buffer_capacity": (i32, [P, P, i32]),
        "sdfibm_set_csg_shapes": (i32, [P, P, i64]),
        "sdfibm_set_velocity_cells_noalloc": (i32, [P, i64, i64, P]),
        "sdfibm_set_tuning": (i32, [P, i32, i32, i32]),
        "sdfibm_autotune": (i32, [P, d, C.c_char_p, C.c_char_p, i64]),
        "sdfibm_interact_steps": (i32, [P, d, d, i64, P]),
        "sdfibm_create_polymesh": (i32, [P, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(status: int, ctx=None):
    if status != SDFIBM_OK:
        L = load()
        msg = L.sdfibm_last_error(ctx)
        raise SdfibmError(status, (msg or b"").decode())
    return status
