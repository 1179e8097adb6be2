"""ctypes binding of libcgs.so (include/cgs.h).

The library is loaded from the package directory; a missing library or a
missing device raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CGS_LIBRARY selects another build of the library (e.g. a tuning variant).
LIB_PATH = os.environ.get("CGS_LIBRARY") or os.path.join(HERE, "libcgs.so")

_P = C.c_void_p
_U32, _I32, _U64, _F32, _SZ = C.c_uint32, C.c_int32, C.c_uint64, C.c_float, C.c_size_t

CGS_OK = 0
STATUS = {
    0: "CGS_OK", 1: "CGS_ERR_INVALID", 2: "CGS_ERR_HIP", 3: "CGS_ERR_HASH_FAILED",
    4: "CGS_ERR_HASH_DUPLICATE", 5: "CGS_ERR_BUCKET_OVERFLOW", 6: "CGS_ERR_BUSY",
    7: "CGS_ERR_EMPTY_SPLIT", 8: "CGS_ERR_STACK_OVERFLOW", 9: "CGS_ERR_NO_DEVICE",
    10: "CGS_ERR_OUT_OF_MEMORY",
}
EMPTY = 0xFFFFFFFF
PATCH_OVERFLOW = 0xFFFFFFFE

# Every symbol include/cgs.h declares (checked by tests/test_abi.py).
EXPORTS = [
    "cgs_abi_version", "cgs_status_string", "cgs_last_error", "cgs_open", "cgs_close",
    "cgs_bezier_clip", "cgs_subdivision", "cgs_patch_eval", "cgs_patch_subpatch",
    "cgs_patchbvh_create", "cgs_patchbvh_destroy", "cgs_patchbvh_intersect", "cgs_patchbvh_intersect_any", "cgs_bvh_build", "cgs_patch_elements", "cgs_mesh_elements", "cgs_free_host", "cgs_load_be",
    "cgs_dl_read", "cgs_dl_parse", "cgs_dl_free",
    "cgs_phf_displacement_size", "cgs_phf_build", "cgs_phf_build_static", "cgs_phf_fill_table", "cgs_phf_hash",
    "cgs_phf_build_wide", "cgs_phf_fill_table_wide", "cgs_phf_hash_wide", "cgs_scene_bvh_wide",
    "cgs_scene_bvh_culled", "cgs_load_se", "cgs_patch_table_free", "cgs_patch_table_elements",
    "cgs_patch_table_patches", "cgs_patchbvh_create_elements",
    "cgs_scene_create", "cgs_scene_destroy", "cgs_scene_bvh_info", "cgs_scene_bvh_copy", "cgs_scene_culled",
    "cgs_trace_closest", "cgs_trace_any", "cgs_trace_closest_ex", "cgs_trace_any_ex", "cgs_render",
    "cgs_render_running", "cgs_render_sampled", "cgs_shade", "cgs_work_counts",
    "cgs_trace_batch", "cgs_batch_create", "cgs_batch_launch", "cgs_batch_destroy", "cgs_batch_set_config",
    "cgs_batch_get_config", "cgs_batch_autotune", "cgs_batch_set_trace_options", "cgs_batch_get_trace_options",
    "cgs_batch_autotune_accuracy",
    "cgs_server_create", "cgs_server_submit", "cgs_server_submit_shaded", "cgs_server_poll", "cgs_server_wait", "cgs_server_destroy",
    "cgs_malloc", "cgs_free", "cgs_memcpy_h2d", "cgs_memcpy_d2h", "cgs_memset", "cgs_stream_sync",
    "cgs_stream_create", "cgs_stream_destroy",
]


class CgsError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Ray(C.Structure):
    _fields_ = [(n, _F32) for n in ("ox", "oy", "oz", "tmax", "dx", "dy", "dz", "pad")]


class Hit(C.Structure):
    _fields_ = [("t", _F32), ("u", _F32), ("v", _F32), ("patch", _U32)]


class BVHNode(C.Structure):
    _fields_ = [("lmin", _F32 * 3), ("lmax", _F32 * 3), ("rmin", _F32 * 3), ("rmax", _F32 * 3),
                ("left", _U32), ("right", _U32), ("pad", _U32 * 2)]


class BVHElement(C.Structure):
    _fields_ = [("bmin", _F32 * 3), ("bmax", _F32 * 3), ("centroid", _F32 * 3), ("index", _U32)]


class PHF(C.Structure):
    _fields_ = [("A", _U32), ("B", _U32), ("mask", _U32), ("table_size", _U32), ("attempts", _U32),
                ("displacement_size", _U32)]


class SceneDesc(C.Structure):
    _fields_ = [
        ("vertices", _P), ("indices", _P), ("surf_vbase", _P), ("surf_pbase", _P), ("nsurf", _U32),
        ("obj_surface", _P), ("obj_local2world", _P), ("obj_world2local", _P), ("obj_material", _P),
        ("nobj", _U32), ("lights", _P), ("nlight", _U32),
        ("background", _F32 * 3), ("ambient", _F32 * 3), ("hash_seed", _U32)]


class DlScene(C.Structure):
    _fields_ = [("scene", SceneDesc), ("view_to_world", _F32 * 16), ("view_position", _F32 * 3),
                ("object_names", C.POINTER(C.c_char_p)), ("light_names", C.POINTER(C.c_char_p)),
                ("storage", _P)]


class BatchJob(C.Structure):
    _fields_ = [("scene", _P), ("rays", _P), ("n", _U32), ("any", _I32), ("hits", _P), ("objects", _P),
                ("occluded", _P)]


class BatchConfig(C.Structure):
    _fields_ = [("chunk", _U32), ("blocks_per_cu", _U32)]


class ServerOptions(C.Structure):
    _fields_ = [("hbm_budget", _SZ), ("slots", _U32), ("max_batch_rays", _U32)]


class Camera(C.Structure):
    _fields_ = [("to_world", _F32 * 16), ("position", _F32 * 3), ("fovy", _F32), ("flip_y", _I32)]


class RenderOptions(C.Structure):
    _fields_ = [("eps", _F32), ("recursion_depth", _I32), ("frame_stride", _U32)]


class TraceOptions(C.Structure):
    _fields_ = [("method", _U32), ("clip_tol", _F32), ("subdiv_depth", _U32), ("max_clip_iters", _U32),
                ("polish", _U32), ("reserved", _U32 * 3)]


class PatchRef(C.Structure):
    _fields_ = [("offset", _U32), ("matrix_offset", _I32), ("size", _U32), ("type", _U32)]


class PatchTable(C.Structure):
    _fields_ = [("points", C.POINTER(_F32)), ("npoints", _U32), ("indices", C.POINTER(_U32)), ("nindices", _U32),
                ("matrices", C.POINTER(_F32)), ("nmatrices", _U32), ("patches", C.POINTER(PatchRef)),
                ("npatches", _U32)]


class AccuracyLimits(C.Structure):
    _fields_ = [("max_rel_t", _F32), ("max_uv", _F32), ("max_disagree", _F32), ("allow_polish", _U32)]


class AutotuneEntry(C.Structure):
    _fields_ = [("options", TraceOptions), ("ms", _F32), ("max_rel_t", _F32), ("max_uv", _F32), ("disagree", _F32),
                ("admissible", _U32)]


class RayCounts(C.Structure):
    _fields_ = [("primary", _U64), ("shadow", _U64), ("secondary", _U64), ("overflow", _U64)]


_lib = None


def lib():
    """Loads libcgs.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        sig = {
            "cgs_abi_version": ([], C.c_int),
            "cgs_status_string": ([C.c_int], C.c_char_p),
            "cgs_last_error": ([], C.c_char_p),
            "cgs_open": ([C.c_int, C.POINTER(_P)], C.c_int),
            "cgs_close": ([_P], C.c_int),
            "cgs_bezier_clip": ([_P, _P, _P, _U32, _F32, _P, _P], C.c_int),
            "cgs_subdivision": ([_P, _P, _P, _U32, _F32, _P, _P], C.c_int),
            "cgs_patch_eval": ([_P, _U32, _P, _P, _U32, _P, _P], C.c_int),
            "cgs_patch_subpatch": ([_P, _P, _P, _U32, _P, _P], C.c_int),
            "cgs_patchbvh_create": ([_P, _P, _U32, _P, _U32, C.POINTER(_P)], C.c_int),
            "cgs_patchbvh_destroy": ([_P], C.c_int),
            "cgs_patchbvh_intersect": ([_P, _P, _P, _U32, _P, _P], C.c_int),
            "cgs_patchbvh_intersect_any": ([_P, _P, _P, _U32, _P, _P], C.c_int),
            "cgs_bvh_build": ([_P, _U32, _U32, C.POINTER(_P), C.POINTER(_P), C.POINTER(_U32), C.POINTER(_P),
                               C.POINTER(_U32)], C.c_int),
            "cgs_patch_elements": ([_P, _P, _U32, _P], C.c_int),
            "cgs_mesh_elements": ([_P, _U32, _P, _U32, _P], C.c_int),
            "cgs_free_host": ([_P], None),
            "cgs_load_be": ([C.c_char_p, C.c_size_t, C.POINTER(_P), C.POINTER(_U32), C.POINTER(_P), C.POINTER(_U32)],
                            C.c_int),
            "cgs_dl_read": ([C.c_char_p, C.POINTER(C.POINTER(DlScene))], C.c_int),
            "cgs_dl_parse": ([C.c_char_p, C.c_size_t, C.c_char_p, C.c_char_p, C.POINTER(C.POINTER(DlScene))],
                             C.c_int),
            "cgs_dl_free": ([C.POINTER(DlScene)], None),
            "cgs_phf_displacement_size": ([_U32], _U32),
            "cgs_phf_build": ([_P, _P, _U32, _U32, _P, C.POINTER(PHF), _P], C.c_int),
            "cgs_phf_build_static": ([_P, _P, _U32, _U32, _P, C.POINTER(PHF), _P], C.c_int),
            "cgs_phf_fill_table": ([_P, C.POINTER(PHF), _P, _P, _U32, _P, _P], C.c_int),
            "cgs_phf_hash": ([_P, C.POINTER(PHF), _P, _P, _U32, _P, _P], C.c_int),
            "cgs_phf_build_wide": ([_P, _P, _U32, _U32, _P, C.POINTER(PHF), _P], C.c_int),
            "cgs_phf_fill_table_wide": ([_P, C.POINTER(PHF), _P, _P, _U32, _P, _P], C.c_int),
            "cgs_phf_hash_wide": ([_P, C.POINTER(PHF), _P, _P, _U32, _P, _P], C.c_int),
            "cgs_scene_bvh_wide": ([_P, _I32], C.c_int),
            "cgs_scene_bvh_culled": ([_P, _I32], C.c_int),
            "cgs_load_se": ([_P, _SZ, C.POINTER(C.POINTER(PatchTable))], C.c_int),
            "cgs_patch_table_free": ([C.POINTER(PatchTable)], None),
            "cgs_patch_table_elements": ([C.POINTER(PatchTable), _P], C.c_int),
            "cgs_patch_table_patches": ([C.POINTER(PatchTable), _P], C.c_int),
            "cgs_patchbvh_create_elements": ([_P, _P, _U32, _P, _U32, _P, C.POINTER(_P)], C.c_int),
            "cgs_scene_create": ([_P, C.POINTER(SceneDesc), C.POINTER(_P)], C.c_int),
            "cgs_scene_destroy": ([_P], C.c_int),
            "cgs_scene_bvh_info": ([_P, _I32, C.POINTER(_U32), C.POINTER(_U32), C.POINTER(PHF)], C.c_int),
            "cgs_scene_bvh_copy": ([_P, _I32, _P, _P, _P, _P, _P], C.c_int),
            "cgs_scene_culled": ([_P], C.c_int),
            "cgs_trace_closest": ([_P, _P, _P, _U32, _P, _P, _P], C.c_int),
            "cgs_trace_any": ([_P, _P, _P, _U32, _P, _P], C.c_int),
            "cgs_trace_closest_ex": ([_P, _P, _P, _U32, _P, _P, C.POINTER(TraceOptions), _P], C.c_int),
            "cgs_trace_any_ex": ([_P, _P, _P, _U32, _P, C.POINTER(TraceOptions), _P], C.c_int),
            "cgs_render": ([_P, _P, C.POINTER(Camera), C.POINTER(RenderOptions), _U32, _U32, _P, _P, _P, _P],
                           C.c_int),
            "cgs_render_running": ([_P], C.c_int),
            "cgs_render_sampled": ([_P, _P, C.POINTER(Camera), C.POINTER(RenderOptions), _U32, _U32, _U32, _U32, _P,
                                    _P, _P, _P], C.c_int),
            "cgs_shade": ([_P, _P, _P, _U32, C.POINTER(RenderOptions), _P, _P, _P, _P, _P, _P], C.c_int),
            "cgs_work_counts": ([_P, _P, _U32, C.c_int], C.c_int),
            "cgs_trace_batch": ([_P, C.POINTER(BatchJob), _U32, _P], C.c_int),
            "cgs_batch_create": ([_P, C.POINTER(BatchJob), _U32, C.POINTER(_P)], C.c_int),
            "cgs_batch_launch": ([_P, _P], C.c_int),
            "cgs_batch_destroy": ([_P], C.c_int),
            "cgs_batch_set_config": ([_P, C.POINTER(BatchConfig)], C.c_int),
            "cgs_batch_get_config": ([_P, C.POINTER(BatchConfig)], C.c_int),
            "cgs_batch_autotune": ([_P, C.POINTER(BatchJob), _U32, C.POINTER(BatchConfig)], C.c_int),
            "cgs_batch_set_trace_options": ([_P, C.POINTER(TraceOptions)], C.c_int),
            "cgs_batch_get_trace_options": ([_P, C.POINTER(TraceOptions)], C.c_int),
            "cgs_batch_autotune_accuracy": ([_P, C.POINTER(BatchJob), _U32, C.POINTER(AccuracyLimits),
                                             C.POINTER(TraceOptions), C.POINTER(AutotuneEntry), _U32,
                                             C.POINTER(_U32)], C.c_int),
            "cgs_server_create": ([_P, C.POINTER(_P), _U32, C.POINTER(ServerOptions), C.POINTER(_P)], C.c_int),
            "cgs_server_submit": ([_P, _U32, _P, _U32, _I32, _P, _P, _P, C.POINTER(_U64)], C.c_int),
            "cgs_server_submit_shaded": ([_P, _U32, _P, _U32, C.POINTER(RenderOptions), _P, _P, _P, _P,
                                          C.POINTER(_U64)], C.c_int),
            "cgs_server_poll": ([_P, _U64, C.POINTER(_U64)], C.c_int),
            "cgs_server_wait": ([_P, _U64, C.POINTER(_U64)], C.c_int),
            "cgs_server_destroy": ([_P], C.c_int),
            "cgs_malloc": ([_P, _SZ, C.POINTER(_P)], C.c_int),
            "cgs_free": ([_P, _P], C.c_int),
            "cgs_memcpy_h2d": ([_P, _P, _P, _SZ, _P], C.c_int),
            "cgs_memcpy_d2h": ([_P, _P, _P, _SZ, _P], C.c_int),
            "cgs_memset": ([_P, _P, C.c_int, _SZ, _P], C.c_int),
            "cgs_stream_sync": ([_P, _P], C.c_int),
            "cgs_stream_create": ([_P, C.POINTER(_P)], C.c_int),
            "cgs_stream_destroy": ([_P, _P], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(status: int, what: str = "") -> None:
    if status != CGS_OK:
        msg = lib().cgs_last_error().decode(errors="replace")
        raise CgsError(status, f"{what}: {msg}" if what else msg)


def ptr(a) -> int | None:
    """Address of a numpy array, a torch tensor, a DeviceArray or an int."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    if hasattr(a, "dptr"):
        return a.dptr
    raise TypeError(f"cannot take the address of {type(a)}")
