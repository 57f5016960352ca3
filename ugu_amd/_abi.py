"""ctypes binding of the C ABI in include/ugu_hip.h (libugu_hip.so).

The library is built in-tree (``ugu_amd/libugu_hip.so``) by
``__graft_entry__.build()`` / ``make -C ugu_amd/csrc``. Importing this module
on a machine without the built library raises immediately: there is no CPU
fallback for the rendering path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libugu_hip.so")

# status codes of include/ugu_hip.h (enum UguHipStatus) beyond the validation codes
UGU_HIP_ERR_INVALID_ARGUMENT = -20
UGU_HIP_ERR_HIP = -21
UGU_HIP_ERR_OUT_OF_MEMORY = -22
UGU_HIP_MEMORY_HOST = 0
UGU_HIP_MEMORY_DEVICE = 1

# Error codes (include/ugu_hip.h); -1..-10 are the reference's LOGE branches.
OK = 0
ERR_NO_CAMERA = -1
ERR_MESH_NOT_PREPARED = -2
ERR_CULLING_NO_FACE_NORMAL = -3
ERR_TEXTURE_EMPTY = -4
ERR_TEXTURE_NTH_EMPTY = -5
ERR_VERTEX_COLOR_EMPTY = -6
ERR_FACE_NORMAL_EMPTY = -7
ERR_VERTEX_NORMAL_EMPTY = -8
ERR_ALL_NULL = -9
ERR_DEPTH_NULL = -10
ERR_INVALID_ARGUMENT = -20
ERR_HIP = -21
ERR_OUT_OF_MEMORY = -22
ERR_NO_MESH = -23

_f = ctypes.c_float
_i = ctypes.c_int32


class UguTexture(ctypes.Structure):
    _fields_ = [("width", _i), ("height", _i), ("data", ctypes.c_void_p)]


class UguMeshDesc(ctypes.Structure):
    _fields_ = [
        ("num_vertices", _i), ("vertices", ctypes.c_void_p),
        ("num_faces", _i), ("faces", ctypes.c_void_p),
        ("face_normals", ctypes.c_void_p),
        ("num_normals", _i), ("normals", ctypes.c_void_p),
        ("normal_indices", ctypes.c_void_p),
        ("num_vertex_colors", _i), ("vertex_colors", ctypes.c_void_p),
        ("num_uv", _i), ("uv", ctypes.c_void_p), ("uv_indices", ctypes.c_void_p),
        ("material_ids", ctypes.c_void_p),
        ("num_materials", _i), ("textures", ctypes.POINTER(UguTexture)),
    ]


class UguCamera(ctypes.Structure):
    _fields_ = [
        ("width", _i), ("height", _i),
        ("fx", _f), ("fy", _f), ("cx", _f), ("cy", _f),
        ("w2c_R", _f * 9), ("w2c_t", _f * 3), ("c2w_R", _f * 9), ("c2w_t", _f * 3),
    ]


class UguOptions(ctypes.Structure):
    _fields_ = [
        ("diffuse_color", _i), ("interp", _i), ("shading_normal", _i),
        ("diffuse_shading", _i), ("depth_scale", _f), ("backface_culling", _i),
        ("oren_nayar_sigma", _f),
    ]


class UguOutputs(ctypes.Structure):
    _fields_ = [
        ("color", ctypes.c_void_p), ("depth", ctypes.c_void_p),
        ("normal", ctypes.c_void_p), ("mask", ctypes.c_void_p),
        ("face_id", ctypes.c_void_p),
    ]


class UguGBuffer(ctypes.Structure):
    _fields_ = [
        ("bary", ctypes.c_void_p), ("uv", ctypes.c_void_p), ("pos_wld", ctypes.c_void_p),
        ("pos_cam", ctypes.c_void_p), ("normal_wld", ctypes.c_void_p),
        ("geo_id", ctypes.c_void_p),
    ]


class UguVisibilityOptions(ctypes.Structure):
    _fields_ = [("backface_culling_angle_rad_th", _f), ("use_mask", _i), ("use_depth", _i),
                ("interp", _i), ("max_viewing_angle", _f),
                ("max_distance_from_camera_to_vertex", _f), ("max_depth_difference", _f),
                ("eps", _f)]


class UguKeyframe(ctypes.Structure):
    _fields_ = [("camera", UguCamera), ("width", _i), ("height", _i),
                ("color", ctypes.c_void_p), ("depth", ctypes.c_void_p), ("mask", ctypes.c_void_p)]


# UguVertexVisibility as a numpy record
VIS_DTYPE = np.dtype([("visible", np.int32), ("px", np.float32), ("py", np.float32),
                      ("angle", np.float32), ("dist", np.float32), ("c0", np.float32),
                      ("c1", np.float32), ("c2", np.float32), ("intensity", np.float32)])


class UguTexFace(ctypes.Structure):
    _fields_ = [("keyframe", _i), ("src", _f * 6), ("dst", _f * 6)]


class UguTiling(ctypes.Structure):
    _fields_ = [("tile_w", _i), ("tile_h", _i), ("bin_capacity", _i), ("persistent", _i)]


class UguTimings(ctypes.Structure):
    _fields_ = [
        ("clear_ms", _f), ("setup_bin_ms", _f),
        ("raster_ms", _f), ("total_ms", _f), ("max_bin_count", _i), ("retries", _i),
        ("bwd_sort_ms", _f), ("bwd_items_ms", _f), ("bwd_items", _i),
    ]


class UguDiffParams(ctypes.Structure):
    _fields_ = [
        ("vertices", ctypes.c_void_p), ("vertex_colors", ctypes.c_void_p),
        ("texture", ctypes.c_void_p), ("tex_width", _i), ("tex_height", _i),
        ("backface_culling", _i),
    ]


# Every symbol the header declares, with its ctypes signature.
_P = ctypes.c_void_p
SIGNATURES = {
    "ugu_hip_abi_version": (ctypes.c_int, []),
    "ugu_hip_create": (_P, [ctypes.c_int]),
    "ugu_hip_destroy": (None, [_P]),
    "ugu_hip_last_error": (ctypes.c_char_p, [_P]),
    "ugu_hip_set_stream": (ctypes.c_int, [_P, _P]),
    "ugu_hip_prepare_mesh": (ctypes.c_int, [_P, ctypes.POINTER(UguMeshDesc)]),
    "ugu_hip_render": (ctypes.c_int, [_P, ctypes.POINTER(UguCamera), ctypes.POINTER(UguOptions),
                                      ctypes.POINTER(UguOutputs), ctypes.c_int]),
    "ugu_hip_render_batch": (ctypes.c_int, [_P, ctypes.POINTER(UguCamera), ctypes.c_int,
                                            ctypes.POINTER(UguOptions), ctypes.POINTER(UguOutputs),
                                            ctypes.c_int]),
    "ugu_hip_render_batch_eager": (ctypes.c_int, [_P, ctypes.POINTER(UguCamera), ctypes.c_int,
                                                  ctypes.POINTER(UguOptions),
                                                  ctypes.POINTER(UguOutputs), ctypes.c_int]),
    "ugu_hip_render_batch_noalloc": (ctypes.c_int, [_P, ctypes.POINTER(UguCamera), ctypes.c_int,
                                                    ctypes.POINTER(UguOptions),
                                                    ctypes.POINTER(UguOutputs)]),
    "ugu_hip_render_gbuffer": (ctypes.c_int, [_P, ctypes.POINTER(UguCamera), ctypes.c_int,
                                              ctypes.POINTER(UguOptions),
                                              ctypes.POINTER(UguOutputs),
                                              ctypes.POINTER(UguGBuffer), ctypes.c_int]),
    "ugu_hip_prepare_raytracer": (ctypes.c_int, [_P]),
    "ugu_hip_raytrace": (ctypes.c_int, [_P, ctypes.POINTER(UguCamera), ctypes.c_int,
                                        ctypes.POINTER(UguOptions), ctypes.POINTER(UguOutputs),
                                        ctypes.c_int]),
    "ugu_hip_visibility_test": (ctypes.c_int, [_P, ctypes.POINTER(UguKeyframe),
                                               ctypes.POINTER(UguVisibilityOptions), ctypes.c_int,
                                               _P, _P, _P, _P]),
    "ugu_hip_texture_map": (ctypes.c_int, [_P, ctypes.POINTER(UguTexture), ctypes.c_int,
                                           ctypes.POINTER(UguTexFace), ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int, _P, _P]),
    "ugu_hip_set_tiling": (ctypes.c_int, [_P, ctypes.POINTER(UguTiling)]),
    "ugu_hip_get_tiling": (ctypes.c_int, [_P, ctypes.POINTER(UguTiling)]),
    "ugu_hip_autotune": (ctypes.c_int, [_P, ctypes.POINTER(UguCamera), ctypes.c_int,
                                        ctypes.POINTER(UguOptions), ctypes.c_int,
                                        ctypes.POINTER(UguTiling)]),
    "ugu_hip_autotune_cached": (ctypes.c_int, [_P, ctypes.POINTER(UguCamera), ctypes.c_int,
                                               ctypes.POINTER(UguOptions), ctypes.c_int,
                                               ctypes.c_char_p, ctypes.POINTER(UguTiling),
                                               ctypes.POINTER(ctypes.c_int)]),
    "ugu_hip_enable_profiling": (ctypes.c_int, [_P, ctypes.c_int]),
    "ugu_hip_get_timings": (ctypes.c_int, [_P, ctypes.POINTER(UguTimings)]),
    "ugu_hip_depth_to_u16": (ctypes.c_int, [_P, _P, ctypes.c_int64, _P, ctypes.c_int]),
    "ugu_hip_diff_forward": (ctypes.c_int, [_P, ctypes.POINTER(UguCamera), ctypes.c_int,
                                            ctypes.POINTER(UguDiffParams),
                                            ctypes.POINTER(_P), ctypes.POINTER(_P)]),
    "ugu_hip_diff_backward": (ctypes.c_int, [_P, ctypes.POINTER(UguCamera), ctypes.c_int,
                                             ctypes.POINTER(UguDiffParams), ctypes.POINTER(_P),
                                             ctypes.POINTER(_P), _P, _P, _P]),
}

_lib = None


def load_library(path=None):
    """Load libugu_hip.so and attach signatures; raises if it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} is not built: run `make -C ugu_amd/csrc` (or __graft_entry__.build()); "
            "ugu_amd has no CPU fallback")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class HipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ugu_hip error {code}: {msg.strip()}")
        self.code = code
