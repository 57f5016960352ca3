"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY. Imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg -- as the checker or the CPU baseline, never by
the product package ugu_amd.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# UGU_ORACLE_LIB selects another build of the same source (the sanitizer
# build of tests/sanitize); the default is the one `make -C oracle` writes.
LIB_PATH = os.environ.get("UGU_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")

_P = ctypes.c_void_p
_i = ctypes.c_int32
_f = ctypes.c_float


class OracleMesh(ctypes.Structure):
    _fields_ = [
        ("num_vertices", _i), ("vertices", _P),
        ("num_faces", _i), ("faces", _P),
        ("face_normals", _P),
        ("num_normals", _i), ("normals", _P), ("normal_indices", _P),
        ("num_vertex_colors", _i), ("vertex_colors", _P),
        ("num_uv", _i), ("uv", _P), ("uv_indices", _P),
        ("material_ids", _P),
        ("num_materials", _i), ("tex_width", _P), ("tex_height", _P), ("tex_data", _P),
    ]


class OracleCamera(ctypes.Structure):
    _fields_ = [("width", _i), ("height", _i), ("fx", _f), ("fy", _f), ("cx", _f), ("cy", _f),
                ("w2c_R", _f * 9), ("w2c_t", _f * 3), ("c2w_R", _f * 9)]


class OracleGBuf(ctypes.Structure):
    _fields_ = [("bary", _P), ("uv", _P), ("pos_wld", _P), ("pos_cam", _P),
                ("normal_wld", _P), ("geo_id", _P)]


class OracleOptions(ctypes.Structure):
    _fields_ = [("diffuse_color", _i), ("interp", _i), ("shading_normal", _i),
                ("diffuse_shading", _i), ("depth_scale", _f), ("backface_culling", _i),
                ("oren_nayar_sigma", _f)]


_lib = None
_lib_path = LIB_PATH


def use_library(path=None):
    """Switch to another build of the oracle (the planted-mistake builds of the
    mutation tests); None restores oracle/liboracle.so."""
    global _lib, _lib_path
    _lib = None
    _lib_path = path or LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_lib_path):
            raise RuntimeError(f"{_lib_path} missing: run `make -C oracle`")
        L = ctypes.CDLL(_lib_path)
        L.oracle_render.restype = ctypes.c_int
        L.oracle_render.argtypes = [ctypes.POINTER(OracleMesh), ctypes.POINTER(OracleCamera), _P,
                                    ctypes.POINTER(OracleOptions), _P, _P, _P, _P, _P,
                                    ctypes.POINTER(ctypes.c_double)]
        L.oracle_render_gbuffer.restype = ctypes.c_int
        L.oracle_render_gbuffer.argtypes = [ctypes.POINTER(OracleMesh), ctypes.POINTER(OracleCamera),
                                            _P, ctypes.POINTER(OracleOptions), _P, _P, _P, _P, _P,
                                            ctypes.POINTER(OracleGBuf),
                                            ctypes.POINTER(ctypes.c_double)]
        L.oracle_raytrace.restype = ctypes.c_int
        L.oracle_raytrace.argtypes = [ctypes.POINTER(OracleMesh), ctypes.POINTER(OracleCamera), _P,
                                      ctypes.POINTER(OracleOptions), _P, _P, _P, _P, _P,
                                      ctypes.POINTER(ctypes.c_double)]
        L.oracle_depth_to_u16.restype = None
        L.oracle_depth_to_u16.argtypes = [_P, ctypes.c_int64, _P]
        L.oracle_edge_function.restype = ctypes.c_float
        L.oracle_edge_function.argtypes = [_P, _P, _P]
        L.oracle_ray_w.restype = None
        L.oracle_ray_w.argtypes = [ctypes.POINTER(OracleCamera), _f, _f, _P]
        L.oracle_num_threads.restype = ctypes.c_int
        L.oracle_num_threads.argtypes = []
        if hasattr(L, "oracle_diff_forward_f64"):
            L.oracle_diff_forward_f64.restype = ctypes.c_int
            L.oracle_diff_backward_f64.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_P)


def camera_struct(w2c_R, w2c_t, c2w_R, width, height, fx, fy, cx, cy):
    c = OracleCamera()
    c.width, c.height = int(width), int(height)
    c.fx, c.fy, c.cx, c.cy = float(fx), float(fy), float(cx), float(cy)
    for i, x in enumerate(np.asarray(w2c_R, np.float32).reshape(-1)):
        c.w2c_R[i] = float(x)
    for i, x in enumerate(np.asarray(w2c_t, np.float32).reshape(-1)):
        c.w2c_t[i] = float(x)
    for i, x in enumerate(np.asarray(c2w_R, np.float32).reshape(-1)):
        c.c2w_R[i] = float(x)
    return c


def camera_from_abi(abi_cam):
    """Same float values as an ugu_amd._abi.UguCamera."""
    c = OracleCamera()
    c.width, c.height = abi_cam.width, abi_cam.height
    c.fx, c.fy, c.cx, c.cy = abi_cam.fx, abi_cam.fy, abi_cam.cx, abi_cam.cy
    for i in range(9):
        c.w2c_R[i] = abi_cam.w2c_R[i]
        c.c2w_R[i] = abi_cam.c2w_R[i]
    for i in range(3):
        c.w2c_t[i] = abi_cam.w2c_t[i]
    return c


class _MeshHold:
    def __init__(self, mesh):
        self.keep = []

        def arr(a, dt):
            if a is None or len(a) == 0:
                return None
            a = np.ascontiguousarray(a, dt)
            self.keep.append(a)
            return a

        m = OracleMesh()
        v = arr(mesh.vertices, np.float32)
        f = arr(mesh.vertex_indices, np.int32)
        m.num_vertices = 0 if v is None else len(v)
        m.vertices = _p(v)
        m.num_faces = 0 if f is None else len(f)
        m.faces = _p(f)
        m.face_normals = _p(arr(mesh.face_normals, np.float32))
        n = arr(mesh.normals, np.float32)
        m.num_normals = 0 if n is None else len(n)
        m.normals = _p(n)
        m.normal_indices = _p(arr(mesh.normal_indices, np.int32))
        vc = arr(mesh.vertex_colors, np.float32)
        m.num_vertex_colors = 0 if vc is None else len(vc)
        m.vertex_colors = _p(vc)
        uv = arr(mesh.uv, np.float32)
        m.num_uv = 0 if uv is None else len(uv)
        m.uv = _p(uv)
        m.uv_indices = _p(arr(mesh.uv_indices, np.int32))
        m.material_ids = _p(arr(mesh.material_ids, np.int32))
        nm = len(mesh.materials)
        m.num_materials = nm
        if nm:
            tw = np.array([t.shape[1] if np.size(t) else 0 for t in mesh.materials], np.int32)
            th = np.array([t.shape[0] if np.size(t) else 0 for t in mesh.materials], np.int32)
            texs = [np.ascontiguousarray(t, np.uint8) for t in mesh.materials]
            ptrs = (ctypes.c_void_p * nm)(*[t.ctypes.data if t.size else None for t in texs])
            self.keep += [tw, th, texs, ptrs]
            m.tex_width, m.tex_height = _p(tw), _p(th)
            m.tex_data = ctypes.cast(ptrs, _P)
        self.m = m


def options_struct(o):
    s = OracleOptions()
    s.diffuse_color = int(o.diffuse_color)
    s.interp = int(o.interp)
    s.shading_normal = int(o.shading_normal)
    s.diffuse_shading = int(o.diffuse_shading)
    s.depth_scale = float(o.depth_scale)
    s.backface_culling = 1 if o.backface_culling else 0
    s.oren_nayar_sigma = float(o.oren_nayar_sigma)
    return s


def render(mesh, cam, options, c2w_t=None, outputs=("color", "depth", "normal", "mask", "face_id")):
    """Oracle Render(): returns (rc, dict of numpy outputs, elapsed_ms).
    `cam` is an OracleCamera (see camera_from_abi / camera_struct)."""
    h, w = cam.height, cam.width
    out = {}
    if "color" in outputs:
        out["color"] = np.zeros((h, w, 3), np.uint8)
    if "depth" in outputs:
        out["depth"] = np.zeros((h, w), np.float32)
    if "normal" in outputs:
        out["normal"] = np.zeros((h, w, 3), np.float32)
    if "mask" in outputs:
        out["mask"] = np.zeros((h, w), np.uint8)
    if "face_id" in outputs:
        out["face_id"] = np.zeros((h, w), np.int32)
    hold = _MeshHold(mesh)
    opt = options_struct(options)
    ct = None if c2w_t is None else np.ascontiguousarray(c2w_t, np.float32)
    ms = ctypes.c_double(0.0)
    rc = lib().oracle_render(ctypes.byref(hold.m), ctypes.byref(cam), _p(ct), ctypes.byref(opt),
                             _p(out.get("color")), _p(out.get("depth")), _p(out.get("normal")),
                             _p(out.get("mask")), _p(out.get("face_id")), ctypes.byref(ms))
    return rc, out, ms.value


def raytrace(mesh, cam, options, c2w_t=None, outputs=("color", "depth", "normal", "mask", "face_id")):
    """oracle_raytrace: Raytracer::Render with the exact closest hit (every
    face tested) -> (rc, dict of numpy outputs, elapsed_ms)."""
    h, w = cam.height, cam.width
    shapes = {"color": ((h, w, 3), np.uint8), "depth": ((h, w), np.float32),
              "normal": ((h, w, 3), np.float32), "mask": ((h, w), np.uint8),
              "face_id": ((h, w), np.int32)}
    out = {k: np.zeros(*shapes[k]) for k in outputs}
    hold = _MeshHold(mesh)
    opt = options_struct(options)
    ct = None if c2w_t is None else np.ascontiguousarray(c2w_t, np.float32)
    ms = ctypes.c_double(0.0)
    rc = lib().oracle_raytrace(ctypes.byref(hold.m), ctypes.byref(cam), _p(ct), ctypes.byref(opt),
                               _p(out.get("color")), _p(out.get("depth")), _p(out.get("normal")),
                               _p(out.get("mask")), _p(out.get("face_id")), ctypes.byref(ms))
    return rc, out, ms.value


GBUFFER_PLANES = ("bary", "uv", "pos_wld", "pos_cam", "normal_wld", "geo_id")


def render_gbuffer(mesh, cam, options, c2w_t=None):
    """oracle_render_gbuffer: the five Render outputs and the six G-buffer
    planes (ugu_oracle.cc OracleGBuf) -> (rc, dict, elapsed_ms)."""
    h, w = cam.height, cam.width
    out = {"color": np.zeros((h, w, 3), np.uint8), "depth": np.zeros((h, w), np.float32),
           "normal": np.zeros((h, w, 3), np.float32), "mask": np.zeros((h, w), np.uint8),
           "face_id": np.zeros((h, w), np.int32)}
    for k in GBUFFER_PLANES[:5]:
        out[k] = np.zeros((h, w, 3), np.float32)
    out["geo_id"] = np.zeros((h, w), np.int32)
    g = OracleGBuf(*(_p(out[k]) for k in GBUFFER_PLANES))
    hold = _MeshHold(mesh)
    opt = options_struct(options)
    ct = None if c2w_t is None else np.ascontiguousarray(c2w_t, np.float32)
    ms = ctypes.c_double(0.0)
    rc = lib().oracle_render_gbuffer(ctypes.byref(hold.m), ctypes.byref(cam), _p(ct),
                                     ctypes.byref(opt), _p(out["color"]), _p(out["depth"]),
                                     _p(out["normal"]), _p(out["mask"]), _p(out["face_id"]),
                                     ctypes.byref(g), ctypes.byref(ms))
    return rc, out, ms.value


class OracleVisOptions(ctypes.Structure):
    _fields_ = [("backface_culling_angle_rad_th", _f), ("use_mask", _i), ("use_depth", _i),
                ("interp", _i), ("max_viewing_angle", _f),
                ("max_distance_from_camera_to_vertex", _f), ("max_depth_difference", _f),
                ("eps", _f)]


VIS_DTYPE = np.dtype([("visible", np.int32), ("px", np.float32), ("py", np.float32),
                      ("angle", np.float32), ("dist", np.float32), ("c0", np.float32),
                      ("c1", np.float32), ("c2", np.float32), ("intensity", np.float32)])


def visibility(mesh, cam, c2w_t, color, depth, mask, opt, item_pos, item_vertex, item_normal):
    """oracle_visibility: TestVertex for every item (ugu_oracle.cc); returns a
    structured array (VIS_DTYPE). `opt` is an OracleVisOptions."""
    L = lib()
    L.oracle_visibility.restype = ctypes.c_int
    hold = _MeshHold(mesh)
    h, w = color.shape[:2]
    color = np.ascontiguousarray(color, np.uint8)
    depth = None if depth is None else np.ascontiguousarray(depth, np.float32)
    mask = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    pos = np.ascontiguousarray(item_pos, np.float32)
    vert = np.ascontiguousarray(item_vertex, np.float32)
    nrm = np.ascontiguousarray(item_normal, np.float32)
    ct = np.ascontiguousarray(c2w_t, np.float32)
    out = np.zeros(len(pos), VIS_DTYPE)
    rc = L.oracle_visibility(ctypes.byref(hold.m), ctypes.byref(cam), _p(ct), ctypes.c_int32(w),
                             ctypes.c_int32(h), _p(color), _p(depth), _p(mask), ctypes.byref(opt),
                             ctypes.c_int32(len(pos)), _p(pos), _p(vert), _p(nrm),
                             out.ctypes.data_as(_P))
    assert rc == 0
    return out


class OracleTexFace(ctypes.Structure):
    _fields_ = [("keyframe", _i), ("src", _f * 6), ("dst", _f * 6)]


def texture_map(images, faces, tex_w, tex_h, padding_kernel):
    """oracle_texture_map: `images` = list of (H, W, 3) uint8, `faces` = a
    ctypes array of OracleTexFace -> (texture, mask)."""
    L = lib()
    L.oracle_texture_map.restype = ctypes.c_int
    imgs = [np.ascontiguousarray(im, np.uint8) for im in images]
    iw = np.array([im.shape[1] for im in imgs] or [0], np.int32)
    ih = np.array([im.shape[0] for im in imgs] or [0], np.int32)
    ptrs = (ctypes.c_void_p * max(1, len(imgs)))(*[im.ctypes.data for im in imgs])
    tex = np.zeros((tex_h, tex_w, 3), np.uint8)
    mask = np.zeros((tex_h, tex_w), np.uint8)
    rc = L.oracle_texture_map(_p(iw), _p(ih), ptrs, ctypes.c_int32(len(faces)), faces,
                              ctypes.c_int32(tex_w), ctypes.c_int32(tex_h),
                              ctypes.c_int32(padding_kernel), _p(tex), _p(mask))
    assert rc == 0
    return tex, mask


def depth_to_u16(depth):
    d = np.ascontiguousarray(depth, np.float32)
    out = np.empty(d.shape, np.uint16)
    lib().oracle_depth_to_u16(_p(d), d.size, _p(out))
    return out


def edge_function(a, b, c):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    c = np.asarray(c, np.float32)
    return lib().oracle_edge_function(_p(a), _p(b), _p(c))


def ray_w(cam, x, y):
    out = np.zeros(3, np.float32)
    lib().oracle_ray_w(ctypes.byref(cam), float(x), float(y), _p(out))
    return out


def num_threads():
    return lib().oracle_num_threads()


def _diff_args(cam, mesh, verts, colors, tex):
    L = lib()
    for fn in (L.oracle_diff_forward_f64, L.oracle_diff_backward_f64):
        fn.restype = ctypes.c_int
    f = np.ascontiguousarray(mesh.vertex_indices, np.int32)
    uv = np.ascontiguousarray(mesh.uv, np.float32) if len(mesh.uv) else None
    ui = np.ascontiguousarray(mesh.uv_indices, np.int32) if len(mesh.uv_indices) else None
    v = np.ascontiguousarray(verts, np.float64)
    c = None if colors is None else np.ascontiguousarray(colors, np.float64)
    t = None if tex is None else np.ascontiguousarray(tex, np.float64)
    th, tw = (0, 0) if t is None else t.shape[:2]
    keep = (f, uv, ui, v, c, t)
    args = [ctypes.byref(cam), _p(f), _p(uv), _p(ui), _p(v), _p(c), _p(t),
            ctypes.c_int(tw), ctypes.c_int(th)]
    return keep, args


def diff_forward_f64(cam, mesh, verts, colors, tex, face_id):
    """Float64 image (H, W, 3) of the differentiable model at fixed visibility."""
    keep, args = _diff_args(cam, mesh, verts, colors, tex)
    fid = np.ascontiguousarray(face_id, np.int32)
    img = np.zeros((cam.height, cam.width, 3), np.float64)
    lib().oracle_diff_forward_f64(*args, _p(fid), _p(img))
    return img


def diff_backward_f64(cam, mesh, verts, colors, tex, face_id, grad_image):
    """Float64 gradients (verts, colors, texture) of sum(grad_image * image)."""
    keep, args = _diff_args(cam, mesh, verts, colors, tex)
    fid = np.ascontiguousarray(face_id, np.int32)
    g = np.ascontiguousarray(grad_image, np.float64)
    gv = np.zeros((len(verts), 3), np.float64)
    gc = None if colors is None else np.zeros_like(np.asarray(colors, np.float64))
    gt = None if tex is None else np.zeros_like(np.asarray(tex, np.float64))
    lib().oracle_diff_backward_f64(*args, _p(fid), _p(g), _p(gv), _p(gc), _p(gt))
    return gv, gc, gt
