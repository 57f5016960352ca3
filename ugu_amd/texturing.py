"""Texture mapping (ugu::TextureMapping with kSimpleProjection,
src/texturing/texture_mapper.cc).

``TextureMapping(keyframes, info, mesh, option)`` picks the best keyframe of
every face from its FaceInfo statistics (kMinViewingAngle, kMinDistance or
kMaxArea; texture_mapper.cc:924-967), then builds the texture for the uv type:

* kGenerateSimpleTile (the default, texture_mapper.cc:30-136): the keyframe
  images tiled on a grid of floor(sqrt(n)) columns, each face's uv its chosen
  keyframe's projected corners moved into that keyframe's tile (host only: a
  copy and per-corner arithmetic);
* kUseOriginalMeshUv (texture_mapper.cc:295-345): one record per face -- the
  keyframe's projected triangle and the face's uv triangle in texel
  coordinates (u * tex_w - 0.5, (1 - v) * tex_h - 0.5) -- rasterized into the
  texture on the GPU (``ugu_hip_texture_map``: bilinear samples of the
  keyframe colour, the face with the larger index winning where uv triangles
  overlap, then PaddingSimple). The reference indexes the keyframe list with
  the keyframe id here, so ids must be list positions.

The result becomes the mesh's single material. The packing atlas generators
(kGenerateSimpleTriangles / kGenerateSimpleCharts) are not provided.
"""
import ctypes
import enum
import math
import sys

import numpy as np

from . import _abi
from .visibility import FaceInfo


class ViewSelectionCriteria(enum.IntEnum):
    """ViewSelectionCriteria (visibility_tester.h:22-35); TextureMapping
    accepts kMinViewingAngle, kMinDistance and kMaxArea, as the reference."""
    kMinViewingAngle = 0
    kMinDistance = 1
    kMeanViewingAngle = 2
    kMedianViewingAngle = 3
    kMeanDistance = 4
    kMedianDistance = 5
    kMaxArea = 6
    kMinIntensity = 7
    kMedianIntensity = 8
    kMode = 9
    kModeViewingAngle = 10
    kCustom = 11


class TexturingOutputUvType(enum.IntEnum):
    kGenerateSimpleTile = 0
    kUseOriginalMeshUv = 1
    kGenerateSimpleTriangles = 2
    kGenerateSimpleCharts = 3


class TextureMappingOption:
    """TextureMappingOption (include/ugu/texturing/texture_mapper.h:21-29)."""

    def __init__(self, **kw):
        self.criteria = ViewSelectionCriteria.kMaxArea
        self.uv_type = TexturingOutputUvType.kGenerateSimpleTile
        self.texture_base_name = "ugutex"
        self.tex_w = 1024
        self.tex_h = 1024
        self.padding_kernel = 3
        for k, v in kw.items():
            if not hasattr(self, k):
                raise TypeError(f"unknown option {k}")
            setattr(self, k, v)


def _log(msg):
    sys.stderr.write(msg if msg.endswith("\n") else msg + "\n")


def _restat(fi):
    c = FaceInfo()
    c.visible_keyframes = fi.visible_keyframes
    c.CalcStat()
    return c


def best_keyframes(info, criteria):
    """Per face: the chosen FaceInfoPerKeyframe, or None (no visible keyframe)."""
    out = []
    for fi in info.face_info_list:
        if not fi.visible_keyframes:
            out.append(None)
            continue
        if fi.max_area_index < 0:  # statistics never computed
            fi = _restat(fi)
        idx = {ViewSelectionCriteria.kMinViewingAngle: fi.min_viewing_angle_index,
               ViewSelectionCriteria.kMinDistance: fi.min_distance_index,
               ViewSelectionCriteria.kMaxArea: fi.max_area_index}[criteria]
        out.append(fi.visible_keyframes[idx])
    return out


def face_records(mesh, best, tex_w, tex_h):
    """UguTexFace array: keyframe = the chosen keyframe's id, src = its
    projected triangle, dst = the uv triangle in texel coordinates."""
    n = len(mesh.vertex_indices)
    recs = (_abi.UguTexFace * max(1, n))()
    uv = mesh.uv[mesh.uv_indices].astype(np.float32)           # F x 3 x 2
    dx = (uv[..., 0] * np.float32(tex_w) - np.float32(0.5)).astype(np.float32)
    dy = ((np.float32(1.0) - uv[..., 1]) * np.float32(tex_h) - np.float32(0.5)).astype(np.float32)
    for i in range(n):
        b = best[i]
        r = recs[i]
        if b is None or b.kf_id < 0:
            r.keyframe = -1
            continue
        r.keyframe = int(b.kf_id)
        for j in range(3):
            r.src[2 * j] = float(b.projected_tri[j][0])
            r.src[2 * j + 1] = float(b.projected_tri[j][1])
            r.dst[2 * j] = float(dx[i, j])
            r.dst[2 * j + 1] = float(dy[i, j])
    return recs, n


def texture_map_faces(images, recs, n, tex_w, tex_h, padding_kernel, device=0):
    """ugu_hip_texture_map -> (texture (tex_h, tex_w, 3) uint8, mask)."""
    lib = _abi.load_library()
    ctx = lib.ugu_hip_create(int(device))
    if not ctx:
        raise RuntimeError(f"ugu_hip_create({device}) failed: no usable HIP device")
    try:
        imgs = [np.ascontiguousarray(im, np.uint8) for im in images]
        tex_arr = (_abi.UguTexture * max(1, len(imgs)))()
        for i, im in enumerate(imgs):
            tex_arr[i] = _abi.UguTexture(im.shape[1], im.shape[0], im.ctypes.data)
        tex = np.zeros((tex_h, tex_w, 3), np.uint8)
        mask = np.zeros((tex_h, tex_w), np.uint8)
        rc = lib.ugu_hip_texture_map(ctx, tex_arr, len(imgs), recs, n, tex_w, tex_h,
                                     padding_kernel, tex.ctypes.data, mask.ctypes.data)
        if rc != 0:
            raise _abi.HipError(rc, lib.ugu_hip_last_error(ctx).decode())
        return tex, mask
    finally:
        lib.ugu_hip_destroy(ctx)


def simple_tile(keyframes, best, mesh, option):
    """kGenerateSimpleTile (texture_mapper.cc:30-136): keyframe i at tile
    (i % cols, i // cols) of a grid with cols = floor(sqrt(n)); face corners
    at ((x + 0.5) / W, 1 - (y + 0.5) / H) after the move into the chosen
    keyframe's tile; faces seen by no keyframe get uv 0. The grid has
    ceil(n / cols) rows (the reference's n // cols rows put the last
    keyframes outside its image), and keyframes must share one size."""
    n = len(keyframes)
    h0, w0 = keyframes[0].color.shape[:2]
    if any(k.color.shape[:2] != (h0, w0) for k in keyframes):
        _log("kGenerateSimpleTile needs keyframes of one size")
        return False
    cols = max(1, int(math.sqrt(n)))
    rows = (n + cols - 1) // cols
    tex_w, tex_h = cols * w0, rows * h0
    tex = np.zeros((tex_h, tex_w, 3), np.uint8)
    for i, k in enumerate(keyframes):
        x0, y0 = (i % cols) * w0, (i // cols) * h0
        tex[y0:y0 + h0, x0:x0 + w0] = np.asarray(k.color, np.uint8).reshape(h0, w0, 3)
    position = {}
    for i, k in enumerate(keyframes):
        position.setdefault(k.id, i)
    F = len(best)
    uv = np.zeros((F, 3, 2), np.float32)
    fw, fh = np.float32(tex_w), np.float32(tex_h)
    half, one = np.float32(0.5), np.float32(1.0)
    for f, b in enumerate(best):
        if b is None or b.kf_id < 0:
            continue
        slot = position.get(b.kf_id, 0)
        ox, oy = np.float32((slot % cols) * w0), np.float32((slot // cols) * h0)
        for j in range(3):
            tx = np.float32(np.float32(b.projected_tri[j][0]) + ox)
            ty = np.float32(np.float32(b.projected_tri[j][1]) + oy)
            uv[f, j, 0] = np.float32(tx + half) / fw
            uv[f, j, 1] = one - np.float32(np.float32(ty + half) / fh)
    mesh.set_uv(uv.reshape(-1, 2))
    mesh.set_uv_indices(np.arange(F * 3, dtype=np.int32).reshape(F, 3))
    _set_texture(mesh, tex, option)
    return True


def _set_texture(mesh, tex, option):
    mesh.materials = [tex]
    mesh.material_ids = np.zeros(len(mesh.vertex_indices), np.int32)
    from .io import ObjMaterial
    m = ObjMaterial(option.texture_base_name)
    m.diffuse_tex = tex
    mesh.obj_materials = [m]


def TextureMapping(keyframes, info, mesh, option):
    """ugu::TextureMapping (texture_mapper.cc:894-1017); returns bool and
    sets the mesh's material (and, for kGenerateSimpleTile, its uv)."""
    if option.criteria not in (ViewSelectionCriteria.kMinViewingAngle,
                               ViewSelectionCriteria.kMinDistance,
                               ViewSelectionCriteria.kMaxArea):
        _log(f"ViewSelectionCriteria {int(option.criteria)} is not implemented")
        return False
    if not keyframes:
        _log("no keyframe")
        return False
    best = best_keyframes(info, option.criteria)
    if option.uv_type == TexturingOutputUvType.kGenerateSimpleTile:
        return simple_tile(keyframes, best, mesh, option)
    if option.uv_type != TexturingOutputUvType.kUseOriginalMeshUv:
        _log(f"OutputUvType {int(option.uv_type)} is not implemented")
        return False
    if len(mesh.uv) == 0 or len(mesh.uv_indices) != len(mesh.vertex_indices):
        _log("OutputUvType kUseOriginalMeshUv is specified but UV on mesh is invalid")
        return False
    ids = sorted({b.kf_id for b in best if b is not None and b.kf_id >= 0})
    if ids and ids[-1] >= len(keyframes):
        _log("keyframe ids must index the keyframe list")
        return False
    recs, n = face_records(mesh, best, option.tex_w, option.tex_h)
    tex, _ = texture_map_faces([k.color for k in keyframes], recs, n, option.tex_w,
                               option.tex_h, option.padding_kernel)
    _set_texture(mesh, tex, option)
    return True
