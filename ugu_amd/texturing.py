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

* kGenerateSimpleTriangles (texture_mapper.cc:347-452): face i rasterized
  (on the GPU, as above) into one half of rectangle i // 2 of a fixed grid,
  then PaddingSimple with a 3 x 3 kernel;
* kGenerateSimpleCharts (texture_mapper.cc:454-922): charts of edge-adjacent
  faces that chose the same keyframe, each the keyframe's pixels over its
  bbox, packed into one texture (host only: copies, a breadth-first walk and
  a guillotine packer).

The result becomes the mesh's single material.
"""
import collections
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


def triangle_grid(n_faces, tex_w, tex_h):
    """Layout of kGenerateSimpleTriangles (texture_mapper.cc:357-395): a
    rectangle of (s + 2) x (s + 1) texels and a 2-texel gap per face pair,
    s the largest side not above min(floor(sqrt(tex_w * tex_h // rects)), 100)
    for which all rectangles fit. Returns (s, pitch_x, pitch_y, per_row) or
    None where the reference gives up (fewer than 6 texels per rectangle, s
    below 3) or would divide by zero (no faces)."""
    rects = (n_faces + 1) // 2
    if rects < 1 or tex_w < 1 or tex_h < 1:
        return None
    texels = tex_h * tex_w // rects
    if texels < 6:
        return None
    side = min(int(math.sqrt(texels)), 100)

    def fits(sd):
        return (tex_w // (sd + 4)) * (tex_h // (sd + 3)) >= rects
    while not fits(side):
        side -= 1
        if side < 3:
            return None
    return side, side + 4, side + 3, tex_w // (side + 4)


def triangle_corners(grid, i):
    """Texel corners of face i (texture_mapper.cc:413-433): the upper-right
    half of rectangle i // 2 for even i, the lower-left half for odd i."""
    s, px, py, per_row = grid
    r = i // 2
    rx, ry = r % per_row, r // per_row
    if i % 2 == 0:
        x0, y0 = px * rx + 2, py * ry
        return [(x0, y0), (x0 + s - 1, y0), (x0 + s - 1, y0 + s - 1)]
    x0, y0 = px * rx, py * ry + 1
    return [(x0, y0), (x0, y0 + s - 1), (x0 + s - 1, y0 + s - 1)]


def simple_triangles(keyframes, best, mesh, option):
    """kGenerateSimpleTriangles: each seen face's projected triangle
    rasterized into its grid half (ugu_hip_texture_map), PaddingSimple with a
    3 x 3 kernel whatever option.padding_kernel is; uv at the corners' texel
    centres, 0 for faces seen by no keyframe."""
    grid = triangle_grid(len(mesh.vertex_indices), option.tex_w, option.tex_h)
    if grid is None:
        _log(f"kGenerateSimpleTriangles: a {option.tex_w} x {option.tex_h} texture cannot "
             f"hold {len(mesh.vertex_indices)} faces")
        return False
    F = len(best)
    recs = (_abi.UguTexFace * max(1, F))()
    uv = np.zeros((F, 3, 2), np.float32)
    fw, fh = np.float32(option.tex_w), np.float32(option.tex_h)
    half, one = np.float32(0.5), np.float32(1.0)
    for i, b in enumerate(best):
        r = recs[i]
        if b is None or b.kf_id < 0:
            r.keyframe = -1
            continue
        if b.kf_id >= len(keyframes):
            _log("keyframe ids must index the keyframe list")
            return False
        r.keyframe = int(b.kf_id)
        for j, (x, y) in enumerate(triangle_corners(grid, i)):
            r.src[2 * j] = float(b.projected_tri[j][0])
            r.src[2 * j + 1] = float(b.projected_tri[j][1])
            r.dst[2 * j], r.dst[2 * j + 1] = float(x), float(y)
            uv[i, j, 0] = np.float32(np.float32(x) + half) / fw
            uv[i, j, 1] = one - np.float32(np.float32(np.float32(y) + half) / fh)
    tex, _ = texture_map_faces([k.color for k in keyframes], recs, F, option.tex_w,
                               option.tex_h, 3)
    mesh.set_uv(uv.reshape(-1, 2))
    mesh.set_uv_indices(np.arange(F * 3, dtype=np.int32).reshape(F, 3))
    _set_texture(mesh, tex, option)
    return True


class _EdgeFaces:
    """Face adjacency of the chart growth (face_adjacency.h:34-263): the
    directed edge (a, b) of face i holds i + 1, summed over faces sharing
    it; a face's neighbours are the non-zero entries of its reversed edges,
    in edge order, minus one; removing a face clears its own entries."""

    def __init__(self, faces):
        self.f = faces
        self.m = collections.defaultdict(int)
        for i, v in enumerate(faces):
            for k in range(3):
                self.m[(int(v[k]), int(v[(k + 1) % 3]))] += i + 1

    def neighbours(self, face):
        v = self.f[face]
        out = []
        for k in range(3):
            e = self.m.get((int(v[(k + 1) % 3]), int(v[k])), 0)
            if e > 0:
                out.append(e - 1)
        return out

    def remove(self, face):
        v = self.f[face]
        for k in range(3):
            key = (int(v[k]), int(v[(k + 1) % 3]))
            if key in self.m:
                self.m[key] = 0


def pack_rects(rects, x0, y0, w, h):
    """BinPacking2D (bin_packer_2d.cc:14-115) for (w, h) sizes in order:
    guillotine best fit, smallest spare area first (earliest on ties), a
    larger free rectangle cut along the axis with the larger leftover (the
    rest, then the strip, appended to the free list) until one matches.
    Returns (placements [(x, y)], free list [(x, y, w, h)]) or None."""
    free = [(x0, y0, w, h)]
    placed = []
    for tw, th in rects:
        if not free:
            return None
        while True:
            best_i, best_d = -1, 2**31 - 1
            for i, (fx, fy, fw, fh) in enumerate(free):
                if th <= fh and tw <= fw and fw * fh - tw * th < best_d:
                    best_i, best_d = i, fw * fh - tw * th
            if best_i < 0:
                return None
            fx, fy, fw, fh = free.pop(best_i)
            if fw == tw and fh == th:
                placed.append((fx, fy))
                break
            if fh - th > fw - tw:
                free += [(fx, fy + th, fw, fh - th), (fx, fy, fw, th)]
            else:
                free += [(fx + tw, fy, fw - tw, fh), (fx, fy, tw, fh)]
    return placed, free


def _chart_patch(members, best, color, faces):
    """Chart::Finalize (texture_mapper.cc:469-617): keyframe pixels over the
    chart's integer bbox with a 1-pixel edge border, and each distinct
    vertex id's position in it (first member naming it), in id order.
    Corners outside the keyframe image are an error (the reference reads
    outside the image there)."""
    h, w = color.shape[:2]
    pts = np.array([best[m].projected_tri for m in members], np.float32).reshape(-1, 2)
    ok = (pts[:, 0] >= 0) & (pts[:, 0] <= w - 1) & (pts[:, 1] >= 0) & (pts[:, 1] <= h - 1)
    if not ok.all():
        _log("projected corner outside the keyframe image")
        return None
    xmin, xmax = int(math.floor(pts[:, 0].min())), int(math.ceil(pts[:, 0].max()))
    ymin, ymax = int(math.floor(pts[:, 1].min())), int(math.ceil(pts[:, 1].max()))
    core = np.asarray(color, np.uint8).reshape(h, w, 3)[ymin:ymax + 1, xmin:xmax + 1]
    patch = np.pad(core, ((1, 1), (1, 1), (0, 0)), mode="edge")
    fx, fy, one = np.float32(xmin), np.float32(ymin), np.float32(1)
    pos = {}
    for m in members:
        for j, vid in enumerate(faces[best[m].face_id]):
            if int(vid) not in pos:
                p = best[m].projected_tri[j]
                pos[int(vid)] = (np.float32(np.float32(p[0]) - fx) + one,
                                 np.float32(np.float32(p[1]) - fy) + one)
    order = sorted(pos)
    slot = {v: k for k, v in enumerate(order)}
    corners = [[slot[int(v)] for v in faces[best[m].face_id]] for m in members]
    return patch, [pos[v] for v in order], corners


def simple_charts(keyframes, best, mesh, option):
    """kGenerateSimpleCharts (texture_mapper.cc:816-922, GenerateAtlas
    :658-814): charts grown breadth-first from the lowest free face through
    shared edges while the keyframe is the same; patches ordered by area,
    largest first, and packed with a 1-texel margin, the texture growing to
    floor(tex * sqrt(2)^k) until they fit (at most 8192); chart vertices at
    ((x + X + 0.5) / W, 1 - (y + Y + 0.5) / H); faces without a keyframe
    share one uv triple at the first free rectangle (or 0). Ties in the area
    order follow libstdc++'s std::sort in the C++ drop-in; this mirror uses
    a stable order, identical whenever patch areas differ."""
    faces = np.asarray(mesh.vertex_indices)
    F = len(faces)
    if len(best) != F:
        _log(f"VisibilityInfo has {len(best)} faces, the mesh {F}")
        return False
    kf = [(-1 if b is None else int(b.kf_id)) for b in best]
    fid = [(i if b is None else int(b.face_id)) for i, b in enumerate(best)]
    for i in range(F):
        if not (0 <= fid[i] < F) or kf[i] >= len(keyframes):
            _log(f"face {i}: face id {fid[i]} or keyframe id {kf[i]} out of range")
            return False
    adj = _EdgeFaces(faces)
    taken = np.zeros(F, bool)
    valid, invalid = [], []
    for seed in range(F):
        if taken[seed]:
            continue
        c = kf[seed]
        members = [seed]
        taken[seed] = True
        queue = collections.deque([seed])
        while queue:
            f = queue.popleft()
            near = adj.neighbours(f)
            adj.remove(f)
            add = [a for a in near if a < F and kf[a] == c and not taken[a]]
            for a in add:
                queue.append(a)
                members.append(a)
                taken[a] = True
        if c < 0:
            invalid.append(members)
        else:
            got = _chart_patch(members, best, keyframes[c].color, faces)
            if got is None:
                return False
            valid.append((members,) + got)
    order = sorted(range(len(valid)), key=lambda k: -valid[k][1].shape[0] * valid[k][1].shape[1])
    rects = [(valid[k][1].shape[1], valid[k][1].shape[0]) for k in order]
    ratio = float(np.float32(math.sqrt(2.0)))
    tw, th, tries = option.tex_w, option.tex_h, 0
    while True:
        got = pack_rects(rects, 1, 1, tw - 1, th - 1)
        if got is not None:
            break
        tries += 1
        tw = int(option.tex_w * ratio ** tries)
        th = int(option.tex_h * ratio ** tries)
        if tw > 8192 or th > 8192:
            _log("Failed to pack in one texture")
            return False
    placed, free = got
    if tries:
        _log(f"Texture size is adjusted ({option.tex_w}, {option.tex_h}) -> ({tw}, {th})")
    tex = np.zeros((th, tw, 3), np.uint8)
    uv, uvi = [], np.zeros((F, 3), np.int32)
    fw, fh = np.float32(tw), np.float32(th)
    half, one = np.float32(0.5), np.float32(1.0)
    for k, (X, Y) in zip(order, placed):
        members, patch, pos, corners = valid[k]
        tex[Y:Y + patch.shape[0], X:X + patch.shape[1]] = patch
        off = len(uv)
        for x, y in pos:
            uv.append((np.float32(np.float32(x + np.float32(X)) + half) / fw,
                       one - np.float32(np.float32(np.float32(y + np.float32(Y)) + half) / fh)))
        for m, cs in zip(members, corners):
            uvi[fid[m]] = [off + cs[0], off + cs[1], off + cs[2]]
    none = (np.float32(0), np.float32(0))
    if free:
        none = (np.float32(np.float32(free[0][0]) + half) / fw,
                one - np.float32(np.float32(np.float32(free[0][1]) + half) / fh))
    n0 = len(uv)
    uv += [none] * 3
    for members in invalid:
        for m in members:
            uvi[fid[m]] = [n0, n0 + 1, n0 + 2]
    mesh.set_uv(np.array(uv, np.float32).reshape(-1, 2))
    mesh.set_uv_indices(uvi)
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
    sets the mesh's material (and, for the generated uv types, its uv)."""
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
    if option.uv_type == TexturingOutputUvType.kGenerateSimpleTriangles:
        return simple_triangles(keyframes, best, mesh, option)
    if option.uv_type == TexturingOutputUvType.kGenerateSimpleCharts:
        return simple_charts(keyframes, best, mesh, option)
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
