"""Visibility tester on the gfx950 raytracer (ugu::VisibilityTester,
src/texturing/visibility_tester.cc, include/ugu/texturing/visibility_tester.h).

The per-vertex test of one keyframe -- projection, mask and depth checks,
back-face angle, the ray through the projected point against the mesh
(closest hit), occlusion (hit farther than `eps` from the vertex), depth
difference, colour sampling -- runs as one HIP kernel over all vertices
(`ugu_hip_visibility_test`, BVH traversal per lane). The host assembles
VisibilityInfo exactly as the reference's loops do:

* ``Test(info, facewise=False)``: TestVertices, then TestFaces (a face is
  visible from a keyframe when its three vertices are);
* ``Test(info, facewise=True)`` (the default): TestFacewiseVertices -- each
  corner tested 1e-3 of the way towards its face centre, a face visible when
  its corners pass in order (the test stops at the first failing corner), a
  vertex's record taken from the first face (lowest id) whose test of it
  passes. The reference runs this loop under OpenMP with a racy
  "first face" flag; here the order is the serial one.

CalcStat fills the per-vertex / per-face statistics that do not depend on
library internals: means, minima (with index, keyframe id and colour),
weighted averages and weighted medians (NormalizeWeights, math_util.h:29-77),
medians for lists of 1, 2 or an even number of entries and odd ones by the
reference's formula on the sorted list (see `median`), minimum intensity,
and the colour modes (`mode`, plain and weighted by inverse viewing angle).
"""
import ctypes
import math
import sys

import numpy as np

from . import _abi
from .types import ColorInterpolation, Mesh, radians_f32

FLT_MAX = float(np.finfo(np.float32).max)


def _log(msg):
    sys.stderr.write(msg if msg.endswith("\n") else msg + "\n")


class VisibilityTesterOption:
    """VisibilityTesterOption (visibility_tester.h:38-57)."""

    def __init__(self, **kw):
        self.backface_culling_angle_rad_th = float(radians_f32(90.0))
        self.use_mask = True
        self.use_depth = True
        self.interp = ColorInterpolation.kBilinear
        self.calc_stat_vertex_info = True
        self.collect_face_info = True
        self.calc_stat_face_info = True
        self.max_viewing_angle = float(radians_f32(90.0))
        self.max_distance_from_camera_to_vertex = FLT_MAX
        self.max_depth_difference = FLT_MAX
        self.eps = 0.1
        self.invalid_color = np.zeros(3, np.float32)
        for k, v in kw.items():
            if not hasattr(self, k):
                raise TypeError(f"unknown option {k}")
            setattr(self, k, v)

    def CopyTo(self, dst):
        dst.__dict__.update({k: (v.copy() if isinstance(v, np.ndarray) else v)
                             for k, v in self.__dict__.items()})

    def to_abi(self):
        o = _abi.UguVisibilityOptions()
        o.backface_culling_angle_rad_th = self.backface_culling_angle_rad_th
        o.use_mask = 1 if self.use_mask else 0
        o.use_depth = 1 if self.use_depth else 0
        o.interp = int(self.interp)
        o.max_viewing_angle = self.max_viewing_angle
        o.max_distance_from_camera_to_vertex = self.max_distance_from_camera_to_vertex
        o.max_depth_difference = self.max_depth_difference
        o.eps = self.eps
        return o


class Keyframe:
    """Keyframe (visibility_tester.h:229-246): id, colour (H, W, 3 uint8),
    depth (H, W float32), mask (H, W uint8), camera (PinholeCamera)."""

    def __init__(self, id=-1, color=None, depth=None, mask=None, camera=None):
        self.id, self.color, self.depth, self.mask, self.camera = id, color, depth, mask, camera


class VertexInfoPerKeyframe:
    __slots__ = ("kf_id", "color", "intensity", "projected_pos", "viewing_angle", "distance")

    def __init__(self, kf_id, rec):
        self.kf_id = kf_id
        self.color = np.array([rec["c0"], rec["c1"], rec["c2"]], np.float32)
        self.intensity = np.float32(rec["intensity"])
        self.projected_pos = np.array([rec["px"], rec["py"]], np.float32)
        self.viewing_angle = np.float32(rec["angle"])
        self.distance = np.float32(rec["dist"])


class FaceInfoPerKeyframe:
    __slots__ = ("kf_id", "face_id", "area", "viewing_angle", "distance", "projected_tri")

    def __init__(self, kf_id, face_id):
        self.kf_id, self.face_id = kf_id, face_id
        self.area = np.float32(-1.0)
        self.viewing_angle = np.float32(0.0)
        self.distance = np.float32(0.0)
        self.projected_tri = np.zeros((3, 2), np.float32)


def normalize_weights(weights):
    """NormalizeWeights (math_util.h:29-48) in float: every weight plus 1e-6,
    divided by the float sum of those taken in order; all 1/n when that sum
    is not above 1e-6."""
    w = np.asarray(weights, np.float32) + np.float32(0.000001)
    s = np.float32(0.0)
    for x in w:
        s = np.float32(s + x)
    if s > 0.000001:
        return (w / s).astype(np.float32)
    return np.full(len(w), np.float32(1.0) / np.float32(len(w)), np.float32)


def mode(colors, weights=None):
    """Mode (math_util.h:130-181): the first colour, in input order, whose
    running (weighted) count becomes the largest so far. Returns (mode,
    frequency, occurrence dict keyed by the colour's float tuple)."""
    occ = {}
    best, freq = None, (0 if weights is None else np.float32(0.0))
    for i, c in enumerate(colors):
        key = tuple(float(x) for x in np.asarray(c, np.float32))
        w = 1 if weights is None else np.float32(weights[i])
        occ[key] = (occ[key] + w) if key in occ else w
        if weights is not None:
            occ[key] = np.float32(occ[key])
        if freq < occ[key]:
            freq, best = occ[key], np.asarray(c, np.float32)
    return best, freq, occ


def weighted_average(colors, weights):
    """WeightedAverage (math_util.h:50-70): double accumulation of
    float(colour * weight), cast to float."""
    nw = normalize_weights(weights)
    acc = [0.0, 0.0, 0.0]
    for c, w in zip(colors, nw):
        for k in range(3):
            acc[k] += float(np.float32(c[k] * w))
    return np.array(acc, np.float32)


def weighted_median(colors, weights):
    """WeightedMedian (math_util.h:72-101): sort by normalized weight, first
    entry whose running float sum exceeds 0.5 (entry 0 if none)."""
    nw = normalize_weights(weights)
    order = sorted(range(len(nw)), key=lambda i: nw[i])  # equal weights: unspecified upstream
    s = np.float32(0.0)
    for i in order:
        s = np.float32(s + nw[i])
        if s > 0.5:
            return np.asarray(colors[i], np.float32)
    return np.asarray(colors[order[0]], np.float32)


def median(data, force_org_val=False):
    """Median (math_util.h:103-126). Sizes 1, 2 and even sizes are exact (the
    n-th element). For odd sizes >= 3 the reference averages the element
    nth_element leaves at n with the (n+1)-th smallest; that element is the
    n-th smallest whenever the selection ends with an insertion sort (all
    lists of 3), which is what is returned here (parity unpinned beyond 3)."""
    d = [np.float32(x) for x in data]
    if len(d) == 1:
        return d[0]
    if len(d) == 2:
        return d[0] if force_org_val else np.float32((d[0] + d[1]) * np.float32(0.5))
    s = sorted(d)
    n = len(d) // 2
    if force_org_val or len(d) % 2 == 0:
        return s[n]
    return np.float32((s[n] + s[n + 1]) * np.float32(0.5))


class VertexInfo:
    def __init__(self):
        self.visible_keyframes = []

    def VisibleFrom(self, kf_id):
        for i, k in enumerate(self.visible_keyframes):
            if k.kf_id == kf_id:
                return i
        return -1

    def CalcStat(self, vert_custom_func=None):
        """VertexInfo::CalcStat (visibility_tester.cc:68-153). Distances of
        FLT_MAX (no hit) overflow float sums to inf, as in the reference.
        ``vert_custom_func(self)`` runs last when given."""
        with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
            self._calc_stat()
        if vert_custom_func is not None and self.visible_keyframes:
            vert_custom_func(self)

    def _calc_stat(self):
        kfs = self.visible_keyframes
        if not kfs:
            return
        n = np.float32(len(kfs))
        mc = np.zeros(3, np.float32)
        md = np.float32(0.0)
        ma = np.float32(0.0)
        for k in kfs:
            mc = (mc + k.color).astype(np.float32)
            md = np.float32(md + k.distance)
            ma = np.float32(ma + k.viewing_angle)
        self.mean_color = (mc / n).astype(np.float32)
        self.mean_distance = np.float32(md / n)
        self.mean_viewing_angle = np.float32(ma / n)
        colors = [k.color for k in kfs]
        dists = [k.distance for k in kfs]
        angles = [k.viewing_angle for k in kfs]
        inten = [k.intensity for k in kfs]
        inv_d = [np.float32(1.0) / d for d in dists]
        inv_a = [np.float32(1.0) / a for a in angles]
        i = int(np.argmin(angles))  # first minimum, as std::min_element
        self.min_viewing_angle_index, self.min_viewing_angle_id = i, kfs[i].kf_id
        self.min_viewing_angle_color = kfs[i].color
        i = int(np.argmin(dists))
        self.min_distance_index, self.min_distance_id = i, kfs[i].kf_id
        self.min_distance_color = kfs[i].color
        self.median_viewing_angle = median(angles)
        self.median_distance = median(dists)
        self.median_color = np.array([median([c[k] for c in colors]) for k in range(3)],
                                     np.float32)
        self.mean_viewing_angle_color = weighted_average(colors, inv_a)
        self.mean_distance_color = weighted_average(colors, inv_d)
        self.median_viewing_angle_color = weighted_median(colors, inv_a)
        self.median_distance_color = weighted_median(colors, inv_d)
        i = int(np.argmin(inten))
        self.min_intensity, self.min_intensity_color = inten[i], colors[i]
        self.median_intensity = median(inten, True)
        self.median_intensity_color = colors[inten.index(self.median_intensity)]
        self.mode, self.mode_frequency, self.occurrence = mode(colors)
        (self.mode_viewing_angle_color, self.mode_frequency_viewing_angle_color,
         self.occurrence_viewing_angle) = mode(colors, inv_a)


class FaceInfo:
    def __init__(self):
        self.visible_keyframes = []
        self.max_area_index = self.max_area_id = -1
        self.max_area = np.float32(-1.0)
        self.min_viewing_angle_index = self.min_viewing_angle_id = -1
        self.min_viewing_angle = np.float32(np.finfo(np.float32).max)
        self.min_distance_index = self.min_distance_id = -1
        self.min_distance = np.float32(np.finfo(np.float32).max)

    def CalcStat(self):
        """FaceInfo::CalcStat (visibility_tester.cc:185-224): the indices and
        keyframe ids of the largest area and the smallest angle and distance
        (first one on ties); max_area, min_viewing_angle and min_distance keep
        their initial values, as in the reference."""
        kfs = self.visible_keyframes
        if not kfs:
            return
        i = int(np.argmax([k.area for k in kfs]))
        self.max_area_index, self.max_area_id = i, kfs[i].kf_id
        i = int(np.argmin([k.viewing_angle for k in kfs]))
        self.min_viewing_angle_index, self.min_viewing_angle_id = i, kfs[i].kf_id
        i = int(np.argmin([k.distance for k in kfs]))
        self.min_distance_index, self.min_distance_id = i, kfs[i].kf_id


class VisibilityInfo:
    """VisibilityInfo (visibility_tester.h:183-199)."""

    def __init__(self, mesh=None, num_vertices=0, num_faces=0):
        if mesh is not None:
            num_vertices, num_faces = len(mesh.vertices), len(mesh.vertex_indices)
        self.vertex_info_list = [VertexInfo() for _ in range(num_vertices)]
        self.face_info_list = [FaceInfo() for _ in range(num_faces)]
        self.has_vertex_stat = False
        self.has_face_stat = False

    def CalcStatVertexInfo(self, vert_custom_func=None):
        for v in self.vertex_info_list:
            v.CalcStat(vert_custom_func)
        self.has_vertex_stat = True

    def CalcStatFaceInfo(self):
        for f in self.face_info_list:
            f.CalcStat()
        self.has_face_stat = True


def _edge(a, b, c):
    # EdgeFunction (rasterizer.cc:25-28) in float
    return np.float32(np.float32((c[0] - a[0]) * (b[1] - a[1])) -
                      np.float32((c[1] - a[1]) * (b[0] - a[0])))


def _face_record(kf_id, fid, corners):
    with np.errstate(over="ignore", invalid="ignore"):
        return _face_record_f32(kf_id, fid, corners)


def _face_record_f32(kf_id, fid, corners):
    fi = FaceInfoPerKeyframe(kf_id, fid)
    a = np.float32(0.0)
    d = np.float32(0.0)
    for j, c in enumerate(corners):
        a = np.float32(a + c.viewing_angle)
        d = np.float32(d + c.distance)
        fi.projected_tri[j] = c.projected_pos
    fi.viewing_angle = np.float32(a / np.float32(3))
    fi.distance = np.float32(d / np.float32(3))
    p = fi.projected_tri
    fi.area = np.float32(np.float32(0.5) * abs(_edge(p[0], p[1], p[2])))
    return fi


class VisibilityTester:
    """ugu::VisibilityTester: set_option / set_data / PrepareData /
    set_keyframe / Test, with the vertex test on the GPU."""

    def __init__(self, option=None, device=0):
        self._lib = _abi.load_library()
        self._ctx = self._lib.ugu_hip_create(int(device))
        if not self._ctx:
            raise RuntimeError(f"ugu_hip_create({device}) failed: no usable HIP device")
        self.option_ = VisibilityTesterOption()
        if option is not None:
            self.set_option(option)
        self.mesh_initialized_ = False
        self.keyframe_ = None

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx:
            self._lib.ugu_hip_destroy(ctx)
            self._ctx = None

    def set_option(self, option):
        option.CopyTo(self.option_)

    def set_data(self, vertices, normals, indices):
        self.mesh_initialized_ = False
        self.vertices_ = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        self.normals_ = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        self.indices_ = np.ascontiguousarray(indices, np.int32).reshape(-1, 3)

    def PrepareData(self):
        """visibility_tester.cc:309-333: BVH over the vertices / faces and the
        face centres ((v0 + v1) + v2) / 3 in float."""
        if len(self.vertices_) == 0 or len(self.normals_) == 0 or len(self.indices_) == 0:
            _log("mesh is empty\n")
            return False
        from .renderer import MeshDesc
        desc = MeshDesc(Mesh(self.vertices_, self.indices_))
        for fn in (lambda: self._lib.ugu_hip_prepare_mesh(self._ctx, ctypes.byref(desc.desc)),
                   lambda: self._lib.ugu_hip_prepare_raytracer(self._ctx)):
            if fn() != 0:
                _log(self._lib.ugu_hip_last_error(self._ctx).decode())
                return False
        v = self.vertices_[self.indices_]
        c = np.zeros((len(self.indices_), 3), np.float32)
        for k in range(3):
            c = (c + v[:, k]).astype(np.float32)
        self.face_centers_ = (c / np.float32(3)).astype(np.float32)
        self.mesh_initialized_ = True
        return True

    def set_keyframe(self, keyframe):
        self.keyframe_ = keyframe

    def _validate(self, info):
        kf, o = self.keyframe_, self.option_
        msg = None
        if info is None:
            msg = "info is nullptr"
        elif kf is None:
            msg = "keyframe has not been set"
        elif not self.mesh_initialized_:
            msg = "mesh has not been initialized"
        elif kf.color is None or np.size(kf.color) == 0:
            msg = "keyframe color is empty"
        elif o.use_mask and (kf.mask is None or np.size(kf.mask) == 0):
            msg = "use_mask is true but keyframe mask is empty"
        elif o.use_depth and (kf.depth is None or np.size(kf.depth) == 0):
            msg = "use_depth is true but keyframe depth is empty"
        elif len(info.vertex_info_list) != len(self.vertices_):
            msg = "info->vertex_info_list size is different from mesh"
        elif len(info.vertex_info_list) != len(self.normals_):
            msg = "info->vertex_info_list size is different from mesh normal"
        elif len(info.face_info_list) != len(self.indices_):
            msg = "info->face_info_list size is different from mesh"
        if msg:
            _log(msg + "\n")
            return False
        for vi in info.vertex_info_list:
            for k in ("min_viewing_angle_color", "min_distance_color", "mean_color",
                      "median_color", "mean_viewing_angle_color", "median_viewing_angle_color",
                      "mean_distance_color", "median_distance_color"):
                setattr(vi, k, o.invalid_color.copy())
        return True

    def test_items(self, pos, vert, nrm):
        """ugu_hip_visibility_test on the current keyframe: one record
        (_abi.VIS_DTYPE) per item."""
        kf = self.keyframe_
        cam = kf.camera.to_abi()
        lin = kf.camera.w2c[:3, :3].astype(np.float32)  # w2c().cast<float>() linear part
        for i in range(9):
            cam.w2c_R[i] = float(lin.reshape(-1)[i])
        color = np.ascontiguousarray(kf.color, np.uint8)
        h, w = color.shape[:2]
        depth = None if kf.depth is None else np.ascontiguousarray(kf.depth, np.float32)
        mask = None if kf.mask is None else np.ascontiguousarray(kf.mask, np.uint8)
        k = _abi.UguKeyframe()
        k.camera, k.width, k.height = cam, w, h
        k.color = color.ctypes.data
        k.depth = None if depth is None else depth.ctypes.data
        k.mask = None if mask is None else mask.ctypes.data
        pos = np.ascontiguousarray(pos, np.float32)
        vert = np.ascontiguousarray(vert, np.float32)
        nrm = np.ascontiguousarray(nrm, np.float32)
        out = np.zeros(len(pos), _abi.VIS_DTYPE)
        opt = self.option_.to_abi()
        rc = self._lib.ugu_hip_visibility_test(
            self._ctx, ctypes.byref(k), ctypes.byref(opt), len(pos), pos.ctypes.data,
            vert.ctypes.data, nrm.ctypes.data, out.ctypes.data)
        if rc != 0:
            raise _abi.HipError(rc, self._lib.ugu_hip_last_error(self._ctx).decode())
        return out, cam

    def facewise_items(self):
        """Item arrays of TestFacewiseVertices: corner j of face i moved by
        (centre - v) * 1e-3 (visibility_tester.cc:588-592)."""
        v = self.vertices_[self.indices_]                      # F x 3 x 3
        off = ((self.face_centers_[:, None, :] - v) * np.float32(1e-3)).astype(np.float32)
        pos = (v + off).astype(np.float32).reshape(-1, 3)
        nrm = self.normals_[self.indices_].reshape(-1, 3)
        return pos, v.reshape(-1, 3), nrm

    def Test(self, info, facewise=True, keyframes=None, vert_custom_func=None):
        """Test(info, facewise) for the current keyframe, or, with
        `keyframes`, Test(keyframes, info, facewise) (statistics only after
        the last keyframe, visibility_tester.cc:693-727).
        ``vert_custom_func(vertex_info)`` runs at the end of every vertex's
        CalcStat."""
        cf = vert_custom_func
        if keyframes is not None:
            org = VisibilityTesterOption()
            self.option_.CopyTo(org)
            self.option_.calc_stat_face_info = False
            self.option_.calc_stat_vertex_info = False
            ok = True
            for kf in keyframes[:-1]:
                self.set_keyframe(kf)
                if not self.Test(info, facewise, vert_custom_func=cf):
                    ok = False
                    break
            self.set_option(org)
            self.set_keyframe(keyframes[-1])
            return self.Test(info, facewise, vert_custom_func=cf) and ok
        if not self._validate(info):
            return False
        kf_id = self.keyframe_.id
        o = self.option_
        if facewise:
            pos, vert, nrm = self.facewise_items()
            rec, _ = self.test_items(pos, vert, nrm)
            rec = rec.reshape(-1, 3)
            done = np.zeros(len(self.vertices_), bool)
            for f, face in enumerate(self.indices_):
                corners = []
                for j in range(3):
                    if not rec[f, j]["visible"]:
                        break
                    vi = VertexInfoPerKeyframe(kf_id, rec[f, j])
                    if not done[face[j]]:
                        done[face[j]] = True
                        info.vertex_info_list[face[j]].visible_keyframes.append(vi)
                    corners.append(vi)
                if len(corners) == 3:
                    info.face_info_list[f].visible_keyframes.append(
                        _face_record(kf_id, f, corners))
            if o.calc_stat_vertex_info:
                info.CalcStatVertexInfo(cf)
            if o.calc_stat_face_info:
                info.CalcStatFaceInfo()
            return True
        rec, _ = self.test_items(self.vertices_, self.vertices_, self.normals_)
        for i in np.nonzero(rec["visible"])[0]:
            info.vertex_info_list[i].visible_keyframes.append(VertexInfoPerKeyframe(kf_id, rec[i]))
        if o.calc_stat_vertex_info:
            info.CalcStatVertexInfo(cf)
        if o.collect_face_info:
            for f, face in enumerate(self.indices_):
                corners = []
                for j in range(3):
                    vinfo = info.vertex_info_list[face[j]]
                    idx = vinfo.VisibleFrom(kf_id)
                    if idx < 0:
                        break
                    corners.append(vinfo.visible_keyframes[idx])
                if len(corners) == 3:
                    info.face_info_list[f].visible_keyframes.append(
                        _face_record(kf_id, f, corners))
            if o.calc_stat_face_info:
                info.CalcStatFaceInfo()
        return True
