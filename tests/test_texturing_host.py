"""CPU: the visibility statistics and kGenerateSimpleTile texture mapping
against values computed by hand from the definitions in
include/ugu/util/math_util.h:29-181, src/texturing/visibility_tester.cc:68-224
and src/texturing/texture_mapper.cc:30-136 -- for the C++ drop-in (through
ugu_amd/cpp/ugu_stats_tool, no GPU) and for the Python host, each against the
same expected numbers, never against each other."""
import os
import subprocess

import numpy as np
import pytest

from ugu_amd import texturing as T
from ugu_amd import visibility as V

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# UGU_STATS_TOOL: another build of the same driver (tests/sanitize)
TOOL = os.environ.get("UGU_STATS_TOOL") or os.path.join(ROOT, "ugu_amd", "cpp", "ugu_stats_tool")
FLT_MAX = float(np.finfo(np.float32).max)

C0, C1, C2 = (200.0, 100.0, 50.0), (10.0, 20.0, 30.0), (200.0, 100.0, 50.0)
# kf_id, colour, intensity, viewing angle, distance
VERTEX_A = [(0, C0, 0.3, 0.5, 2.0), (1, C1, 0.1, 0.25, np.inf), (2, C2, 0.7, 1.0, 4.0)]
VERTEX_B = [(0, (1.0, 2.0, 3.0), 0.5, 1.0, np.inf), (1, (4.0, 5.0, 6.0), 0.5, 1.0, 2.0),
            (2, (7.0, 8.0, 9.0), 0.5, 1.0, 2.0)]

# VERTEX_A by hand:
#  mean colour      (C0 + C1 + C2) / 3 = (136.667, 73.333, 43.333)
#  median colour    3 entries: the mean of the 2nd and 3rd smallest (nth_element
#                   at n + 1 = 2 leaves the sorted triple) -> C0 per channel
#  angle weights    1/angle = (2, 4, 1) (+1e-6 each) -> weighted average
#                   (2 C0 + 4 C1 + C2) / 7 = (91.4286, 54.2857, 38.5714);
#                   weighted median: ascending (1/7 C2, 2/7 C0, 4/7 C1), the
#                   running sum first exceeds 0.5 at C1
#  distance weights 1/dist = (0.5, 0, 0.25) plus 1e-6 each, over 0.750003:
#                   average = C0 (1 - 1.3333e-6) + C1 1.3333e-6 per channel
#                   (without the 1e-6 it would be exactly C0 = 200, 100);
#                   median: ascending (C1, C2, C0) crosses 0.5 at C0
#  intensities      min 0.1 (C1); median with force_org_val: nth_element at 1
#                   -> 0.3, the colour of the first keyframe with 0.3 (C0)
#  angles           mean 0.58333, median (0.5 + 1.0) / 2 = 0.75
#  distances        mean inf, median (4 + inf) / 2 = inf
#  mode             C0 twice (frequency 2); weighted by 1/angle: C0 reaches 2,
#                   C1 4, C0 3 -> C1 with 4
# VERTEX_B: distance weights (0, 0.5, 0.5) + 1e-6: normalised (1e-6, 0.4999995,
#  0.4999995); ascending order keeps list order for equal weights, and the
#  running sum 0.5000005 exceeds 0.5 at the second entry -> (4, 5, 6). Without
#  the 1e-6 it reaches exactly 0.5 there and the median would be (7, 8, 9).
EXPECT_A = {
    "mean_color": (136.666667, 73.333333, 43.333333),
    "median_color": C0,
    "min_viewing_angle_index": 1, "min_viewing_angle_id": 1, "min_viewing_angle_color": C1,
    "mean_viewing_angle_color": (640.0 / 7, 380.0 / 7, 270.0 / 7),
    "median_viewing_angle_color": C1,
    "min_distance_index": 0, "min_distance_id": 0, "min_distance_color": C0,
    "mean_distance_color": (200.0 - 190.0 * 1e-6 / 0.750003, 100.0 - 80.0 * 1e-6 / 0.750003,
                            50.0 - 20.0 * 1e-6 / 0.750003),
    "median_distance_color": C0,
    "min_intensity": 0.1, "min_intensity_color": C1,
    "median_intensity": 0.3, "median_intensity_color": C0,
    "mean_viewing_angle": 0.5833333, "median_viewing_angle": 0.75,
    "mean_distance": np.inf, "median_distance": np.inf,
    "mode": C0, "mode_frequency": 2,
    "mode_viewing_angle_color": C1, "mode_frequency_viewing_angle_color": 4.0,
}
# float32 sums of a few terms of size <= 200: a few ulps of 200
TOL = 3e-5


def _tool(*args):
    r = subprocess.run([TOOL, *map(str, args)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _cpp_vertex(records, tmp_path):
    p = tmp_path / "v.txt"
    p.write_text("".join(f"{k} {c[0]} {c[1]} {c[2]} {i} {a} {d}\n"
                         for k, c, i, a, d in records))
    out = {}
    for line in _tool("vertex", p).splitlines():
        name, *vals = line.split()
        v = [float.fromhex(x) if ("0x" in x or "inf" in x or "nan" in x) else int(x)
             for x in vals]
        out[name] = v[0] if len(v) == 1 else tuple(v)
    return out


def _py_vertex(records):
    vi = V.VertexInfo()
    for k, c, i, a, d in records:
        r = V.VertexInfoPerKeyframe.__new__(V.VertexInfoPerKeyframe)
        r.kf_id, r.color = k, np.array(c, np.float32)
        r.intensity, r.viewing_angle, r.distance = np.float32(i), np.float32(a), np.float32(d)
        r.projected_pos = np.zeros(2, np.float32)
        vi.visible_keyframes.append(r)
    calls = []

    def custom(v):
        calls.append(1)
        v.custom_best = v.mode

    vi.CalcStat(custom)
    out = {k: getattr(vi, k) for k in EXPECT_A}
    out["custom_calls"] = len(calls)
    out["custom_best"] = vi.custom_best
    return out


def _check(got, expect):
    for name, exp in expect.items():
        g = np.asarray(got[name], np.float64)
        e = np.asarray(exp, np.float64)
        if np.issubdtype(np.asarray(exp).dtype, np.integer):
            assert np.array_equal(g, e), f"{name}: {g} != {e}"
        else:
            assert np.allclose(g, e, rtol=0, atol=TOL, equal_nan=False) or \
                np.array_equal(g, e), f"{name}: {g} != {e}"


@pytest.fixture(params=["cpp", "python"])
def stats(request, tmp_path):
    if request.param == "cpp":
        return lambda recs: _cpp_vertex(recs, tmp_path)
    return _py_vertex


def test_vertex_statistics_by_hand(stats):
    got = stats(VERTEX_A)
    _check(got, EXPECT_A)
    assert got["custom_calls"] == 1
    _check({"custom_best": got["custom_best"]}, {"custom_best": C0})
    # the 1e-6 added to every weight moves the inverse-distance average away
    # from C0 by more than the tolerance
    assert abs(got["mean_distance_color"][0] - 200.0) > 5 * TOL


def test_weighted_median_depends_on_weight_eps(stats):
    got = stats(VERTEX_B)
    _check(got, {"median_distance_color": (4.0, 5.0, 6.0),
                 "mean_viewing_angle_color": (4.0, 5.0, 6.0),
                 "median_intensity": 0.5, "median_intensity_color": (1.0, 2.0, 3.0),
                 "mode": (1.0, 2.0, 3.0), "mode_frequency": 1,
                 "median_viewing_angle": 1.0})


FACE_RECS = [(3, 10.0, 0.5, 2.0), (5, 30.0, 0.2, 3.0), (7, 30.0, 0.9, 1.0)]
# first largest area -> index 1 (id 5); smallest angle index 1 (id 5);
# smallest distance index 2 (id 7); the value fields keep their initial values
EXPECT_FACE = {"max_area_index": 1, "max_area_id": 5, "min_viewing_angle_index": 1,
               "min_viewing_angle_id": 5, "min_distance_index": 2, "min_distance_id": 7,
               "max_area": -1.0, "min_viewing_angle": FLT_MAX, "min_distance": FLT_MAX}


def test_face_statistics_by_hand(tmp_path):
    p = tmp_path / "f.txt"
    p.write_text("".join(f"{k} {a} {g} {d}\n" for k, a, g, d in FACE_RECS))
    cpp = {}
    for line in _tool("face", p).splitlines():
        name, val = line.split()
        cpp[name] = float.fromhex(val) if "0x" in val else int(val)
    fi = V.FaceInfo()
    for k, a, g, d in FACE_RECS:
        r = V.FaceInfoPerKeyframe(k, 0)
        r.area, r.viewing_angle, r.distance = np.float32(a), np.float32(g), np.float32(d)
        fi.visible_keyframes.append(r)
    fi.CalcStat()
    py = {k: getattr(fi, k) for k in EXPECT_FACE}
    for got in (cpp, py):
        for k, e in EXPECT_FACE.items():
            assert float(got[k]) == float(np.float32(e)), (k, got[k], e)


# ---------------------------------------------------------------------------
# kGenerateSimpleTile

W, H = 6, 4
IDS = [10, 11, 12, 13, 14]
TRI = [(0.25, 0.5), (3.75, 1.0), (2.0, 2.5)]
# face 0: seen by keyframes 12 (area 5) and 10 (area 9) -> kMaxArea picks 10
# (slot 0, tile (0, 0)); face 1: unseen; face 2: only keyframe 14 (slot 4 of
# a grid with floor(sqrt(5)) = 2 columns and ceil(5 / 2) = 3 rows: tile
# (0, 2), offset (0, 8)); texture 12 x 12
FACES = [[(12, 5.0), (10, 9.0)], [], [(14, 2.0)]]


def _keyframe_images():
    rng = np.random.default_rng(3)
    return [rng.integers(0, 256, (H, W, 3), dtype=np.uint8) for _ in IDS]


def _expected_tile():
    imgs = _keyframe_images()
    tex = np.zeros((12, 12, 3), np.uint8)
    for i, im in enumerate(imgs):
        tex[(i // 2) * H:(i // 2 + 1) * H, (i % 2) * W:(i % 2 + 1) * W] = im
    uv = np.zeros((3, 3, 2), np.float64)
    for f, (ox, oy) in ((0, (0, 0)), (2, (0, 8))):
        for j, (x, y) in enumerate(TRI):
            uv[f, j] = ((x + ox + 0.5) / 12.0, 1.0 - (y + oy + 0.5) / 12.0)
    return tex, uv


def _cpp_tile(tmp_path):
    imgs = _keyframe_images()
    (tmp_path / "meta.txt").write_text(f"{len(IDS)} {W} {H} {len(FACES)} 6\n")
    (tmp_path / "ids.txt").write_text(" ".join(map(str, IDS)))
    for i, im in enumerate(imgs):
        (tmp_path / f"kf{i}.raw").write_bytes(im.tobytes())
    lines = []
    for recs in FACES:
        lines.append(str(len(recs)))
        for k, a in recs:
            lines.append(f"{k} {a} 0.5 1.0 " + " ".join(f"{x} {y}" for x, y in TRI))
    (tmp_path / "faces.txt").write_text("\n".join(lines) + "\n")
    ok, tw, th = map(int, _tool("tile", tmp_path).split())
    assert ok == 1
    tex = np.fromfile(tmp_path / "tex.u8", np.uint8).reshape(th, tw, 3)
    uv = np.fromfile(tmp_path / "uv.f32", np.float32).reshape(-1, 2)
    uvi = np.fromfile(tmp_path / "uvi.i32", np.int32).reshape(-1, 3)
    return tex, uv[uvi]


def _py_tile():
    from ugu_amd import Mesh
    imgs = _keyframe_images()
    kfs = []
    for i, im in enumerate(imgs):
        k = V.Keyframe()
        k.id, k.color = IDS[i], im
        kfs.append(k)
    mesh = Mesh(np.zeros((3, 3), np.float32), np.tile(np.arange(3, dtype=np.int32), (3, 1)))
    info = V.VisibilityInfo(num_vertices=3, num_faces=3)
    for f, recs in enumerate(FACES):
        for k, a in recs:
            r = V.FaceInfoPerKeyframe(k, f)
            r.area, r.viewing_angle, r.distance = np.float32(a), np.float32(0.5), np.float32(1)
            r.projected_tri = np.array(TRI, np.float32)
            info.face_info_list[f].visible_keyframes.append(r)
    info.CalcStatFaceInfo()
    opt = T.TextureMappingOption()
    assert opt.uv_type == T.TexturingOutputUvType.kGenerateSimpleTile  # the default
    assert T.TextureMapping(kfs, info, mesh, opt)
    return mesh.materials[0], np.asarray(mesh.uv)[np.asarray(mesh.uv_indices)]


@pytest.mark.parametrize("impl", ["cpp", "python"])
def test_simple_tile_by_hand(impl, tmp_path):
    tex, uv = _cpp_tile(tmp_path) if impl == "cpp" else _py_tile()
    exp_tex, exp_uv = _expected_tile()
    assert np.array_equal(tex, exp_tex)
    assert np.allclose(uv, exp_uv, rtol=0, atol=1e-7)


# ---------------------------------------------------------------------------
# kGenerateSimpleCharts (texture_mapper.cc:454-922) and the triangle grid of
# kGenerateSimpleTriangles (:347-452)

CW, CH = 8, 6  # keyframe size
# five faces over ten vertices; faces 0-3 share edges (0-4, 1-4, 1-5), face 4
# stands alone
CHART_VI = [(0, 3, 4), (0, 4, 1), (1, 4, 5), (1, 5, 2), (6, 7, 8)]
CHART_KF = [0, 0, 1, None, 0]  # face 3 is seen by no keyframe
# projected corner of (vertex, keyframe)
CHART_PROJ = {(0, 0): (1.2, 0.5), (1, 0): (4.0, 0.2), (3, 0): (1.0, 3.5), (4, 0): (3.5, 3.0),
              (6, 0): (5.0, 1.0), (7, 0): (7.0, 1.0), (8, 0): (6.0, 2.0),
              (1, 1): (2.0, 1.0), (4, 1): (2.5, 4.5), (5, 1): (6.5, 4.0)}


def _chart_images():
    rng = np.random.default_rng(11)
    return [rng.integers(0, 256, (CH, CW, 3), dtype=np.uint8) for _ in range(2)]


def _f32(x):
    return np.float32(x)


# placements (B, A, D) and the first free rectangle's corner, by hand below
CHART_LAYOUT = {16: ((1, 1), (9, 1), (1, 12), (15, 1)),
                15: ((1, 8), (9, 8), (1, 1), (6, 1))}


def _expected_charts(size):
    """By hand. Charts in growth order: A = faces 0, 1 (keyframe 0; face 0's
    only neighbour is 1; face 1's neighbour 2 chose keyframe 1), B = face 2
    (keyframe 1; face 3 is unseen), C = face 3 (no keyframe), D = face 4.
    Patches: A bbox x 1..4, y 0..4 -> 7 x 6 (rows x cols, 1-pixel border);
    B x 2..7, y 1..5 -> 7 x 8; D x 5..7, y 1..2 -> 4 x 5. By area: B (56),
    A (42), D (20). Packing into 16 x 16 from the free rectangle (1, 1, 15,
    15): B cuts it into (1, 8, 15, 8) + (1, 1, 15, 7), then (1, 1, 15, 7)
    into (9, 1, 7, 7) + (1, 1, 8, 7): B at (1, 1). A: (9, 1, 7, 7) is the
    tightest, cut into (15, 1, 1, 7) + (9, 1, 6, 7): A at (9, 1). D: (1, 8,
    15, 8) cut along x into (6, 8, 10, 8) + (1, 8, 5, 8), that along y into
    (1, 12, 5, 4) + (1, 8, 5, 4); the first exact fit is (1, 12): D at
    (1, 12). Free list then starts with (15, 1, 1, 7): the unseen uv.
    From an 8 x 8 request: B does not fit 8 x 8 (7 texels wide after the
    margin); in floor(8 sqrt(2)) = 11, B takes (1, 1) and A (7 x 6) fits
    neither (1, 8, 10, 3) nor (9, 1, 2, 7); the next size is floor(8 *
    1.41421354^2) = 15 (the float sqrt(2) lies below sqrt(2)). There B cuts
    (1, 1, 14, 14) into (1, 8, 14, 7) + (1, 1, 14, 7), both 42 texels too
    large, takes the first, cut into (9, 8, 6, 7) + (1, 8, 8, 7): B at
    (1, 8); A fits (9, 8, 6, 7) exactly; D cuts (1, 1, 14, 7) into
    (6, 1, 9, 7) + (1, 1, 5, 7), the latter into (1, 5, 5, 3) + (1, 1, 5, 4):
    D at (1, 1), and the free list starts with (6, 1, 9, 7)."""
    kf = _chart_images()
    (bx, by), (ax, ay), (dx, dy), (fx, fy) = CHART_LAYOUT[size]
    tex = np.zeros((size, size, 3), np.uint8)

    def put(img, x0, x1, y0, y1, X, Y):
        p = np.pad(img[y0:y1 + 1, x0:x1 + 1], ((1, 1), (1, 1), (0, 0)), mode="edge")
        tex[Y:Y + p.shape[0], X:X + p.shape[1]] = p
    put(kf[1], 2, 7, 1, 5, bx, by)  # B
    put(kf[0], 1, 4, 0, 4, ax, ay)  # A
    put(kf[0], 5, 7, 1, 2, dx, dy)  # D
    n = _f32(size)
    uv = []
    for kfi, vids, x0, y0, X, Y in ((1, (1, 4, 5), 2, 1, bx, by), (0, (0, 1, 3, 4), 1, 0, ax, ay),
                                    (0, (6, 7, 8), 5, 1, dx, dy)):
        for v in vids:
            x, y = CHART_PROJ[(v, kfi)]
            lx = _f32(_f32(x) - _f32(x0)) + _f32(1)
            ly = _f32(_f32(y) - _f32(y0)) + _f32(1)
            uv.append((_f32(_f32(lx + _f32(X)) + _f32(0.5)) / n,
                       _f32(1) - _f32(_f32(_f32(ly + _f32(Y)) + _f32(0.5)) / n)))
    uv += [(_f32(_f32(fx) + _f32(0.5)) / n, _f32(1) - _f32(_f32(fy) + _f32(0.5)) / n)] * 3
    uvi = [(3, 5, 6), (3, 6, 4), (0, 1, 2), (10, 11, 12), (7, 8, 9)]
    return tex, np.array(uv, np.float32), np.array(uvi, np.int32)


def _map_cpp(tmp_path, imgs, vi, kfs_of_face, proj, uv_type, tex_w, tex_h):
    """TextureMapping of the C++ drop-in through ugu_stats_tool map: one
    record per seen face (area 1)."""
    h, w = imgs[0].shape[:2]
    (tmp_path / "meta.txt").write_text(
        f"{len(imgs)} {w} {h} {len(vi)} 6 {int(uv_type)} {tex_w} {tex_h}\n")
    (tmp_path / "ids.txt").write_text(" ".join(str(i) for i in range(len(imgs))))
    for i, im in enumerate(imgs):
        (tmp_path / f"kf{i}.raw").write_bytes(im.tobytes())
    (tmp_path / "vi.txt").write_text("\n".join(" ".join(map(str, f)) for f in vi) + "\n")
    lines = []
    for f, k in zip(vi, kfs_of_face):
        if k is None:
            lines.append("0")
            continue
        xy = " ".join(f"{float(np.float32(c)):.9g}" for v in f for c in proj[(v, k)])
        lines += ["1", f"{k} 1.0 0.5 1.0 {xy}"]
    (tmp_path / "faces.txt").write_text("\n".join(lines) + "\n")
    # the result is the last line (the library logs before it)
    ok, tw, th = map(int, _tool("map", tmp_path).splitlines()[-1].split())
    if not ok:
        return None
    return (np.fromfile(tmp_path / "tex.u8", np.uint8).reshape(th, tw, 3),
            np.fromfile(tmp_path / "uv.f32", np.float32).reshape(-1, 2),
            np.fromfile(tmp_path / "uvi.i32", np.int32).reshape(-1, 3))


def _map_py(imgs, vi, kfs_of_face, proj, uv_type, tex_w, tex_h):
    from ugu_amd import Mesh
    kfs = []
    for i, im in enumerate(imgs):
        k = V.Keyframe()
        k.id, k.color = i, im
        kfs.append(k)
    nv = max(max(f) for f in vi) + 1
    mesh = Mesh(np.zeros((nv, 3), np.float32), np.array(vi, np.int32))
    info = V.VisibilityInfo(num_vertices=nv, num_faces=len(vi))
    for fi, (f, k) in enumerate(zip(vi, kfs_of_face)):
        if k is None:
            continue
        r = V.FaceInfoPerKeyframe(k, fi)
        r.area, r.viewing_angle, r.distance = np.float32(1), np.float32(0.5), np.float32(1)
        r.projected_tri = np.array([proj[(v, k)] for v in f], np.float32)
        info.face_info_list[fi].visible_keyframes.append(r)
    info.CalcStatFaceInfo()
    opt = T.TextureMappingOption(uv_type=uv_type, tex_w=tex_w, tex_h=tex_h)
    if not T.TextureMapping(kfs, info, mesh, opt):
        return None
    return (mesh.materials[0], np.asarray(mesh.uv, np.float32).reshape(-1, 2),
            np.asarray(mesh.uv_indices, np.int32).reshape(-1, 3))


def _map(impl, tmp_path, *args):
    return _map_cpp(tmp_path, *args) if impl == "cpp" else _map_py(*args)


@pytest.mark.parametrize("impl", ["cpp", "python"])
@pytest.mark.parametrize("tex,size", [(16, 16), (8, 15)])
def test_simple_charts_by_hand(impl, tex, size, tmp_path):
    got = _map(impl, tmp_path, _chart_images(), CHART_VI, CHART_KF, CHART_PROJ,
               T.TexturingOutputUvType.kGenerateSimpleCharts, tex, tex)
    assert got is not None
    tex_got, uv, uvi = got
    exp_tex, exp_uv, exp_uvi = _expected_charts(size)
    assert tex_got.shape == exp_tex.shape and np.array_equal(tex_got, exp_tex)
    assert np.array_equal(uvi, exp_uvi)
    assert np.array_equal(uv, exp_uv)


@pytest.mark.parametrize("impl", ["cpp", "python"])
def test_simple_charts_rejects(impl, tmp_path):
    # a corner outside the keyframe image (the reference reads outside it)
    proj = dict(CHART_PROJ)
    proj[(7, 0)] = (7.5, 1.0)
    assert _map(impl, tmp_path, _chart_images(), CHART_VI, CHART_KF, proj,
                T.TexturingOutputUvType.kGenerateSimpleCharts, 16, 16) is None
    # patches that do not fit 8192 x 8192 (a 4 x 4 request grows as
    # floor(4 sqrt(2)^k) and stops past 8192): one chart is already too wide
    imgs = [np.zeros((4, 9000, 3), np.uint8)]
    vi = [(0, 1, 2)]
    far = {(0, 0): (0.0, 0.0), (1, 0): (8999.0, 0.0), (2, 0): (0.0, 3.0)}
    assert _map(impl, tmp_path, imgs, vi, [0], far,
                T.TexturingOutputUvType.kGenerateSimpleCharts, 4, 4) is None


def _grid_case(seed):
    """A 10 x 10 grid of quads (two triangles each, one winding), faces
    assigned to keyframes -1..2 by 3 x 3 blocks with a few changed at
    random; every vertex projected into each keyframe at x, y = 3 i + 2,
    3 j + 2 plus a per-keyframe offset, some at quarter positions."""
    rng = np.random.default_rng(seed)
    n = 10
    vi, kfs = [], []
    for j in range(n):
        for i in range(n):
            v00, v10 = j * (n + 1) + i, j * (n + 1) + i + 1
            v01, v11 = v00 + n + 1, v10 + n + 1
            for f in ((v00, v01, v11), (v00, v11, v10)):
                k = (i // 3 + j // 3) % 4 - 1
                if rng.random() < 0.1:
                    k = int(rng.integers(-1, 3))
                vi.append(f)
                kfs.append(None if k < 0 else k)
    proj = {}
    for v in range((n + 1) ** 2):
        i, j = v % (n + 1), v // (n + 1)
        for k in range(3):
            q = 0.25 if (v + k) % 3 == 0 else 0.0
            proj[(v, k)] = (3 * i + 2 + 4 * k + q, 3 * j + 2 + 2 * k + q)
    imgs = [rng.integers(0, 256, (48, 64, 3), dtype=np.uint8) for _ in range(3)]
    return imgs, vi, kfs, proj


@pytest.mark.parametrize("impl", ["cpp", "python"])
@pytest.mark.parametrize("seed", [1, 2])
def test_simple_charts_properties(impl, seed, tmp_path):
    """What the atlas must satisfy whatever its layout: the texel under each
    seen corner's uv is the keyframe pixel under its projected position (the
    patch copies the keyframe over the chart's bbox and the uv moves the
    corner with it); two faces that share an edge and a keyframe share the uv
    entries of the edge's vertices, faces with different keyframes share
    none; unseen faces share the last three entries; a grown texture has the
    size floor(128 * sqrt(2)^k)."""
    imgs, vi, kfs, proj = _grid_case(seed)
    got = _map(impl, tmp_path, imgs, vi, kfs, proj,
               T.TexturingOutputUvType.kGenerateSimpleCharts, 128, 128)
    assert got is not None
    tex, uv, uvi = got
    th, tw = tex.shape[:2]
    sizes = {int(128 * float(np.float32(np.sqrt(2.0))) ** k) for k in range(12)}
    assert th == tw and tw in sizes
    n_uv = len(uv)
    assert ((uv >= 0) & (uv <= 1)).all()
    checked = 0
    for f, (face, k) in enumerate(zip(vi, kfs)):
        if k is None:
            assert tuple(uvi[f]) == (n_uv - 3, n_uv - 2, n_uv - 1)
            continue
        assert (uvi[f] < n_uv - 3).all()
        for j, v in enumerate(face):
            u, w = uv[uvi[f, j]]
            tx = float(u) * tw - 0.5
            ty = (1.0 - float(w)) * th - 0.5
            x, y = proj[(v, k)]
            assert abs(tx - round(tx) - (x - round(x))) < 1e-3
            assert np.array_equal(tex[round(ty), round(tx)], imgs[k][round(y), round(x)])
            checked += 1
    assert checked > 400
    edges = {}
    for f, face in enumerate(vi):
        for a in range(3):
            edges.setdefault(frozenset((face[a], face[(a + 1) % 3])), []).append(f)
    pairs = same = 0
    for e, fs in edges.items():
        if len(fs) != 2:
            continue
        f, g = fs
        if kfs[f] is None or kfs[g] is None:
            continue
        pairs += 1
        for v in e:
            a = uvi[f, list(vi[f]).index(v)]
            b = uvi[g, list(vi[g]).index(v)]
            if kfs[f] == kfs[g]:
                assert a == b
                same += 1
            else:
                assert a != b
    assert pairs > 100 and same > 100


def test_triangle_grid_by_hand():
    # 3 faces in 16 x 16: 2 rectangles of 128 texels, s from floor(sqrt(128))
    # = 11 down to 5, the first side with (16 // (s + 4)) * (16 // (s + 3))
    # >= 2 (1 x 2 at s = 5); pitch 9 x 8, one rectangle per row
    g = T.triangle_grid(3, 16, 16)
    assert g == (5, 9, 8, 1)
    assert T.triangle_corners(g, 0) == [(2, 0), (6, 0), (6, 4)]
    assert T.triangle_corners(g, 1) == [(0, 1), (0, 5), (4, 5)]
    assert T.triangle_corners(g, 2) == [(2, 8), (6, 8), (6, 12)]
    # 1000 faces in 1024 x 1024: 500 rectangles of 2097 texels, s from
    # floor(sqrt(2097)) = 45: 20 * 21 = 420 rectangles fit at 45, 21 * 21 at
    # 44, 21 * 22 at 43, 22 * 22 at 42, 22 * 23 = 506 at 41; 22 per row, so
    # face 47 (odd) takes the lower-left half of rectangle 23 at (1, 1)
    assert T.triangle_grid(1000, 1024, 1024) == (41, 45, 44, 22)
    assert T.triangle_corners(T.triangle_grid(1000, 1024, 1024), 47) == \
        [(45, 44 + 1), (45, 44 + 41), (45 + 40, 44 + 41)]
    # the reference gives up: fewer than 6 texels, the side below 3
    assert T.triangle_grid(10, 4, 4) is None
    assert T.triangle_grid(40, 20, 20) is None
    assert T.triangle_grid(0, 64, 64) is None


@pytest.mark.parametrize("impl", ["cpp", "python"])
def test_simple_charts_shared_directed_edge(impl, tmp_path):
    """Faces 1 and 2 both hold the directed edge 0 -> 1 (inconsistent
    winding): its entry is the sum of their ids + 1, (1 + 1) + (2 + 1) = 5, as
    the reference's setFromTriplets sums duplicates. Face 0 (edge 1 -> 0) then
    finds "face 4" across that edge -- an isolated face -- and grows the chart
    {0, 4}; faces 1, 2 and 3 are charts of their own (face 0's edges are
    cleared before they are reached). Uv entries: 6 + 3 + 3 + 3 chart
    vertices and the 3 shared unseen entries = 18. Keeping the first or the
    last face of the edge instead would pair face 0 with face 1 or 2 (16)."""
    vi = [(1, 0, 4), (0, 1, 2), (0, 1, 3), (5, 6, 7), (8, 9, 10)]
    kfs = [0] * 5
    proj = {(v, 0): (1.0 + v % 6, 1.0 + v // 6) for v in range(11)}
    imgs = [np.random.default_rng(4).integers(0, 256, (CH, CW, 3), dtype=np.uint8)]
    got = _map(impl, tmp_path, imgs, vi, kfs, proj,
               T.TexturingOutputUvType.kGenerateSimpleCharts, 32, 32)
    assert got is not None
    tex, uv, uvi = got
    assert len(uv) == 18
    # face 0 and face 4 share one chart, whose patch spans the bbox of both:
    # x 1..5, y 1..2 -> 4 x 7 = 28 texels with the border. By patch area the
    # charts are {3} (x 1..6, y 1..2: 32), {0, 4} (28), {2} (x 1..4: 18),
    # {1} (x 1..3: 15), so {0, 4} holds uv entries 3..8 (vertex ids in order)
    assert set(uvi[0]) | set(uvi[4]) == set(range(3, 9))
    assert tuple(uvi[3]) == (0, 1, 2) and set(uvi[2]) == set(range(9, 12))
    assert set(uvi[1]) == set(range(12, 15))
