"""GPU: the C++ texturing API (ugu::VisibilityTester, ugu::TextureMapping in
ugu_amd/libugu_rasterizer.so) driven like UGU's texturing example by
ugu_amd/cpp/test_texturing_api on the bundled Spot mesh -- LoadObj, keyframes
rendered by the C++ Rasterizer, Test(keyframes, info) face-wise and
vertex-wise, TextureMapping on the mesh uv -- against the Python pipeline
run on the same keyframe images. Records, the selection statistics and the
textures are bit-exact; the Python pipeline itself is checked against the
oracle in test_gpu_visibility.py and test_gpu_texturing.py. Also runs the
driver's checks of the reference's error conventions."""
import copy
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O

from ugu_amd import PinholeCamera, io as uio, look_at_c2w, texturing as T, visibility as V

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "tests", "golden", "data")
DRIVER = os.path.join(ROOT, "ugu_amd", "cpp", "test_texturing_api")
W, H, NCAM = 256, 192, 6


def _cameras():
    cams = []
    for i in range(NCAM):
        a = 2 * np.pi * i / NCAM
        pos = np.array([2.0 * np.sin(a), 0.5, 0.35 - 2.0 * np.cos(a)])
        cams.append(PinholeCamera(W, H, look_at_c2w(pos, (0, 0, 0.35), (0, 1, 0)), (127.5, 95.5),
                                  (220.0, 220.0)))
    return cams


def _vertex_arrays(info):
    rec, stat = [], []
    for i, vi in enumerate(info.vertex_info_list):
        for k in vi.visible_keyframes:
            rec.append([i, k.kf_id, *k.color, k.intensity, *k.projected_pos, k.viewing_angle,
                        k.distance])
        if vi.visible_keyframes:
            stat.append([*vi.mean_color, *vi.mean_distance_color, vi.min_viewing_angle_id,
                         vi.min_distance_id])
        else:
            stat.append([0, 0, 0, 0, 0, 0, -1, -1])
    return np.array(rec, np.float32).reshape(-1, 10), np.array(stat, np.float32)


def _face_arrays(info):
    rec, stat = [], []
    for fi in info.face_info_list:
        for k in fi.visible_keyframes:
            rec.append([k.face_id, k.kf_id, k.area, k.viewing_angle, k.distance,
                        *np.asarray(k.projected_tri, np.float32).reshape(-1)])
        if fi.visible_keyframes:
            stat.append([fi.max_area_index, fi.min_viewing_angle_index, fi.min_distance_index])
        else:
            stat.append([-1, -1, -1])
    return np.array(rec, np.float32).reshape(-1, 11), np.array(stat, np.int32)


def _same(a, b):
    # bit-exact, with NaN == NaN and +inf == +inf
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.fixture(scope="module")
def cpp_run(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("cpp_tex"))
    for n in ("spot.mtl", "spot_texture.png"):
        shutil.copy(os.path.join(DATA, n), d)
    src = open(os.path.join(DATA, "spot_quadrangulated.obj")).read()
    obj = os.path.join(d, "spot_q.obj")
    open(obj, "w").write("mtllib spot.mtl\nusemtl material000\n" + src)
    cams = _cameras()
    rows = [[c.width, c.height, *c.focal_length, *c.principal_point, *c.c2w.reshape(-1)]
            for c in cams]
    np.array(rows, np.float64).tofile(os.path.join(d, "cams.f64"))
    out = os.path.join(d, "out")
    os.makedirs(out)
    print(f"[cpp-texturing] driver on {obj}", flush=True)
    res = subprocess.run([DRIVER, obj, os.path.join(d, "cams.f64"), str(NCAM), out],
                         capture_output=True, text=True, timeout=120,
                         env=dict(os.environ, UGU_HIP_TUNE_CACHE=os.path.join(d, "tiling.txt")))
    print(res.stdout[-2000:], res.stderr[-2000:], flush=True)
    mesh, ok = uio.load_obj(obj)
    assert ok
    kfs = []
    for i, cam in enumerate(cams):
        p = os.path.join(out, f"kf{i}_")
        kfs.append(V.Keyframe(i, np.fromfile(p + "color.u8", np.uint8).reshape(H, W, 3),
                              np.fromfile(p + "depth.f32", np.float32).reshape(H, W),
                              np.fromfile(p + "mask.u8", np.uint8).reshape(H, W), cam))
    return res, out, mesh, kfs


def test_driver_checks(gpu_available, cpp_run):
    res = cpp_run[0]
    assert res.returncode == 0, res.stdout + res.stderr
    assert "FAILED" not in res.stdout
    assert res.stdout.count("CHECK ") >= 20


@pytest.mark.parametrize("facewise", [True, False])
def test_cpp_matches_python_pipeline(gpu_available, cpp_run, facewise):
    res, out, mesh, kfs = cpp_run
    assert res.returncode == 0, res.stdout + res.stderr
    vt = V.VisibilityTester()
    vt.set_data(mesh.vertices, mesh.normals, mesh.vertex_indices)
    assert vt.PrepareData()
    info = V.VisibilityInfo(mesh)
    assert vt.Test(info, facewise=facewise, keyframes=kfs)
    p = os.path.join(out, "fw_" if facewise else "vw_")
    vrec, vstat = _vertex_arrays(info)
    frec, fstat = _face_arrays(info)
    c_vrec = np.fromfile(p + "vrec.f32", np.float32).reshape(-1, 10)
    c_frec = np.fromfile(p + "frec.f32", np.float32).reshape(-1, 11)
    c_vstat = np.fromfile(p + "vstat.f32", np.float32).reshape(-1, 8)
    c_fstat = np.fromfile(p + "fstat.i32", np.int32).reshape(-1, 3)
    print(f"[cpp-texturing] facewise={facewise}: vertex records {len(vrec)} "
          f"face records {len(frec)}", flush=True)
    assert len(vrec) > len(mesh.vertices) and len(frec) > len(mesh.vertex_indices)
    assert _same(c_vrec, vrec)
    assert _same(c_frec, frec)
    assert _same(c_vstat, vstat)
    assert np.array_equal(c_fstat, fstat)
    for c, crit in enumerate([T.ViewSelectionCriteria.kMaxArea,
                              T.ViewSelectionCriteria.kMinViewingAngle]):
        mm = copy.copy(mesh)  # TextureMapping replaces the materials
        opt = T.TextureMappingOption(criteria=crit, tex_w=512,
                                     uv_type=T.TexturingOutputUvType.kUseOriginalMeshUv, tex_h=512)
        assert T.TextureMapping(kfs, info, mm, opt)
        tex = np.fromfile(p + f"tex{c}.u8", np.uint8).reshape(512, 512, 3)
        assert np.array_equal(tex, mm.materials[0]), crit.name
    # default option (kGenerateSimpleTile): the C++ tiled texture and uvs
    # equal the Python host's
    mm = copy.copy(mesh)
    assert T.TextureMapping(kfs, info, mm, T.TextureMappingOption())
    tex = np.fromfile(p + "tile_tex.u8", np.uint8).reshape(mm.materials[0].shape)
    assert np.array_equal(tex, mm.materials[0])
    uv = np.fromfile(p + "tile_uv.f32", np.float32).reshape(-1, 2)
    assert np.array_equal(uv, np.asarray(mm.uv, np.float32).reshape(-1, 2))


def _atlas(out, p, name):
    h, w = np.fromfile(os.path.join(out, p + name + "size.i32"), np.int32)
    tex = np.fromfile(os.path.join(out, p + name + "tex.u8"), np.uint8).reshape(h, w, 3)
    uv = np.fromfile(os.path.join(out, p + name + "uv.f32"), np.float32).reshape(-1, 2)
    uvi = np.fromfile(os.path.join(out, p + name + "uvi.i32"), np.int32).reshape(-1, 3)
    return tex, uv, uvi


def _corner_texels(tex, uv, uvi):
    """Texel under every face corner's uv (nearest texel centre)."""
    h, w = tex.shape[:2]
    q = uv[uvi].astype(np.float64)
    x = np.rint(q[..., 0] * w - 0.5).astype(np.int64).clip(0, w - 1)
    y = np.rint((1.0 - q[..., 1]) * h - 0.5).astype(np.int64).clip(0, h - 1)
    return tex[y, x]


@pytest.mark.parametrize("facewise", [True, False])
def test_cpp_atlas_generators(gpu_available, cpp_run, facewise):
    """kGenerateSimpleTriangles and kGenerateSimpleCharts through the C++
    drop-in on Spot. Triangles: texture and uvs equal the Python host's and
    oracle_texture_map of the faces into their grid halves. Charts: each
    seen corner's uv lands on its projected position in the chosen keyframe
    moved by a whole number of texels (within 0.1 texel: the face-wise test
    projects a vertex nudged by 1e-3 of its distance towards each face's
    centre, up to 0.022 pixel apart on Spot, and a chart keeps the position
    of its first face naming the vertex), the texel there is
    the keyframe pixel moved by the same amount, and the texel colours at
    every corner equal the Python host's (the two may place equal-area
    charts differently: the C++ drop-in orders them with std::sort as the
    reference does, the Python host with a stable sort)."""
    res, out, mesh, kfs = cpp_run
    assert res.returncode == 0, res.stdout + res.stderr
    vt = V.VisibilityTester()
    vt.set_data(mesh.vertices, mesh.normals, mesh.vertex_indices)
    assert vt.PrepareData()
    info = V.VisibilityInfo(mesh)
    assert vt.Test(info, facewise=facewise, keyframes=kfs)
    p = "fw_" if facewise else "vw_"
    best = T.best_keyframes(info, T.ViewSelectionCriteria.kMaxArea)
    F = len(mesh.vertex_indices)

    tex, uv, uvi = _atlas(out, p, "tri_")
    mm = copy.copy(mesh)
    opt = T.TextureMappingOption(uv_type=T.TexturingOutputUvType.kGenerateSimpleTriangles,
                                 tex_w=512, tex_h=512)
    assert T.TextureMapping(kfs, info, mm, opt)
    assert np.array_equal(tex, mm.materials[0])
    assert np.array_equal(uv, np.asarray(mm.uv, np.float32).reshape(-1, 2))
    assert np.array_equal(uvi, np.arange(3 * F).reshape(F, 3))
    grid = T.triangle_grid(F, 512, 512)
    orecs = (O.OracleTexFace * F)()
    for f, b in enumerate(best):
        r = orecs[f]
        r.keyframe = -1 if b is None else int(b.kf_id)
        if b is None:
            continue
        for j, (x, y) in enumerate(T.triangle_corners(grid, f)):
            r.src[2 * j], r.src[2 * j + 1] = (float(b.projected_tri[j][0]),
                                              float(b.projected_tri[j][1]))
            r.dst[2 * j], r.dst[2 * j + 1] = float(x), float(y)
    rt, _ = O.texture_map([k.color for k in kfs], orecs, 512, 512, 3)
    assert np.array_equal(tex, rt)

    tex, uv, uvi = _atlas(out, p, "chart_")
    h, w = tex.shape[:2]
    checked = 0
    clear = np.zeros((F, 3), bool)  # corners whose nearest texel is unambiguous
    for f, b in enumerate(best):
        if b is None:
            continue
        for j in range(3):
            x, y = (float(c) for c in b.projected_tri[j])
            tx = float(uv[uvi[f, j], 0]) * w - 0.5
            ty = (1.0 - float(uv[uvi[f, j], 1])) * h - 0.5
            ox, oy = int(np.rint(tx - x)), int(np.rint(ty - y))
            assert abs(tx - ox - x) < 0.1 and abs(ty - oy - y) < 0.1, (f, j)
            if abs(tx - np.floor(tx) - 0.5) < 0.01 or abs(ty - np.floor(ty) - 0.5) < 0.01:
                continue
            clear[f, j] = True
            px, py = int(np.rint(tx)), int(np.rint(ty))
            assert np.array_equal(tex[py, px], kfs[b.kf_id].color[py - oy, px - ox]), (f, j)
            checked += 1
    mm = copy.copy(mesh)
    opt = T.TextureMappingOption(uv_type=T.TexturingOutputUvType.kGenerateSimpleCharts,
                                 tex_w=512, tex_h=512)
    assert T.TextureMapping(kfs, info, mm, opt)
    py = _corner_texels(mm.materials[0], np.asarray(mm.uv, np.float32).reshape(-1, 2),
                        np.asarray(mm.uv_indices))
    print(f"[cpp-texturing] facewise={facewise}: charts atlas {w}x{h} (python "
          f"{mm.materials[0].shape[1]}x{mm.materials[0].shape[0]}), corners checked {checked}",
          flush=True)
    assert checked > F
    assert np.array_equal(_corner_texels(tex, uv, uvi)[clear], py[clear])
