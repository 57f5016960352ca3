"""GPU: texture mapping on the mesh's own uv (ugu_hip_texture_map) against
oracle_texture_map -- bit-exact texture and mask -- on random overlapping
faces, and the whole pipeline on the bundled Spot mesh: keyframes rendered by
the rasterizer, the GPU visibility tester, TextureMapping; the same pipeline
on the oracle gives the same texture, which also reproduces Spot's own
texture where it was seen."""
import copy
import os
import shutil

import numpy as np
import pytest

import oracle as O
import visibility_util as U
from ugu_amd import (DiffuseColor, PinholeCamera, Rasterizer, RendererCpuOption, io as uio,
                     look_at_c2w, texturing as T, visibility as V)

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "data")


@pytest.mark.parametrize("pad", [0, 3, 5])
def test_random_faces_match_oracle(gpu_available, pad):
    rng = np.random.default_rng(pad)
    imgs = [rng.integers(0, 256, (60 + 7 * k, 80 + 5 * k, 3), dtype=np.uint8) for k in range(3)]
    n = 300
    recs = (T._abi.UguTexFace * n)()
    orecs = (O.OracleTexFace * n)()
    for i in range(n):
        kf = int(rng.integers(-1, 3))
        img = imgs[max(kf, 0)]
        src = rng.uniform([1, 1], [img.shape[1] - 2, img.shape[0] - 2], (3, 2)).reshape(-1)
        c = rng.uniform(-10, 200, 2)
        dst = (c + rng.uniform(-25, 25, (3, 2))).reshape(-1)
        for r in (recs[i], orecs[i]):
            r.keyframe = kf
            for k in range(6):
                r.src[k] = float(np.float32(src[k]))
                r.dst[k] = float(np.float32(dst[k]))
    tex, mask = T.texture_map_faces(imgs, recs, n, 192, 160, pad)
    rt, rm = O.texture_map(imgs, orecs, 192, 160, pad)
    print(f"[texturing] random pad={pad}: set texels {int((rm == 255).sum())}", flush=True)
    assert np.array_equal(mask, rm)
    assert np.array_equal(tex, rt)


@pytest.fixture(scope="module")
def spot(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("spot_tex"))
    for n in ("spot.mtl", "spot_texture.png"):
        shutil.copy(os.path.join(DATA, n), d)
    src = open(os.path.join(DATA, "spot_quadrangulated.obj")).read()
    path = os.path.join(d, "spot_q.obj")
    open(path, "w").write("mtllib spot.mtl\nusemtl material000\n" + src)
    mesh, ok = uio.load_obj(path)
    assert ok
    return mesh


def _keyframes(mesh, n=6):
    r = Rasterizer(RendererCpuOption(diffuse_color=DiffuseColor.kTexture))
    r.set_mesh(mesh)
    assert r.PrepareMesh()
    kfs = []
    for i in range(n):
        a = 2 * np.pi * i / n
        pos = np.array([2.0 * np.sin(a), 0.5, 0.35 - 2.0 * np.cos(a)])
        cam = PinholeCamera(256, 192, look_at_c2w(pos, (0, 0, 0.35), (0, 1, 0)), (127.5, 95.5),
                            (220.0, 220.0))
        out = r.render_batch([cam])[0]
        kfs.append(V.Keyframe(i, out["color"], out["depth"], out["mask"], cam))
    return kfs


@pytest.mark.parametrize("criteria", [T.ViewSelectionCriteria.kMaxArea,
                                      T.ViewSelectionCriteria.kMinViewingAngle])
def test_spot_pipeline_matches_oracle(gpu_available, spot, criteria):
    mesh = spot
    kfs = _keyframes(mesh)
    vt = V.VisibilityTester()
    vt.set_data(mesh.vertices, mesh.normals, mesh.vertex_indices)
    assert vt.PrepareData()
    info = V.VisibilityInfo(mesh)
    assert vt.Test(info, facewise=True, keyframes=kfs)
    opt = T.TextureMappingOption(criteria=criteria, tex_w=512,
                                 uv_type=T.TexturingOutputUvType.kUseOriginalMeshUv, tex_h=512)
    m_gpu = copy.copy(mesh)  # TextureMapping replaces the materials
    assert T.TextureMapping(kfs, info, m_gpu, opt)
    tex = m_gpu.materials[0]
    # oracle pipeline: oracle visibility test, same host logic, oracle texture map
    ot = U.OracleTester()
    ot.set_data(mesh.vertices, mesh.normals, mesh.vertex_indices)
    ot.PrepareData()
    oinfo = V.VisibilityInfo(mesh)
    assert ot.Test(oinfo, facewise=True, keyframes=kfs)
    best = T.best_keyframes(oinfo, criteria)
    recs, n = T.face_records(mesh, best, 512, 512)
    orecs = (O.OracleTexFace * n).from_buffer_copy(bytes(recs)[:n * 52])
    rt, rm = O.texture_map([k.color for k in kfs], orecs, 512, 512, 3)
    seen = rm == 255
    print(f"[texturing] spot criteria={criteria.name}: set texels {int(seen.sum())}", flush=True)
    assert np.array_equal(tex, rt)
    # the recovered texture matches Spot's texture where it was seen (both
    # sampled at 512^2 from the 1024^2 original, so compare to its 2x2 mean)
    ref = uio.imread(os.path.join(DATA, "spot_texture.png")).astype(np.float64)
    ref = ref.reshape(512, 2, 512, 2, 3).mean((1, 3))
    err = np.abs(tex.astype(np.float64) - ref)[seen].mean()
    assert seen.sum() > 0.3 * seen.size and err < 8.0, (seen.sum(), err)


@pytest.mark.parametrize("seed", [0, 1])
def test_simple_triangles_match_oracle(gpu_available, seed):
    """kGenerateSimpleTriangles: the texture equals oracle_texture_map of the
    seen faces' projected triangles into their grid halves (the layout
    pinned by hand in tests/test_texturing_host.py) with a 3 x 3 padding
    kernel whatever the option says; uvs at the corners' texel centres, 0
    for unseen faces."""
    from ugu_amd import Mesh
    rng = np.random.default_rng(seed)
    imgs = [rng.integers(0, 256, (40 + 9 * k, 56 + 4 * k, 3), dtype=np.uint8) for k in range(3)]
    kfs = []
    for i, im in enumerate(imgs):
        k = V.Keyframe()
        k.id, k.color = i, im
        kfs.append(k)
    F, tw, th = 23, 96, 72
    mesh = Mesh(np.zeros((30, 3), np.float32), rng.integers(0, 30, (F, 3)).astype(np.int32))
    info = V.VisibilityInfo(num_vertices=30, num_faces=F)
    chosen = []
    for f in range(F):
        k = int(rng.integers(-1, 3))
        chosen.append(k)
        if k < 0:
            continue
        r = V.FaceInfoPerKeyframe(k, f)
        r.area = np.float32(1)
        h, w = imgs[k].shape[:2]
        r.projected_tri = rng.uniform([0, 0], [w - 1, h - 1], (3, 2)).astype(np.float32)
        info.face_info_list[f].visible_keyframes.append(r)
    info.CalcStatFaceInfo()
    opt = T.TextureMappingOption(uv_type=T.TexturingOutputUvType.kGenerateSimpleTriangles,
                                 tex_w=tw, tex_h=th, padding_kernel=5)
    assert T.TextureMapping(kfs, info, mesh, opt)
    grid = T.triangle_grid(F, tw, th)
    orecs = (O.OracleTexFace * F)()
    exp_uv = np.zeros((F, 3, 2), np.float32)
    for f in range(F):
        r = orecs[f]
        r.keyframe = chosen[f]
        if chosen[f] < 0:
            continue
        src = info.face_info_list[f].visible_keyframes[0].projected_tri
        for j, (x, y) in enumerate(T.triangle_corners(grid, f)):
            r.src[2 * j], r.src[2 * j + 1] = float(src[j][0]), float(src[j][1])
            r.dst[2 * j], r.dst[2 * j + 1] = float(x), float(y)
            exp_uv[f, j] = ((np.float32(x) + np.float32(0.5)) / np.float32(tw),
                            np.float32(1) - (np.float32(y) + np.float32(0.5)) / np.float32(th))
    rt, rm = O.texture_map(imgs, orecs, tw, th, 3)
    print(f"[texturing] triangles seed={seed}: grid {grid}, set texels {int((rm == 255).sum())}",
          flush=True)
    assert np.array_equal(mesh.materials[0], rt)
    assert np.array_equal(np.asarray(mesh.uv, np.float32).reshape(F, 3, 2)[np.asarray(
        mesh.uv_indices)[:, 0] // 3], exp_uv)
    assert np.array_equal(np.asarray(mesh.uv_indices), np.arange(3 * F).reshape(F, 3))
