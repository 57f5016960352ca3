"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs. face_id and mask bit-exact; colour (u8) exact; depth and
normal within 1e-5 relative; Oren-Nayar colour within 1 (libm transcendentals
differ between the device and glibc)."""
import math

import numpy as np
import pytest

import oracle as O
import tiny_scenes as TS
from ugu_amd import (ColorInterpolation, DiffuseColor, DiffuseShading, PinholeCamera,
                     Rasterizer, RendererCpuOption, ShadingNormal, look_at_c2w, scenes)

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def _oracle(mesh, cam, opt):
    rc, ref, _ = O.render(mesh, O.camera_from_abi(cam.to_abi()), opt, c2w_t=cam.c2w[:3, 3])
    assert rc == 0
    return ref


def _compare(out, ref, label="", color_tol=0):
    print(f"[parity] {label}: covered={int((ref['face_id'] >= 0).sum())}", flush=True)
    bad = np.argwhere(out["face_id"] != ref["face_id"])
    assert len(bad) == 0, f"{label}: face_id differs at {bad[:5]} " \
        f"gpu={out['face_id'][tuple(bad[0])]} ref={ref['face_id'][tuple(bad[0])]}"
    assert np.array_equal(out["mask"], ref["mask"]), f"{label}: mask differs"
    if color_tol == 0:
        bad = np.argwhere((out["color"] != ref["color"]).any(-1))
        assert len(bad) == 0, f"{label}: colour differs at {bad[:5]}"
    else:
        d = np.abs(out["color"].astype(int) - ref["color"].astype(int))
        assert d.max() <= color_tol, f"{label}: colour differs by {d.max()}"
    np.testing.assert_allclose(out["depth"], ref["depth"], rtol=RTOL, atol=0)
    np.testing.assert_allclose(out["normal"], ref["normal"], rtol=RTOL, atol=1e-7)


def _render(mesh, cam, opt, tile=(16, 16), cap=1024):
    r = Rasterizer(opt)
    r.set_tiling(*tile, bin_capacity=cap)
    r.set_mesh(mesh)
    assert r.PrepareMesh()
    r.set_camera(cam)
    return r.render_batch([cam])[0]


@pytest.fixture(scope="module")
def uv_scene():
    return scenes.scene_uv_sphere_128()


@pytest.fixture(scope="module")
def ico_scene():
    return scenes.scene_icosphere8()


@pytest.mark.parametrize("tile", [(8, 8), (16, 16), (32, 8)])
def test_uv_sphere_vertex_color(gpu_available, uv_scene, tile):
    mesh, cam = uv_scene
    opt = RendererCpuOption(diffuse_color=DiffuseColor.kVertex)
    out = _render(mesh, cam, opt, tile, cap=64)
    _compare(out, _oracle(mesh, cam, opt), f"uv128 tile{tile}")
    # closed form: silhouette disk of radius f R / sqrt(d^2 - R^2)
    rad = 500.0 / math.sqrt(8.0)
    assert abs((out["mask"] > 0).sum() / (math.pi * rad * rad) - 1.0) < 0.01


def test_reference_api_render(gpu_available, uv_scene):
    from ugu_amd import Image1b, Image1f, Image1i, Image1w, Image3b, Image3f
    mesh, cam = uv_scene
    opt = RendererCpuOption(diffuse_color=DiffuseColor.kVertex,
                            diffuse_shading=DiffuseShading.kLambertian)
    r = Rasterizer(opt)
    r.set_mesh(mesh)
    assert r.PrepareMesh()
    r.set_camera(cam)
    c, d, n, m, f = Image3b(), Image1f(), Image3f(), Image1b(), Image1i()
    assert r.Render(c, d, n, m, f)
    out = dict(color=c.data, depth=d.data, normal=n.data, mask=m.data, face_id=f.data)
    ref = _oracle(mesh, cam, opt)
    _compare(out, ref, "Render() lambertian")
    # single-output entry points agree with Render
    f2 = Image1i()
    assert r.RenderFaceId(f2) and np.array_equal(f2.data, f.data)
    m2 = Image1b()
    assert r.RenderMask(m2) and np.array_equal(m2.data, m.data)
    w = Image1w()
    assert r.RenderDepthW(w)
    assert np.array_equal(w.data, O.depth_to_u16(ref["depth"]))


@pytest.mark.parametrize("variant", ["bilinear", "nn", "face_lambert", "noculling"])
def test_icosphere_1080p(gpu_available, ico_scene, variant):
    mesh, cam = ico_scene
    kw = dict(diffuse_color=DiffuseColor.kTexture, interp=ColorInterpolation.kBilinear)
    if variant == "nn":
        kw["interp"] = ColorInterpolation.kNn
    if variant == "face_lambert":
        kw.update(shading_normal=ShadingNormal.kFace, diffuse_shading=DiffuseShading.kLambertian)
    if variant == "noculling":
        kw.update(backface_culling=False)
    opt = RendererCpuOption(**kw)
    out = _render(mesh, cam, opt, (16, 16))
    _compare(out, _oracle(mesh, cam, opt), f"ico8 1080p {variant}")


def test_oren_nayar(gpu_available, uv_scene):
    mesh, cam = uv_scene
    opt = RendererCpuOption(diffuse_color=DiffuseColor.kVertex,
                            diffuse_shading=DiffuseShading.kOrenNayar)
    out = _render(mesh, cam, opt)
    _compare(out, _oracle(mesh, cam, opt), "oren-nayar", color_tol=1)


def test_multiview_batched_equals_per_view_and_oracle(gpu_available, ico_scene):
    import torch
    mesh, _ = ico_scene
    cams = scenes.multiview_cameras(64)
    opt = RendererCpuOption(diffuse_color=DiffuseColor.kTexture)
    r = Rasterizer(opt)
    r.set_tiling(16, 16, 512)
    r.set_mesh(mesh)
    assert r.PrepareMesh()
    batched = r.render_batch(cams)
    # per-view loop through the single-view entry point
    for i in (0, 1, 2, 17, 63):
        r.set_camera(cams[i])
        single = r.render_batch([cams[i]])[0]
        for k in single:
            assert np.array_equal(single[k], batched[i][k]), f"view {i} {k}: batched != per-view"
    # every view bitwise equal across tile shapes, and a few views vs the oracle
    r.set_tiling(32, 8, 512)
    other = r.render_batch(cams)
    for i in range(64):
        for k in batched[i]:
            assert np.array_equal(other[i][k], batched[i][k]), f"view {i} {k}: tile dependence"
    for i in (0, 2, 5):
        _compare(batched[i], _oracle(mesh, cams[i], opt), f"ring view {i}")
    # device-memory path (graph replay) gives the same bits
    bufs = [{"face_id": torch.empty((c.height, c.width), dtype=torch.int32, device="cuda"),
             "depth": torch.empty((c.height, c.width), dtype=torch.float32, device="cuda")}
            for c in cams]
    for _ in range(3):
        r.render_device(cams, bufs)
        for i in (0, 31, 63):
            assert np.array_equal(bufs[i]["face_id"].cpu().numpy(), batched[i]["face_id"])
            assert np.array_equal(bufs[i]["depth"].cpu().numpy(), batched[i]["depth"])
    # the bound entry: same bits; a pose changed on the binding is rendered
    binding = r.bind_device(cams, bufs)
    for b in bufs:
        b["face_id"].fill_(-7)
    r.render_bound(binding)
    for i in (0, 31, 63):
        assert np.array_equal(bufs[i]["face_id"].cpu().numpy(), batched[i]["face_id"])
    assert (cams[5].width, cams[5].height) == (cams[10].width, cams[10].height)
    binding.set_camera(5, cams[10])
    r.render_bound(binding)
    assert np.array_equal(bufs[5]["face_id"].cpu().numpy(), batched[10]["face_id"])
    assert np.array_equal(bufs[5]["depth"].cpu().numpy(), batched[10]["depth"])
    with pytest.raises(ValueError):
        binding.set_camera(5, cams[1])  # another image size than the bound tensors


@pytest.mark.parametrize("tile,cull", [((16, 16), True), ((8, 8), False), ((32, 8), True)])
def test_layered_occlusion_multichunk(gpu_available, tile, cull):
    """Thousands of faces per tile in shuffled id order: faces skipped as
    hidden behind the tile's current depths must not change any pixel."""
    tris = TS.layered_grid_scene(seed=3)
    mesh = TS.mesh_from_pixels(tris, seed=3)
    cam = TS.exact_camera(96, 64)
    opt = RendererCpuOption(diffuse_color=DiffuseColor.kVertex, backface_culling=cull)
    out = _render(mesh, cam, opt, tile, cap=8192)
    _compare(out, _oracle(mesh, cam, opt), f"layered tile{tile} cull={cull}")


@pytest.mark.parametrize("seed", list(range(12)))
def test_tiny_random_exact_ties(gpu_available, seed):
    w, h = [(8, 8), (16, 16), (32, 32), (13, 29)][seed % 4]
    tris = TS.random_triangles(6 + seed * 3, w, h, seed)
    if seed % 3 == 0:
        tris = np.concatenate([tris, tris[:3]], 0)   # duplicated faces: exact ties
    mesh = TS.mesh_from_pixels(tris, seed=seed)
    TS.add_texture(mesh, 7, seed)
    cam = TS.exact_camera(w, h)
    for opt in (RendererCpuOption(diffuse_color=DiffuseColor.kVertex, backface_culling=bool(seed % 2)),
                RendererCpuOption(diffuse_color=DiffuseColor.kTexture, interp=ColorInterpolation.kNn,
                                  shading_normal=ShadingNormal.kFace)):
        for tile in ((8, 8), (32, 8)):
            out = _render(mesh, cam, opt, tile, cap=4)
            _compare(out, _oracle(mesh, cam, opt), f"tiny seed{seed} {w}x{h} tile{tile}")


def test_degenerate_vertices(gpu_available):
    """Vertices exactly at camera z = 0 (infinite/NaN projections), behind the
    camera, zero-area and sliver faces: the slow sequential path must
    reproduce the reference's update rule."""
    rng = np.random.default_rng(5)
    w, h = 24, 20
    tris = TS.random_triangles(40, w, h, 7)
    mesh = TS.mesh_from_pixels(tris, seed=1)
    v = mesh.vertices.copy()
    v[0] = (0.0, 0.0, 0.0)          # at the camera centre
    v[4] = (3.0, 2.0, 0.0)          # on the z = 0 plane
    v[7, 2] = -1.0                  # behind the camera
    v[10] = v[9]                    # zero-area face 3
    v[13] = v[12] + np.float32(1e-6)  # sliver face 4
    mesh.set_vertices(v)
    mesh.CalcFaceNormal()
    cam = TS.exact_camera(w, h)
    for cull in (True, False):
        opt = RendererCpuOption(diffuse_color=DiffuseColor.kVertex, backface_culling=cull)
        out = _render(mesh, cam, opt, (8, 8), cap=8)
        ref = _oracle(mesh, cam, opt)
        assert np.array_equal(out["face_id"], ref["face_id"])
        assert np.array_equal(out["mask"], ref["mask"])
        # NaN depths compare equal position-wise
        assert np.array_equal(np.isnan(out["depth"]), np.isnan(ref["depth"]))
        ok = ~np.isnan(ref["depth"])
        np.testing.assert_allclose(out["depth"][ok], ref["depth"][ok], rtol=RTOL)


def test_huge_screen_coordinates(gpu_available):
    """Corners projected far outside the image, around and beyond 2^30 and
    2^31 pixels (finite): the exact bbox path (std::min, x86 conversions
    that saturate to INT32_MIN) next to the faces on the hardware min / max
    path, for every tile shape."""
    w, h = 40, 24
    big = [2.0 ** 30 - 64.0, 2.0 ** 30, 2.0 ** 30 + 64.0, 2.0 ** 31, 2.0 ** 31 + 2.0 ** 9, 1e12]
    tris = list(TS.random_triangles(30, w, h, 11))
    rng = np.random.default_rng(3)
    for i, b in enumerate(big):
        for sx, sy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
            cx, cy = rng.uniform(2, w - 2), rng.uniform(2, h - 2)
            z = float(rng.choice([1.0, 2.0]))
            far = (cx + sx * b, cy + sy * b)
            side = (cx + sy * 3.0, cy + sx * 3.0)
            tris.append([[cx, cy, z], [far[0], far[1], z], [side[0], side[1], z]])
    mesh = TS.mesh_from_pixels(np.array(tris), seed=4)
    cam = TS.exact_camera(w, h)
    for cull in (False, True):
        opt = RendererCpuOption(diffuse_color=DiffuseColor.kVertex, backface_culling=cull)
        ref = _oracle(mesh, cam, opt)
        for tile in ((8, 8), (16, 16), (32, 8)):
            out = _render(mesh, cam, opt, tile, cap=64)
            _compare(out, ref, f"huge coordinates cull={cull} tile{tile}")


@pytest.mark.parametrize("flip,cull", [(False, False), (True, True), (True, False)])
def test_subnormal_depth_sentinel(gpu_available, flip, cull):
    """Depth 0 (< FLT_MIN) written by a face lets every later face overwrite
    the pixel (rasterizer.cc:229): handled by the sequential slow path."""
    mesh, cam = TS.subnormal_depth_scene(flip=flip)
    opt = RendererCpuOption(diffuse_color=DiffuseColor.kVertex, backface_culling=cull)
    ref = _oracle(mesh, cam, opt)
    for tile in ((8, 8), (16, 16), (32, 8)):
        out = _render(mesh, cam, opt, tile, cap=16)
        _compare(out, ref, f"subnormal flip={flip} cull={cull} tile{tile}")


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_subnormal_depths_in_dense_tiles(gpu_available, seed):
    """Hundreds of faces per tile, many of them with a corner at a subnormal
    depth (pixel depth 0 < FLT_MIN): the slow path's two passes over a large
    bin against the oracle's sequential loop, for every tile shape."""
    mesh, cam = TS.random_subnormal_layers(600, 40, 24, seed)
    for cull in (False, True):
        opt = RendererCpuOption(diffuse_color=DiffuseColor.kVertex, backface_culling=cull)
        ref = _oracle(mesh, cam, opt)
        for tile in ((8, 8), (16, 16), (32, 8)):
            out = _render(mesh, cam, opt, tile, cap=64)
            _compare(out, ref, f"dense subnormal seed{seed} cull={cull} tile{tile}")


@pytest.mark.parametrize("seed", [5, 6])
def test_small_faces_at_tile_edges(gpu_available, seed):
    """Thousands of faces whose bboxes hold at most a few pixel centres, around
    tile boundaries, edges through pixel centres, both windings: setup tests
    such faces at their pixel centres and bins them only into the tiles where
    they may cover one, so every tile shape must still match the oracle."""
    w, h = 72, 40
    mesh = TS.mesh_from_pixels(TS.small_faces_at_tile_edges(4000, w, h, seed), seed=seed)
    cam = TS.exact_camera(w, h)
    for cull in (False, True):
        opt = RendererCpuOption(diffuse_color=DiffuseColor.kVertex, backface_culling=cull)
        ref = _oracle(mesh, cam, opt)
        for tile in ((8, 8), (16, 16), (32, 8)):
            out = _render(mesh, cam, opt, tile, cap=64)
            _compare(out, ref, f"small faces seed{seed} cull={cull} tile{tile}")


def test_offscreen_and_empty(gpu_available, uv_scene):
    mesh, cam = uv_scene
    c2 = scenes.front_camera(64, 48, 500.0, 3.0)
    c2.c2w[0, 3] = 50.0   # looking away from the sphere
    c2.set_c2w(c2.c2w)
    opt = RendererCpuOption(diffuse_color=DiffuseColor.kVertex)
    out = _render(mesh, c2, opt)
    assert (out["face_id"] == -1).all() and (out["depth"] == 0).all()
    assert (out["color"] == 0).all() and (out["mask"] == 0).all()


def test_capacity_growth_without_reallocation(gpu_available, uv_scene):
    """A capacity grown after an overflow must not replay a graph captured
    with the old capacity (the bins buffer is already large enough, so no
    reallocation marks the graph stale)."""
    mesh, cam = uv_scene
    opt = RendererCpuOption(diffuse_color=DiffuseColor.kVertex)
    ref = _oracle(mesh, cam, opt)
    r = Rasterizer(opt)
    r.set_mesh(mesh)
    assert r.PrepareMesh()
    r.set_tiling(16, 16, 8192)
    _compare(r.render_batch([cam])[0], ref, "cap 8192")
    for cap in (8, 64, 8):
        r.set_tiling(16, 16, cap)
        for _ in range(2):
            _compare(r.render_batch([cam])[0], ref, f"cap {cap} -> {r.get_tiling()['bin_capacity']}")


@pytest.mark.parametrize("interp,shading", [(ColorInterpolation.kBilinear, DiffuseShading.kNone),
                                            (ColorInterpolation.kNn, DiffuseShading.kLambertian)])
def test_multi_material_textures(gpu_available, interp, shading):
    """Three materials of different texture sizes with seeded per-face material
    ids: every face's texture comes from its own material (the per-face
    shading record carries the id; with more than one material the texture
    descriptor is looked up per pixel), bit-exact against the oracle; the
    same scene through the raytracer against its oracle."""
    import test_gpu_raytrace as RT
    from ugu_amd import Raytracer
    mesh, _ = scenes.scene_icosphere8(level=5, tex_size=256, seed=3)
    rng = np.random.default_rng(11)
    mesh.materials = [mesh.materials[0],
                      (rng.random((64, 96, 3)) * 255).astype(np.uint8),
                      (rng.random((200, 33, 3)) * 255).astype(np.uint8)]
    mesh.set_material_ids(rng.integers(0, 3, len(mesh.vertex_indices)))
    opt = RendererCpuOption(diffuse_color=DiffuseColor.kTexture, interp=interp,
                            diffuse_shading=shading)
    # small views: the oracle raytracer tests every face at every pixel
    cams = [PinholeCamera(w, h, look_at_c2w(pos, (0, 0, 0), (0, 1, 0)), fov_y_deg=40.0)
            for (w, h), pos in (((200, 150), (0.4, 0.3, -2.6)), ((96, 64), (-2.2, 0.5, -1.5)),
                                ((128, 128), (1.9, -0.8, 1.7)))]
    r = Rasterizer(opt)
    r.set_mesh(mesh)
    assert r.PrepareMesh()
    outs = r.render_batch(cams)
    for i, c in enumerate(cams):
        _compare(outs[i], _oracle(mesh, c, opt), f"multi-material view {i}")
    rt = Raytracer(opt)
    rt.set_mesh(mesh)
    assert rt.PrepareMesh()
    rts = rt.render_batch(cams[:2])
    for i, c in enumerate(cams[:2]):
        RT._compare(rts[i], RT._oracle(mesh, c, opt), f"multi-material raytrace view {i}")


@pytest.mark.parametrize("tile", [(8, 8), (32, 8)])
def test_large_triangles_over_many_tiles(gpu_available, tile):
    """Faces whose pixel boxes span hundreds of tiles (their tiles past the
    fourth are appended compacted across the wavefront, 64 per round) mixed
    with small faces, at several depths and both windings."""
    w, h = 200, 150
    rng = np.random.default_rng(4)
    big = []
    for k in range(40):
        c = rng.uniform([-0.2 * w, -0.2 * h], [1.2 * w, 1.2 * h], (3, 2))
        big.append(np.concatenate([c, np.full((3, 1), [1.0, 2.0, 4.0][k % 3])], 1))
    tris = np.concatenate([np.array(big), TS.random_triangles(300, w, h, 4)], 0)
    mesh = TS.mesh_from_pixels(tris, seed=4)
    cam = TS.exact_camera(w, h)
    for cull in (False, True):
        opt = RendererCpuOption(diffuse_color=DiffuseColor.kVertex, backface_culling=cull)
        out = _render(mesh, cam, opt, tile, cap=16)
        _compare(out, _oracle(mesh, cam, opt), f"large triangles tile{tile} cull={cull}")
