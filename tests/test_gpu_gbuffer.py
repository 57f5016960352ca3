"""GPU: G-buffer planes (ugu_hip_render_gbuffer) against the oracle's
restatement of the same definitions (oracle_render_gbuffer), and the Render
outputs of the G-buffer call against render_batch. Float planes within 1e-5
relative (as depth and normal), geo_id and all Render outputs exact."""
import numpy as np
import pytest

import oracle as O
import tiny_scenes as TS
from ugu_amd import (ColorInterpolation, DiffuseColor, Rasterizer, RendererCpuOption,
                     ShadingNormal, scenes)

pytestmark = pytest.mark.gpu

FLOAT_PLANES = ("bary", "uv", "pos_wld", "pos_cam", "normal_wld")


def _check(gpu, ref, label):
    print(f"[gbuffer] {label}: covered={int((ref['mask'] == 255).sum())}", flush=True)
    assert np.array_equal(gpu["face_id"], ref["face_id"]), f"{label}: face_id"
    assert np.array_equal(gpu["mask"], ref["mask"]), f"{label}: mask"
    assert np.array_equal(gpu["geo_id"], ref["geo_id"]), f"{label}: geo_id"
    for k in FLOAT_PLANES:
        np.testing.assert_allclose(gpu[k], ref[k], rtol=1e-5, atol=1e-6, err_msg=f"{label}: {k}")


def _run(mesh, cams, opt, tile=(16, 16)):
    r = Rasterizer(opt)
    r.set_tiling(*tile, bin_capacity=1024)
    r.set_mesh(mesh)
    assert r.PrepareMesh()
    g = r.render_gbuffer(cams)
    plain = r.render_batch(cams)
    for i, c in enumerate(cams):
        for k in plain[i]:
            assert np.array_equal(g[i][k], plain[i][k]), f"view {i} {k}: G-buffer call changed {k}"
        rc, ref, _ = O.render_gbuffer(mesh, O.camera_from_abi(c.to_abi()), opt, c2w_t=c.c2w[:3, 3])
        assert rc == 0
        yield i, g[i], ref


@pytest.mark.parametrize("normal", ["vertex", "face"])
def test_gbuffer_uv_sphere(gpu_available, normal):
    mesh, cam = scenes.scene_uv_sphere_128()
    opt = RendererCpuOption(diffuse_color=DiffuseColor.kVertex,
                            shading_normal=ShadingNormal.kVertex if normal == "vertex"
                            else ShadingNormal.kFace)
    for i, g, ref in _run(mesh, [cam], opt):
        _check(g, ref, f"uv sphere {normal}")


def test_gbuffer_icosphere_textured_multiview(gpu_available):
    mesh, _ = scenes.scene_icosphere8()
    cams = [c for i, c in enumerate(scenes.multiview_cameras(64)) if i in (0, 5, 17)]
    opt = RendererCpuOption(diffuse_color=DiffuseColor.kTexture,
                            interp=ColorInterpolation.kBilinear)
    for i, g, ref in _run(mesh, cams, opt, tile=(32, 8)):
        _check(g, ref, f"icosphere view {i}")
        m = ref["mask"] == 255
        assert np.abs(g["bary"][m].sum(-1) - 1.0).max() < 1e-6


def test_gbuffer_layered_occlusion(gpu_available):
    tris = TS.layered_grid_scene(seed=5)
    mesh = TS.mesh_from_pixels(tris, seed=5)
    cam = TS.exact_camera(96, 64)
    opt = RendererCpuOption(diffuse_color=DiffuseColor.kVertex, backface_culling=True)
    for i, g, ref in _run(mesh, [cam], opt):
        _check(g, ref, "layered")
