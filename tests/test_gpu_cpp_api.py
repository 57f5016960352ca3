"""GPU: the C++ drop-in API (ugu::Rasterizer and ugu::Raytracer in
ugu_amd/libugu_rasterizer.so)
driven like UGU's own example by ugu_amd/cpp/test_rasterizer_api, compared
with the oracle. Also checks the reference's error conventions through the
C++ interface and that RenderBatch equals the per-view Render loop."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from ugu_amd import DiffuseColor, RendererCpuOption, ShadingNormal, scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "ugu_amd", "cpp", "test_rasterizer_api")


def _dump_scene(d, mesh, cams, opt):
    tex = mesh.materials[0]
    meta = [len(mesh.vertices), len(mesh.vertex_indices), len(mesh.normals), len(mesh.uv),
            tex.shape[1], tex.shape[0], len(cams), int(opt.diffuse_color), int(opt.interp),
            int(opt.shading_normal), int(opt.diffuse_shading), int(opt.backface_culling)]
    open(os.path.join(d, "meta.txt"), "w").write(" ".join(map(str, meta)))
    mesh.vertices.astype(np.float32).tofile(os.path.join(d, "verts.f32"))
    mesh.vertex_indices.astype(np.int32).tofile(os.path.join(d, "faces.i32"))
    mesh.face_normals.astype(np.float32).tofile(os.path.join(d, "fnormals.f32"))
    mesh.normals.astype(np.float32).tofile(os.path.join(d, "normals.f32"))
    mesh.vertex_colors.astype(np.float32).tofile(os.path.join(d, "colors.f32"))
    mesh.uv.astype(np.float32).tofile(os.path.join(d, "uv.f32"))
    tex.astype(np.uint8).tofile(os.path.join(d, "tex.u8"))
    rows = []
    for c in cams:
        rows.append([c.width, c.height, c.focal_length[0], c.focal_length[1],
                     c.principal_point[0], c.principal_point[1]] + list(c.c2w.reshape(-1)))
    np.array(rows, np.float64).tofile(os.path.join(d, "cams.f64"))


@pytest.mark.parametrize("variant", ["vertex_color", "texture_lambert"])
def test_cpp_rasterizer_matches_oracle(gpu_available, tmp_path, variant):
    assert os.path.exists(DRIVER), "C++ driver not built (make -C ugu_amd/cpp)"
    v, f = scenes.icosphere_unit(5)
    mesh = scenes.perturbed_sphere(v, f, texture=scenes.checker_noise_texture(256, 3))
    rng = np.random.default_rng(9)
    mesh.set_vertex_colors(rng.random((len(mesh.vertices), 3), dtype=np.float32) * 255)
    cams = scenes.ring_cameras(5, 3.0, 31, 32, res_choices=[(320, 240), (200, 150), (97, 131)])
    if variant == "vertex_color":
        opt = RendererCpuOption(diffuse_color=DiffuseColor.kVertex)
    else:
        opt = RendererCpuOption(diffuse_color=DiffuseColor.kTexture,
                                shading_normal=ShadingNormal.kFace, diffuse_shading=1)
    scene, out = tmp_path / "scene", tmp_path / "out"
    scene.mkdir()
    out.mkdir()
    _dump_scene(str(scene), mesh, cams, opt)
    env = dict(os.environ, UGU_HIP_TUNE_CACHE=str(tmp_path / "tiling.txt"))
    res = subprocess.run([DRIVER, str(scene), str(out)], capture_output=True, text=True,
                         timeout=120, env=env)
    print(res.stdout[-2000:], res.stderr[-2000:])
    assert res.returncode == 0, res.stdout + res.stderr
    for name in ("prepare_without_mesh_fails", "render_before_prepare_fails", "all_null_fails",
                 "renderw_null_depth_fails", "render_without_camera_fails",
                 "culling_without_face_normals_fails", "vertex_color_without_colors_fails"):
        assert f"CHECK {name} ok" in res.stdout
    for i, c in enumerate(cams):
        h, w = c.height, c.width
        pfx = os.path.join(str(out), f"view{i}_")
        cam_f = np.fromfile(pfx + "camera.f32", np.float32)
        # the C++ host math (Affine3d inverse + rotation()) agrees with the
        # Python host math to float rounding
        a = c.to_abi()
        assert np.allclose(cam_f[:9], np.array(a.w2c_R[:]), atol=2e-7)
        assert np.allclose(cam_f[9:], np.array(a.w2c_t[:]), atol=1e-6)
        oc = O.camera_from_abi(a)
        for k in range(9):
            oc.w2c_R[k] = float(cam_f[k])
        for k in range(3):
            oc.w2c_t[k] = float(cam_f[9 + k])
        rc, ref, _ = O.render(mesh, oc, opt, c2w_t=c.c2w[:3, 3])
        assert rc == 0
        got = {
            "face_id": np.fromfile(pfx + "face_id.i32", np.int32).reshape(h, w),
            "mask": np.fromfile(pfx + "mask.u8", np.uint8).reshape(h, w),
            "color": np.fromfile(pfx + "color.u8", np.uint8).reshape(h, w, 3),
            "depth": np.fromfile(pfx + "depth.f32", np.float32).reshape(h, w),
            "normal": np.fromfile(pfx + "normal.f32", np.float32).reshape(h, w, 3),
        }
        if not np.array_equal(got["face_id"], ref["face_id"]):
            # same float camera through the Python API, for the diagnosis
            from ugu_amd import Rasterizer
            r = Rasterizer(opt)
            r.set_mesh(mesh)
            assert r.PrepareMesh()
            a2 = c.to_abi()
            for k in range(9):
                a2.w2c_R[k] = float(cam_f[k])
            for k in range(3):
                a2.w2c_t[k] = float(cam_f[9 + k])
            py = r.render_batch([a2])[0]
            bad = np.argwhere(got["face_id"] != ref["face_id"])
            y, x = bad[0]
            raise AssertionError(
                f"view {i} ({w}x{h}) face_id differs at {len(bad)} px, first ({y},{x}): "
                f"cpp={got['face_id'][y, x]} oracle={ref['face_id'][y, x]} "
                f"python-gpu={py['face_id'][y, x]}; python-gpu==oracle: "
                f"{np.array_equal(py['face_id'], ref['face_id'])}, python-gpu==cpp: "
                f"{np.array_equal(py['face_id'], got['face_id'])}")
        assert np.array_equal(got["mask"], ref["mask"]), f"view {i} mask"
        assert np.array_equal(got["color"], ref["color"]), f"view {i} colour"
        np.testing.assert_allclose(got["depth"], ref["depth"], rtol=1e-5)
        np.testing.assert_allclose(got["normal"], ref["normal"], rtol=1e-5, atol=1e-7)
        dw = np.fromfile(pfx + "depthw.u16", np.uint16).reshape(h, w)
        assert np.array_equal(dw, O.depth_to_u16(ref["depth"]))
        if i == 0:
            # RenderGBuffer: the Render planes it shares, and the G-buffer
            # planes against the oracle's restatement
            gp = os.path.join(str(out), "gbuf0_")
            rd = lambda n, dt, ch: np.fromfile(gp + n, dt).reshape((h, w, ch) if ch else (h, w))
            assert np.array_equal(rd("color.u8", np.uint8, 3), got["color"])
            assert np.array_equal(rd("shaded.u8", np.uint8, 3), got["color"])
            assert np.array_equal(rd("stencil.u8", np.uint8, 0), got["mask"])
            assert np.array_equal(rd("face_id.i32", np.int32, 0), got["face_id"])
            assert np.array_equal(rd("normal_cam.f32", np.float32, 3), got["normal"])
            rcg, refg, _ = O.render_gbuffer(mesh, oc, opt, c2w_t=c.c2w[:3, 3])
            assert rcg == 0
            assert np.array_equal(rd("geo_id.i32", np.int32, 0), refg["geo_id"])
            for k in ("bary", "uv", "pos_wld", "pos_cam", "normal_wld"):
                np.testing.assert_allclose(rd(k + ".f32", np.float32, 3), refg[k], rtol=1e-5,
                                           atol=1e-6, err_msg=f"RenderGBuffer {k}")
        bpfx = os.path.join(str(out), f"batch{i}_")
        assert np.array_equal(np.fromfile(bpfx + "face_id.i32", np.int32).reshape(h, w),
                              got["face_id"])
        assert np.array_equal(np.fromfile(bpfx + "color.u8", np.uint8).reshape(h, w, 3),
                              got["color"])
    # ugu::Raytracer through the same driver
    for name in ("rt_prepare_without_mesh_fails", "rt_prepare_empty_mesh_fails",
                 "rt_render_before_prepare_fails", "rt_prepare_ok", "rt_batch_equals_single"):
        assert f"CHECK {name} ok" in res.stdout
    from ugu_amd import Raytracer
    rt = Raytracer(opt)
    rt.set_mesh(mesh)
    assert rt.PrepareMesh()
    smallest = min(range(len(cams)), key=lambda i: cams[i].width * cams[i].height)
    for i, c in enumerate(cams):
        h, w = c.height, c.width
        pfx = os.path.join(str(out), f"rt_view{i}_")
        got = {
            "face_id": np.fromfile(pfx + "face_id.i32", np.int32).reshape(h, w),
            "mask": np.fromfile(pfx + "mask.u8", np.uint8).reshape(h, w),
            "color": np.fromfile(pfx + "color.u8", np.uint8).reshape(h, w, 3),
            "depth": np.fromfile(pfx + "depth.f32", np.float32).reshape(h, w),
            "normal": np.fromfile(pfx + "normal.f32", np.float32).reshape(h, w, 3),
        }
        cam_f = np.fromfile(os.path.join(str(out), f"view{i}_camera.f32"), np.float32)
        a2 = c.to_abi()
        for k in range(9):
            a2.w2c_R[k] = float(cam_f[k])
        for k in range(3):
            a2.w2c_t[k] = float(cam_f[9 + k])
        py = rt.render_batch([a2])[0]
        for k in got:
            assert np.array_equal(got[k], py[k]), f"raytracer view {i} {k}: C++ != Python"
        dw = np.fromfile(pfx + "depthw.u16", np.uint16).reshape(h, w)
        assert np.array_equal(dw, O.depth_to_u16(got["depth"]))
        if i == smallest:
            oc = O.camera_from_abi(a2)
            rc, ref, _ = O.raytrace(mesh, oc, opt, c2w_t=c.c2w[:3, 3])
            assert rc == 0
            assert np.array_equal(got["face_id"], ref["face_id"]), f"raytracer view {i} face_id"
            assert np.array_equal(got["mask"], ref["mask"])
            assert np.array_equal(got["color"], ref["color"])
            np.testing.assert_allclose(got["depth"], ref["depth"], rtol=1e-5)
            np.testing.assert_allclose(got["normal"], ref["normal"], rtol=1e-5, atol=1e-7)
