"""GPU: a mesh of more than 2^24 faces. Above 2^24 the rasterizer's depth keys
carry the face id alone (no LDS slot of the staged face), so the resolve
evaluates every winner from its face record instead of the registers filled
during the chunks. 2^24 zero-area faces (every corner vertex 0) are put in
front of a textured icosphere, whose faces then have ids above 2^24; the
views must equal the oracle's, which renders the same padded mesh."""
import numpy as np
import pytest

import oracle as O
from ugu_amd import (ColorInterpolation, DiffuseColor, DiffuseShading, Mesh, Rasterizer,
                     RendererCpuOption, ShadingNormal, scenes)

pytestmark = pytest.mark.gpu

PAD = 1 << 24


def _padded_scene():
    v, f = scenes.icosphere_unit(5)
    base = scenes.perturbed_sphere(v, f, texture=scenes.checker_noise_texture(256, 4))
    F = len(base.vertex_indices)
    zeros = np.zeros((PAD, 3), np.int32)
    m = Mesh(base.vertices, np.concatenate([zeros, base.vertex_indices]))
    m.set_normals(base.normals)
    m.set_normal_indices(np.concatenate([zeros, base.normal_indices]))
    m.set_uv(base.uv)
    m.set_uv_indices(np.concatenate([zeros, base.uv_indices]))
    # the zero-area faces get the normal of a zero cross product; they are
    # rejected before it is read (|area| < FLT_MIN)
    m.face_normals = np.concatenate([np.zeros((PAD, 3), np.float32), base.face_normals])
    m.set_single_material(base.materials[0])
    return m, F


def test_faces_above_2_pow_24(gpu_available):
    mesh, F = _padded_scene()
    cams = scenes.ring_cameras(3, 3.0, 81, 82, res_choices=[(96, 72), (128, 96)])
    print(f"[large] {len(mesh.vertex_indices)} faces ({F} real, ids from {PAD})", flush=True)
    for opt in (RendererCpuOption(diffuse_color=DiffuseColor.kTexture,
                                  interp=ColorInterpolation.kBilinear),
                RendererCpuOption(diffuse_color=DiffuseColor.kTexture,
                                  interp=ColorInterpolation.kNn,
                                  shading_normal=ShadingNormal.kFace,
                                  diffuse_shading=DiffuseShading.kLambertian)):
        r = Rasterizer(opt)
        r.set_tiling(16, 16, bin_capacity=256, persistent=True)
        r.set_mesh(mesh)
        assert r.PrepareMesh()
        outs = r.render_batch(cams)
        for c, out in zip(cams, outs):
            rc, ref, _ = O.render(mesh, O.camera_from_abi(c.to_abi()), opt, c2w_t=c.c2w[:3, 3])
            assert rc == 0
            fid = ref["face_id"]
            assert (fid >= PAD).sum() > 1000 and not ((fid >= 0) & (fid < PAD)).any()
            assert np.array_equal(out["face_id"], fid)
            assert np.array_equal(out["mask"], ref["mask"])
            assert np.array_equal(out["color"], ref["color"])
            np.testing.assert_allclose(out["depth"], ref["depth"], rtol=1e-5, atol=0)
            np.testing.assert_allclose(out["normal"], ref["normal"], rtol=1e-5, atol=1e-7)
        del r
