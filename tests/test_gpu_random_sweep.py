"""GPU: seeded random sweep against the oracle -- pixel-lattice scenes with
exact edge ties and perspective icosphere scenes, random options (colour
source, interpolation, shading normal, Lambert, culling), random tile shapes,
grid or persistent launch and small bin capacities (so bins overflow and are
regrown), multi-view batches of mixed sizes. face_id, mask and colour
bit-exact; depth and normal within 1e-5 relative."""
import numpy as np
import pytest

import oracle as O
import tiny_scenes as TS
from ugu_amd import (ColorInterpolation, DiffuseColor, DiffuseShading, PinholeCamera, Rasterizer,
                     RendererCpuOption, ShadingNormal, look_at_c2w, scenes)

pytestmark = pytest.mark.gpu

RTOL = 1e-5
TILES = [(8, 8, False), (8, 8, True), (16, 16, False), (16, 16, True), (32, 8, False),
         (32, 8, True)]


def _scene(seed, rng):
    if seed % 3 == 0:
        w, h = int(rng.integers(8, 160)), int(rng.integers(8, 120))
        tris = TS.random_triangles(int(rng.integers(5, 1500)), w, h, seed)
        if seed % 2 == 0:
            tris = np.concatenate([tris, tris[: max(1, len(tris) // 4)]], 0)  # exact ties
        mesh = TS.mesh_from_pixels(tris, seed=seed)
        TS.add_texture(mesh, int(rng.integers(3, 48)), seed)
        return mesh, [TS.exact_camera(w, h)]
    mesh, _ = scenes.scene_icosphere8(level=int(rng.integers(2, 7)), tex_size=96, seed=seed)
    mesh.set_vertex_colors((rng.random((len(mesh.vertices), 3)) * 255).astype(np.float32))
    cams = []
    for _ in range(int(rng.integers(1, 4))):
        w, h = int(rng.integers(16, 400)), int(rng.integers(16, 300))
        pos = rng.normal(size=3)
        pos = pos / np.linalg.norm(pos) * rng.uniform(1.2, 5.0)
        cams.append(PinholeCamera(w, h, look_at_c2w(pos, (0, 0, 0), (0, 1, 0)),
                                  fov_y_deg=float(rng.uniform(15, 100))))
    return mesh, cams


@pytest.mark.parametrize("block", range(4))
def test_random_scenes_match_oracle(gpu_available, block):
    views = 0
    for seed in range(block * 15, block * 15 + 15):
        rng = np.random.default_rng(7000 + seed)
        mesh, cams = _scene(seed, rng)
        opt = RendererCpuOption(diffuse_color=DiffuseColor(int(rng.integers(0, 3))),
                                interp=ColorInterpolation(int(rng.integers(0, 2))),
                                shading_normal=ShadingNormal(int(rng.integers(0, 2))),
                                diffuse_shading=DiffuseShading(int(rng.integers(0, 2))),
                                backface_culling=bool(rng.integers(0, 2)))
        tw, th, persistent = TILES[int(rng.integers(0, len(TILES)))]
        r = Rasterizer(opt)
        r.set_tiling(tw, th, int(rng.integers(1, 128)), persistent)
        r.set_mesh(mesh)
        assert r.PrepareMesh()
        outs = r.render_batch(cams)
        for i, cam in enumerate(cams):
            rc, ref, _ = O.render(mesh, O.camera_from_abi(cam.to_abi()), opt,
                                  c2w_t=cam.c2w[:3, 3])
            assert rc == 0
            label = (f"seed {seed} view {i} {cam.width}x{cam.height} "
                     f"F={len(mesh.vertex_indices)} tile {tw}x{th}{'p' if persistent else 'g'}")
            out = outs[i]
            assert np.array_equal(out["face_id"], ref["face_id"]), label + ": face_id"
            assert np.array_equal(out["mask"], ref["mask"]), label + ": mask"
            assert np.array_equal(out["color"], ref["color"]), label + ": colour"
            np.testing.assert_allclose(out["depth"], ref["depth"], rtol=RTOL, atol=0,
                                       err_msg=label)
            np.testing.assert_allclose(out["normal"], ref["normal"], rtol=RTOL, atol=1e-6,
                                       err_msg=label)
            views += 1
    print(f"[sweep] block {block}: {views} views match the oracle", flush=True)
