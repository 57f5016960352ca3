"""GPU: the autotuner (ugu_hip_autotune_cached through the C ABI). Its choice
is one of the six candidate tilings; the bin capacity it settles on holds the
measured largest bin, so a render after tuning needs no retry; the outputs
after tuning equal the oracle's; a second context finds the choice in the
cache file without tuning, and other view sizes miss the file and are tuned."""
import numpy as np
import pytest

import oracle as O
from ugu_amd import DiffuseColor, Rasterizer, RendererCpuOption, scenes

pytestmark = pytest.mark.gpu

CANDIDATES = {(8, 8, False), (16, 16, False), (32, 8, False),
              (8, 8, True), (16, 16, True), (32, 8, True)}


def _scene():
    v, f = scenes.icosphere_unit(7)  # 327,680 faces: bins of thousands of entries
    mesh = scenes.perturbed_sphere(v, f, texture=scenes.checker_noise_texture(512, 3))
    cams = scenes.ring_cameras(4, 3.0, 71, 72, res_choices=[(320, 240), (256, 256), (160, 200)])
    return mesh, cams


def _renderer(mesh, opt):
    r = Rasterizer(opt)
    r.set_mesh(mesh)
    assert r.PrepareMesh()
    return r


def test_autotune_choice_capacity_and_cache(gpu_available, tmp_path):
    mesh, cams = _scene()
    opt = RendererCpuOption(diffuse_color=DiffuseColor.kTexture)
    cache = tmp_path / "tiling.txt"
    r = _renderer(mesh, opt)
    t = r.autotune(cams, repeats=2, cache_path=cache)
    print(f"[autotune] chose {t}", flush=True)
    assert not t["from_cache"]
    assert (t["tile_w"], t["tile_h"], t["persistent"]) in CANDIDATES
    assert t["bin_capacity"] > 0 and t["bin_capacity"] % 64 == 0
    assert r.get_tiling()["tile_w"] == t["tile_w"]
    assert len(cache.read_text().splitlines()) == 1

    # the tuned capacity needs no retry; at half of it a bin overflows, and
    # the occupancy the overflow reports is within the tuned capacity
    r.enable_profiling(True)
    outs = r.render_batch(cams)
    tm = r.timings()
    assert tm["retries"] == 0 and tm["max_bin_count"] == 0, tm
    half = t["bin_capacity"] // 2
    r.set_tiling(t["tile_w"], t["tile_h"], bin_capacity=half, persistent=t["persistent"])
    outs_half = r.render_batch(cams)
    tm = r.timings()
    r.enable_profiling(False)
    print(f"[autotune] largest bin {tm['max_bin_count']}, tuned capacity {t['bin_capacity']}",
          flush=True)
    assert tm["retries"] >= 1
    assert half < tm["max_bin_count"] <= t["bin_capacity"]
    for a, b in zip(outs, outs_half):
        for k in a:
            assert np.array_equal(a[k], b[k]), k
    for c, out in zip(cams, outs):
        rc, ref, _ = O.render(mesh, O.camera_from_abi(c.to_abi()), opt, c2w_t=c.c2w[:3, 3])
        assert rc == 0
        assert np.array_equal(out["face_id"], ref["face_id"])
        assert np.array_equal(out["mask"], ref["mask"])
        assert np.array_equal(out["color"], ref["color"])
        np.testing.assert_allclose(out["depth"], ref["depth"], rtol=1e-5, atol=0)
        np.testing.assert_allclose(out["normal"], ref["normal"], rtol=1e-5, atol=1e-7)

    # a new context finds the same choice in the file, without tuning
    r2 = _renderer(mesh, opt)
    t2 = r2.autotune(cams, repeats=2, cache_path=cache)
    assert t2["from_cache"]
    assert {k: t2[k] for k in ("tile_w", "tile_h", "persistent", "bin_capacity")} == \
        {k: t[k] for k in ("tile_w", "tile_h", "persistent", "bin_capacity")}
    assert r2.get_tiling()["bin_capacity"] == t["bin_capacity"]
    outs2 = r2.render_batch(cams)
    for a, b in zip(outs, outs2):
        for k in a:
            assert np.array_equal(a[k], b[k]), k

    # other view sizes are another key: tuned and appended
    cams[0].set_size(200, 100)
    cams[0].principal_point[:] = (99.5, 49.5)
    t3 = r2.autotune(cams, repeats=1, cache_path=cache)
    assert not t3["from_cache"]
    assert len(cache.read_text().splitlines()) == 2
