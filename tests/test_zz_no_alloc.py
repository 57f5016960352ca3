"""GPU: ugu_hip_render_batch_noalloc (the allocation-free batched render,
through the C ABI) equals ugu_hip_render_batch on the same context, bit for
bit in every output plane, on two calls in a row (the second with fewer and
smaller views), including a batch that does not fit the scratch array at once
and is rendered in groups; sizes beyond the array's maximum are rejected
before anything runs."""
import ctypes

import numpy as np
import pytest

from ugu_amd import (ColorInterpolation, DiffuseColor, DiffuseShading, Rasterizer,
                     RendererCpuOption, ShadingNormal, _abi, scenes)
from ugu_amd.renderer import options_to_abi

pytestmark = pytest.mark.gpu

PLANES = ("color", "depth", "normal", "mask", "face_id")


def _buffers(torch, cams):
    out = []
    for c in cams:
        h, w = c.height, c.width
        out.append({
            "color": torch.full((h, w, 3), 7, dtype=torch.uint8, device="cuda"),
            "depth": torch.full((h, w), 7.0, dtype=torch.float32, device="cuda"),
            "normal": torch.full((h, w, 3), 7.0, dtype=torch.float32, device="cuda"),
            "mask": torch.full((h, w), 7, dtype=torch.uint8, device="cuda"),
            "face_id": torch.full((h, w), 7, dtype=torch.int32, device="cuda"),
        })
    return out


def _noalloc(r, cams, bufs):
    arr = (_abi.UguCamera * len(cams))(*[c.to_abi() for c in cams])
    outs = (_abi.UguOutputs * len(cams))(*[
        _abi.UguOutputs(*[b[k].data_ptr() for k in PLANES]) for b in bufs])
    opt = options_to_abi(r.option_)
    rc = r._lib.ugu_hip_render_batch_noalloc(r._ctx, arr, len(cams), ctypes.byref(opt), outs)
    return rc, r._lib.ugu_hip_last_error(r._ctx).decode()


def _same(torch, a, b, label):
    """Every plane of every view equal bit for bit."""
    torch.cuda.synchronize()
    assert len(a) == len(b)
    for i, (da, db) in enumerate(zip(a, b)):
        for k in PLANES:
            x, y = da[k].cpu().numpy(), db[k].cpu().numpy()
            assert x.dtype == y.dtype and x.shape == y.shape
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), \
                f"{label}: view {i} {k} differs"


@pytest.fixture(scope="module")
def ico():
    return scenes.scene_icosphere8()


def test_noalloc_equals_batched_render(gpu_available, ico):
    import torch
    mesh, _ = ico
    cams = scenes.multiview_cameras(64)
    opt = RendererCpuOption(diffuse_color=DiffuseColor.kTexture,
                            interp=ColorInterpolation.kBilinear,
                            shading_normal=ShadingNormal.kVertex,
                            diffuse_shading=DiffuseShading.kLambertian)
    r = Rasterizer(opt)
    # a small starting capacity: the 64-view batch overflows the bins the
    # scratch array can give 220,800 tiles at once and is split into groups
    r.set_tiling(32, 8, bin_capacity=64, persistent=True)
    r.set_mesh(mesh)
    assert r.PrepareMesh()
    small = [cams[5], cams[17], cams[40]]
    small[1].set_size(200, 120)  # a size the first call did not have
    small[1].principal_point[:] = (99.5, 59.5)
    print(f"[noalloc] F={len(mesh.vertex_indices)} views {len(cams)} then {len(small)}",
          flush=True)
    got1 = _buffers(torch, cams)
    rc, msg = _noalloc(r, cams, got1)
    assert rc == 0, msg
    got2 = _buffers(torch, small)
    rc, msg = _noalloc(r, small, got2)
    assert rc == 0, msg
    # the existing entry point on the same context and inputs
    ref1 = _buffers(torch, cams)
    r.render_device(cams, ref1)
    ref2 = _buffers(torch, small)
    r.render_device(small, ref2)
    _same(torch, got1, ref1, "64 views")
    _same(torch, got2, ref2, "3 views")
    cov = int((ref1[0]["face_id"] >= 0).sum())
    assert cov > 1000
    print(f"[noalloc] equal to ugu_hip_render_batch; view 0 covers {cov} px", flush=True)


def test_noalloc_rejects_sizes_beyond_the_scratch(gpu_available):
    import torch
    mesh, cam = scenes.scene_uv_sphere_128()
    r = Rasterizer(RendererCpuOption(diffuse_color=DiffuseColor.kVertex))
    r.set_tiling(8, 8, bin_capacity=64, persistent=True)
    r.set_mesh(mesh)
    assert r.PrepareMesh()
    # 1025 views: rejected before any work
    bufs = _buffers(torch, [cam])
    many = [cam] * 1025
    rc, msg = _noalloc(r, many, bufs * 1025)
    assert rc == _abi.UGU_HIP_ERR_INVALID_ARGUMENT and "1024 views" in msg, (rc, msg)
    # a 32768 x 32768 view has 16.7 M 8x8 tiles: its counters alone exceed the
    # array. Only the face id plane is requested, and it is allocated at full
    # size, so the buffer is valid whatever happens.
    big = cam.__class__(32768, 32768, cam.c2w, (16383.5, 16383.5), (20000.0, 20000.0))
    fid = torch.empty((32768, 32768), dtype=torch.int32, device="cuda")
    arr = (_abi.UguCamera * 1)(big.to_abi())
    outs = (_abi.UguOutputs * 1)(_abi.UguOutputs(None, None, None, None, fid.data_ptr()))
    opt = options_to_abi(r.option_)
    rc = r._lib.ugu_hip_render_batch_noalloc(r._ctx, arr, 1, ctypes.byref(opt), outs)
    msg = r._lib.ugu_hip_last_error(r._ctx).decode()
    assert rc == _abi.UGU_HIP_ERR_OUT_OF_MEMORY and "too large" in msg, (rc, msg)
    del fid
    # the context still renders afterwards
    got = _buffers(torch, [cam])
    rc, msg = _noalloc(r, [cam], got)
    assert rc == 0, msg
    ref = _buffers(torch, [cam])
    r.render_device([cam], ref)
    _same(torch, got, ref, "after rejections")
