"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the oracle on CPU and the HIP path on GPU must reproduce them."""
import os

import numpy as np
import pytest

import golden_scenes as GS
import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return dict(np.load(os.path.join(HERE, f"{name}.npz")))


def _check(out, g, name):
    for k in ("face_id", "mask", "color"):
        assert np.array_equal(out[k], g[k]), f"{name}: {k} differs from the golden file"
    for k in ("depth", "normal"):
        np.testing.assert_allclose(out[k], g[k], rtol=1e-5, atol=1e-7, err_msg=f"{name}: {k}")


@pytest.mark.parametrize("name", GS.SCENES)
def test_oracle_reproduces_golden(name):
    g = _load(name)
    mesh, cam, opt = GS.build(name)
    a = cam.to_abi()
    camf = np.array([a.width, a.height, a.fx, a.fy, a.cx, a.cy] + list(a.w2c_R) +
                    list(a.w2c_t) + list(a.c2w_R), np.float64)
    assert np.array_equal(camf, g["camera"]), "scene camera changed"
    rc, out, _ = O.render(mesh, O.camera_from_abi(a), opt, c2w_t=cam.c2w[:3, 3])
    assert rc == 0
    _check(out, g, name)


@pytest.mark.gpu
@pytest.mark.parametrize("name", GS.SCENES)
def test_gpu_reproduces_golden(gpu_available, name):
    from ugu_amd import Rasterizer
    g = _load(name)
    mesh, cam, opt = GS.build(name)
    r = Rasterizer(opt)
    r.set_tiling(8, 8, 16)
    r.set_mesh(mesh)
    assert r.PrepareMesh()
    out = r.render_batch([cam])[0]
    _check(out, g, name)
