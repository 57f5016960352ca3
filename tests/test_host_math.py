"""CPU: host-side camera and mesh math of the drop-in API."""
import math

import numpy as np

from ugu_amd import Mesh, PinholeCamera, look_at_c2w, scenes
from ugu_amd.types import affine_inverse, inverse3x3, polar_rotation


def test_affine_inverse_matches_numpy():
    rng = np.random.default_rng(0)
    for _ in range(20):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        T = np.eye(4)
        T[:3, :3] = q * rng.uniform(0.5, 2.0)
        T[:3, 3] = rng.normal(size=3)
        np.testing.assert_allclose(affine_inverse(T), np.linalg.inv(T), rtol=1e-12, atol=1e-12)


def test_inverse3x3_exact_on_permutation():
    P = [[0, 1, 0], [0, 0, 1], [1, 0, 0]]
    assert inverse3x3(P) == [[0, 0, 1], [1, 0, 0], [0, 1, 0]]


def test_polar_rotation_of_scaled_rotation():
    rng = np.random.default_rng(1)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    R = polar_rotation(q @ np.diag([2.0, 3.0, 0.5]))
    np.testing.assert_allclose(R, q, atol=1e-12)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)


def test_look_at_is_orthonormal_and_points_at_target():
    T = look_at_c2w((1.0, 2.0, -3.0), (0, 0, 0), (0, 1, 0))
    R = T[:3, :3]
    np.testing.assert_allclose(R.T @ R, np.eye(3), atol=1e-12)
    z = R[:, 2]
    np.testing.assert_allclose(z, -np.array([1.0, 2.0, -3.0]) / math.sqrt(14), atol=1e-12)
    # world up projects upward in the image (OpenCV: image y down)
    assert R[:, 1] @ np.array([0, 1, 0]) < 0


def test_fov_constructor_semantics():
    cam = PinholeCamera(640, 480, fov_y_deg=45.0)
    assert cam.principal_point[0] == np.float32(319.5)
    assert cam.principal_point[1] == np.float32(239.5)
    f = 240.0 / math.tan(math.radians(45.0) / 2)
    assert abs(cam.focal_length[0] - f) < 1e-3 and cam.focal_length[0] == cam.focal_length[1]


def test_face_normals_match_float64():
    m = scenes.make_uv_sphere(16, 16)
    m.CalcFaceNormal()
    v = m.vertices.astype(np.float64)
    f = m.vertex_indices
    n = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    assert np.abs(m.face_normals - n).max() < 1e-5
    # outward on a sphere (the reference's winding)
    c = v[f].mean(1)
    assert ((n * c).sum(1) > 0).mean() > 0.99 or ((n * c).sum(1) < 0).mean() > 0.99


def test_calc_normal_averages_face_normals():
    m = scenes.make_uv_sphere(24, 24)
    m.CalcNormal()
    assert m.normals.shape == m.vertices.shape
    np.testing.assert_allclose(np.linalg.norm(m.normals, axis=1), 1.0, atol=1e-6)
    assert np.array_equal(m.normal_indices, m.vertex_indices)


def test_scene_sizes():
    v, f = scenes.icosphere_unit(3)
    assert len(f) == 20 * 4 ** 3 and len(v) == 10 * 4 ** 3 + 2
    v, f = scenes.geodesic_unit(71)
    assert len(f) == 100820 and len(v) == 50412
    m = scenes.make_uv_sphere(128, 128)
    assert len(m.vertex_indices) == 32512 and len(m.vertices) == 16258
    cams = scenes.multiview_cameras(64)
    assert len(cams) == 64
    assert {(c.width, c.height) for c in cams} <= set(scenes.RES_CHOICES)


def test_geodesic_is_watertight():
    v, f = scenes.geodesic_unit(9)
    e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
    _, counts = np.unique(e, axis=0, return_counts=True)
    assert (counts == 2).all()
    # Euler characteristic of a sphere
    assert len(v) - len(counts) + len(f) == 2


def test_device_binding_checks_buffer_shapes_and_camera_sizes():
    """DeviceBinding (Rasterizer.bind_device) marshals the arguments once; it
    rejects tensors whose shape is not the camera's image and a camera of
    another size than the bound tensors (the library would write past them)."""
    import pytest
    import torch
    from ugu_amd import RendererCpuOption
    from ugu_amd.renderer import DeviceBinding

    class _R:
        option_ = RendererCpuOption()

    c2w = look_at_c2w((0, 0, -3), (0, 0, 0), (0, 1, 0))
    small = PinholeCamera(64, 48, c2w, (31.5, 23.5), (60.0, 60.0))
    big = PinholeCamera(96, 64, c2w, (47.5, 31.5), (60.0, 60.0))
    bufs = [{"face_id": torch.empty((48, 64), dtype=torch.int32),
             "color": torch.empty((48, 64, 3), dtype=torch.uint8)},
            {"depth": torch.empty((64, 96), dtype=torch.float32)}]
    b = DeviceBinding(_R(), [small, big], bufs)
    assert b.n == 2 and b.sizes == [(64, 48), (96, 64)]
    assert b.outs[0].face_id == bufs[0]["face_id"].data_ptr()
    assert b.outs[0].depth is None and b.outs[1].depth == bufs[1]["depth"].data_ptr()
    moved = PinholeCamera(64, 48, look_at_c2w((1, 0, -3), (0, 0, 0), (0, 1, 0)), (31.5, 23.5),
                          (60.0, 60.0))
    b.set_camera(0, moved)
    assert b.cams[0].w2c_t[0] == moved.to_abi().w2c_t[0]
    with pytest.raises(ValueError):
        b.set_camera(0, big)
    with pytest.raises(ValueError):
        DeviceBinding(_R(), [big, small], bufs)
    with pytest.raises(ValueError):
        DeviceBinding(_R(), [small], bufs)
