"""GPU: the visibility test kernel (ugu_hip_visibility_test) against the
oracle's TestVertex on the same items -- every record field equal (the
viewing angle within 2 ulp: acosf on the device and in glibc may differ in
the last bit) -- and whole VisibilityInfo structures from the GPU tester
equal to those from the oracle-backed tester."""
import ctypes

import numpy as np
import pytest

import oracle_pins as P
import visibility_util as U
from ugu_amd import _abi, visibility as V

pytestmark = pytest.mark.gpu


def _compare(gpu, ref, label):
    print(f"[visibility] {label}: items={len(ref)} visible={int(ref['visible'].sum())}",
          flush=True)
    assert np.array_equal(gpu["visible"], ref["visible"]), label
    vis = ref["visible"] == 1
    for k in ("px", "py", "dist", "c0", "c1", "c2", "intensity"):
        assert np.array_equal(gpu[k][vis].view(np.uint32), ref[k][vis].view(np.uint32)), \
            f"{label} {k}"
    ulp = np.abs(gpu["angle"][vis].view(np.int32) - ref["angle"][vis].view(np.int32))
    assert ulp.max(initial=0) <= 2, f"{label} angle differs by {ulp.max()} ulp"


def _gpu_tester(scene, **opt):
    m, normals, *_ = scene
    t = V.VisibilityTester(V.VisibilityTesterOption(**opt))
    t.set_data(m.vertices, normals, m.vertex_indices)
    assert t.PrepareData()
    return t


@pytest.mark.parametrize("interp,use_depth", [(1, True), (0, True), (1, False)])
def test_items_match_oracle(gpu_available, interp, use_depth):
    scene = P.visibility_scene(seed=interp)
    kw = dict(interp=interp, use_depth=use_depth, max_depth_difference=0.05)
    t = _gpu_tester(scene, **kw)
    ref_t = U.OracleTester(V.VisibilityTesterOption(**kw))
    ref_t.set_data(t.vertices_, t.normals_, t.indices_)
    ref_t.PrepareData()
    for kf in U.keyframes(scene, n=4, seed=interp):
        t.set_keyframe(kf)
        ref_t.set_keyframe(kf)
        for mode, items in (("vertex", (t.vertices_, t.vertices_, t.normals_)),
                            ("facewise", t.facewise_items())):
            gpu, _ = t.test_items(*items)
            ref, _ = ref_t.test_items(*items)
            _compare(gpu, ref, f"kf {kf.id} {mode} interp={interp} depth={use_depth}")


@pytest.mark.parametrize("facewise", [True, False])
def test_visibility_info_matches_oracle(gpu_available, facewise):
    scene = P.visibility_scene(seed=5)
    kfs = U.keyframes(scene, n=5, seed=5)
    t = _gpu_tester(scene)
    ref_t = U.OracleTester()
    ref_t.set_data(t.vertices_, t.normals_, t.indices_)
    ref_t.PrepareData()
    a, b = V.VisibilityInfo(scene[0]), V.VisibilityInfo(scene[0])
    assert t.Test(a, facewise=facewise, keyframes=kfs)
    assert ref_t.Test(b, facewise=facewise, keyframes=kfs)
    ra, rb = U.info_records(a), U.info_records(b)
    n_vis = sum(len(r[0]) for r in ra[:len(a.vertex_info_list)])
    print(f"[visibility] info facewise={facewise}: vertex records={n_vis}", flush=True)
    assert n_vis > 0
    # angles may differ by an ulp; everything else must be identical
    for x, y in zip(ra, rb):
        assert len(x[0]) == len(y[0])
        for p, q in zip(x[0], y[0]):
            for i, (u, w) in enumerate(zip(p, q)):
                if isinstance(u, float) and u != w:
                    assert abs(u - w) <= 4e-7 * max(1.0, abs(w)), (x, y)
                else:
                    assert u == w


def test_visibility_errors(gpu_available):
    scene = P.visibility_scene()
    m, normals, cam, color, depth, mask = scene
    t = V.VisibilityTester()
    t.set_data(m.vertices, normals, m.vertex_indices)
    info = V.VisibilityInfo(m)
    t.set_keyframe(V.Keyframe(0, color, depth, mask, cam))
    assert not t.Test(info)                              # PrepareData not called
    assert t.PrepareData()
    t.set_keyframe(V.Keyframe(0, color, depth, None, cam))
    assert not t.Test(info)                              # use_mask without a mask
    t.set_keyframe(V.Keyframe(0, color, None, mask, cam))
    assert not t.Test(info)                              # use_depth without a depth
    assert not t.Test(V.VisibilityInfo(num_vertices=3, num_faces=1))
    e = V.VisibilityTester()
    e.set_data(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((0, 3)))
    assert not e.PrepareData()                           # mesh is empty
    # the C ABI refuses a keyframe without colour
    k = _abi.UguKeyframe()
    k.camera, k.width, k.height = cam.to_abi(), 96, 80
    o = V.VisibilityTesterOption().to_abi()
    lib = _abi.load_library()
    assert lib.ugu_hip_visibility_test(t._ctx, ctypes.byref(k), ctypes.byref(o), 0, None, None,
                                       None, None) == _abi.ERR_INVALID_ARGUMENT
