"""CPU: the host side of the visibility tester (ugu_amd/visibility.py) --
VisibilityInfo assembly (vertex-wise with TestFaces, face-wise), the
multi-keyframe driver and the statistics -- with the oracle's TestVertex as
the per-item test."""
import numpy as np

import oracle_pins as P
import visibility_util as U
from ugu_amd import visibility as V


def _tester(scene, **opt):
    m, normals, *_ = scene
    t = U.OracleTester(V.VisibilityTesterOption(**opt))
    t.set_data(m.vertices, normals, m.vertex_indices)
    assert t.PrepareData()
    return t


def test_vertexwise_faces_follow_vertices():
    scene = P.visibility_scene()
    t = _tester(scene)
    kfs = U.keyframes(scene, n=3)
    info = V.VisibilityInfo(scene[0])
    assert t.Test(info, facewise=False, keyframes=kfs)
    faces = t.indices_
    for kf in kfs:
        for f, face in enumerate(faces):
            vis = all(info.vertex_info_list[v].VisibleFrom(kf.id) >= 0 for v in face)
            has = any(r.kf_id == kf.id for r in info.face_info_list[f].visible_keyframes)
            assert vis == has, f"face {f} keyframe {kf.id}"
    # the statistics point at the largest projected area (max_area itself
    # keeps its initial value, as in the reference's FaceInfo::CalcStat)
    fi = next(f for f in info.face_info_list if f.visible_keyframes)
    areas = [r.area for r in fi.visible_keyframes]
    assert areas[fi.max_area_index] == max(areas)
    assert fi.max_area == -1.0


def test_facewise_first_face_and_short_circuit():
    scene = P.visibility_scene()
    t = _tester(scene)
    kf = U.keyframes(scene, n=1)[0]
    t.set_keyframe(kf)
    info = V.VisibilityInfo(scene[0])
    assert t.Test(info, facewise=True)
    pos, vert, nrm = t.facewise_items()
    rec = U.oracle_items(t, pos, vert, nrm).reshape(-1, 3)
    first = {}
    for f, face in enumerate(t.indices_):
        for j in range(3):
            if not rec[f, j]["visible"]:
                break
            first.setdefault(int(face[j]), (f, j))
    for v, vi in enumerate(info.vertex_info_list):
        if v in first:
            f, j = first[v]
            assert len(vi.visible_keyframes) == 1
            assert vi.visible_keyframes[0].projected_pos.tobytes() == \
                np.array([rec[f, j]["px"], rec[f, j]["py"]], np.float32).tobytes()
        else:
            assert not vi.visible_keyframes
    # a face is recorded iff its three corners pass, in order
    for f in range(len(t.indices_)):
        assert bool(info.face_info_list[f].visible_keyframes) == bool(rec[f]["visible"].all())


def test_statistics_helpers():
    rng = np.random.default_rng(1)
    for n in (1, 2, 4, 6, 8):
        x = rng.random(n).astype(np.float32)
        s = np.sort(x)
        exp = x[0] if n == 1 else (np.float32((x[0] + x[1]) * np.float32(0.5)) if n == 2
                                   else s[n // 2])
        assert V.median(list(x)) == exp
    x = np.float32([3.0, 1.0, 2.0])
    assert V.median(list(x)) == np.float32(2.5)        # (sorted[1] + sorted[2]) / 2
    assert V.median(list(x), True) == np.float32(2.0)
    cols = [np.float32([10, 20, 30]), np.float32([40, 50, 60]), np.float32([70, 80, 90])]
    w = [1.0, 1.0, 2.0]
    exp = np.sum(np.array(cols, np.float64) * (np.float32(w) / np.float32(4))[:, None], 0)
    assert np.allclose(V.weighted_average(cols, w), exp)
    # weights (1, 1, 2) + 1e-6 each over their float sum 4.0000029: 0.25000006
    # twice, so the running sum passes 0.5 at the second entry (without the
    # 1e-6 it would reach exactly 0.5 there and pick the third)
    assert np.array_equal(V.weighted_median(cols, w), cols[1])
    assert np.allclose(V.normalize_weights([0.0, 0.0]), [0.5, 0.5])
