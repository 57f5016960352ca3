"""Helpers shared by the visibility tests: the oracle-backed item tester
(the oracle's TestVertex in place of the GPU kernel) and info comparison."""
import numpy as np

import oracle as O
import oracle_pins as P
from ugu_amd import visibility as V


def oracle_items(tester, pos, vert, nrm):
    kf = tester.keyframe_
    o = tester.option_
    opt = P.vis_options(backface_culling_angle_rad_th=o.backface_culling_angle_rad_th,
                        use_mask=int(o.use_mask), use_depth=int(o.use_depth), interp=int(o.interp),
                        max_viewing_angle=o.max_viewing_angle,
                        max_distance_from_camera_to_vertex=o.max_distance_from_camera_to_vertex,
                        max_depth_difference=o.max_depth_difference, eps=o.eps)
    from ugu_amd import Mesh
    m = Mesh(tester.vertices_, tester.indices_)
    oc = O.camera_from_abi(kf.camera.to_abi())
    lin = kf.camera.w2c[:3, :3].astype(np.float32)
    for i in range(9):
        oc.w2c_R[i] = float(lin.reshape(-1)[i])
    return O.visibility(m, oc, kf.camera.c2w[:3, 3], kf.color, kf.depth, kf.mask, opt, pos, vert,
                        nrm)


class OracleTester(V.VisibilityTester):
    """VisibilityTester whose per-item test is the CPU oracle (no GPU)."""

    def __init__(self, option=None):  # no HIP context
        self.option_ = V.VisibilityTesterOption()
        if option is not None:
            self.set_option(option)
        self.mesh_initialized_ = False
        self.keyframe_ = None
        self._ctx = None

    def PrepareData(self):
        v = self.vertices_[self.indices_]
        c = np.zeros((len(self.indices_), 3), np.float32)
        for k in range(3):
            c = (c + v[:, k]).astype(np.float32)
        self.face_centers_ = (c / np.float32(3)).astype(np.float32)
        self.mesh_initialized_ = True
        return True

    def test_items(self, pos, vert, nrm):
        return oracle_items(self, pos, vert, nrm), None


def info_records(info):
    """Comparable lists of every per-keyframe record and statistic."""
    out = []
    for vi in info.vertex_info_list:
        recs = [(k.kf_id, k.color.tobytes(), float(k.intensity), k.projected_pos.tobytes(),
                 float(k.viewing_angle), float(k.distance)) for k in vi.visible_keyframes]
        stats = {a: (getattr(vi, a).tobytes() if isinstance(getattr(vi, a), np.ndarray)
                     else getattr(vi, a))
                 for a in ("mean_color", "mean_distance", "mean_viewing_angle",
                           "min_viewing_angle_id", "min_distance_id", "median_color",
                           "mean_viewing_angle_color", "median_distance_color")
                 if hasattr(vi, a)}
        out.append((recs, stats))
    for fi in info.face_info_list:
        recs = [(k.kf_id, float(k.area), float(k.viewing_angle), float(k.distance),
                 k.projected_tri.tobytes()) for k in fi.visible_keyframes]
        out.append((recs, getattr(fi, "max_area_id", None), getattr(fi, "min_distance_id", None)))
    return out


def keyframes(scene, n=4, seed=0):
    """The visibility scene seen from the pin's camera and n-1 more poses."""
    from ugu_amd import PinholeCamera, RendererCpuOption, DiffuseColor, look_at_c2w
    m, normals, cam, color, depth, mask = scene
    kfs = [V.Keyframe(0, color, depth, mask, cam)]
    rng = np.random.default_rng(seed)
    for i in range(1, n):
        pos = np.array([np.sin(0.5 * i), 0.3 * np.cos(i), -np.cos(0.5 * i)]) * 3.0
        c = PinholeCamera(96, 80, look_at_c2w(pos, (0, 0, 0), (0, 1, 0)), (47.5, 39.5),
                          (70.0, 70.0))
        out = P.render(m, c, RendererCpuOption(diffuse_color=DiffuseColor.kVertex,
                                               backface_culling=False))
        col = rng.integers(0, 256, (80, 96, 3), dtype=np.uint8)
        kfs.append(V.Keyframe(i, col, out["depth"], (out["mask"] > 0).astype(np.uint8) * 255, c))
    return kfs
