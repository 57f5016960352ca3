"""Writes tests/golden/*.npz: small seeded scenes rendered by the CPU oracle.

Each file stores the scene name, the render options, the float camera and
every output image, so tests can check the oracle (regression) and the GPU
path against fixed data without re-deriving the scene. Scenes are rebuilt
from their seeds by tests/golden_scenes.py; the meshes themselves are not
stored. Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.dirname(HERE)]

import golden_scenes as GS  # noqa: E402
import oracle as O  # noqa: E402


def main():
    for name in GS.SCENES:
        mesh, cam, opt = GS.build(name)
        a = cam.to_abi()
        rc, out, _ = O.render(mesh, O.camera_from_abi(a), opt, c2w_t=cam.c2w[:3, 3])
        assert rc == 0, (name, rc)
        camf = np.array([a.width, a.height, a.fx, a.fy, a.cx, a.cy] + list(a.w2c_R) +
                        list(a.w2c_t) + list(a.c2w_R), np.float64)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), camera=camf,
                            options=np.array([int(opt.diffuse_color), int(opt.interp),
                                              int(opt.shading_normal), int(opt.diffuse_shading),
                                              int(opt.backface_culling)], np.int32),
                            **out)
        print(name, {k: v.shape for k, v in out.items()}, "covered", int((out["face_id"] >= 0).sum()))


if __name__ == "__main__":
    main()
