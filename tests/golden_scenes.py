"""Seeded scenes behind the golden fixtures in tests/golden/."""
import numpy as np

import tiny_scenes as TS
from ugu_amd import (ColorInterpolation, DiffuseColor, DiffuseShading, RendererCpuOption,
                     ShadingNormal, scenes)

SCENES = ["uvsphere_80x60_vertex", "ico5_96x72_texture_bilinear", "ico5_96x72_texture_nn_lambert",
          "lattice_24x20_ties", "subnormal_16x16"]


def build(name):
    if name == "uvsphere_80x60_vertex":
        mesh, _ = scenes.scene_uv_sphere_128()
        cam = scenes.front_camera(80, 60, 62.5, 3.0)
        return mesh, cam, RendererCpuOption(diffuse_color=DiffuseColor.kVertex)
    if name.startswith("ico5_96x72"):
        v, f = scenes.icosphere_unit(5)
        mesh = scenes.perturbed_sphere(v, f, texture=scenes.checker_noise_texture(64, 11, block=8))
        cam = scenes.ring_cameras(1, 3.0, 51, 52, res_choices=[(96, 72)])[0]
        if name.endswith("bilinear"):
            opt = RendererCpuOption(diffuse_color=DiffuseColor.kTexture,
                                    interp=ColorInterpolation.kBilinear)
        else:
            opt = RendererCpuOption(diffuse_color=DiffuseColor.kTexture,
                                    interp=ColorInterpolation.kNn,
                                    shading_normal=ShadingNormal.kFace,
                                    diffuse_shading=DiffuseShading.kLambertian)
        return mesh, cam, opt
    if name == "lattice_24x20_ties":
        tris = TS.random_triangles(30, 24, 20, 77)
        tris = np.concatenate([tris, tris[:6]], 0)
        mesh = TS.mesh_from_pixels(tris, seed=77)
        return mesh, TS.exact_camera(24, 20), RendererCpuOption(
            diffuse_color=DiffuseColor.kVertex, backface_culling=False)
    if name == "subnormal_16x16":
        mesh, cam = TS.subnormal_depth_scene(flip=True)
        return mesh, cam, RendererCpuOption(diffuse_color=DiffuseColor.kVertex)
    raise KeyError(name)
