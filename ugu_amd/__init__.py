"""ugu_amd: UGU's software rasterizer (ugu::Rasterizer behind ugu::RendererCpu)
rebuilt for MI355X (gfx950) as hand-written HIP kernels behind a C ABI
(include/ugu_hip.h, ugu_amd/libugu_hip.so).

Python entry points:
  Rasterizer          mirror of ugu::Rasterizer (set_option/set_mesh/PrepareMesh/
                      set_camera/Render...), plus batched and device-memory renders
  Raytracer           mirror of ugu::Raytracer (same interface, BVH ray casting)
  Mesh, PinholeCamera, RendererCpuOption, Image*   reference types
  diff                differentiable rendering (DiffRenderer, RasterizeFunction:
                      a torch.autograd.Function)
  scenes              seeded benchmark scenes
"""
from .types import (ColorInterpolation, DiffuseColor, DiffuseShading, Image, Image1b,
                    Image1f, Image1i, Image1w, Image3b, Image3f, Mesh, PinholeCamera,
                    RendererCpuOption, ShadingNormal, look_at_c2w)
from .renderer import Rasterizer, Raytracer
from . import scenes

Renderer = Rasterizer

__all__ = ["Rasterizer", "Raytracer", "Renderer", "Mesh", "PinholeCamera", "RendererCpuOption",
           "DiffuseColor", "DiffuseShading", "ShadingNormal", "ColorInterpolation",
           "Image", "Image1b", "Image1f", "Image1i", "Image1w", "Image3b", "Image3f",
           "look_at_c2w", "scenes"]
