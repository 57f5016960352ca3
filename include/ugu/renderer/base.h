// ugu/renderer/base.h (include/ugu/renderer/base.h:11-77 in UGU): renderer
// enums and the G-buffer struct.
#pragma once

#include "ugu/image.h"

namespace ugu {

// Diffuse color
enum class DiffuseColor {
  kNone = 0,     // Default white color
  kTexture = 1,  // From diffuse uv texture
  kVertex = 2    // From vertex color
};

// Normal used for shading, also returned as output normal
enum class ShadingNormal {
  kFace = 0,   // Face normal
  kVertex = 1  // Vertex normal
};

// Diffuse shading; the light ray is the viewing ray
enum class DiffuseShading {
  kNone = 0,        // No shading
  kLambertian = 1,  // Lambertian reflectance model
  kOrenNayar = 2    // Simplified Oren-Nayar reflectance model
};

// G-buffer (include/ugu/renderer/base.h:34-77 in UGU, same fields). UGU fills
// it from OpenGL; Rasterizer::RenderGBuffer (an extension) fills it from the
// rasterizer's resolve: see that method for what each plane holds.
struct GBuffer {
  Image3b color;
  Image3f normal_cam;
  Image3f normal_wld;
  Image3f pos_cam;
  Image3f pos_wld;
  Image1f depth_01;
  Image1b stencil;
  Image1i face_id;
  Image3f uv;
  Image3f bary;
  Image1i geo_id;
  Image3b shaded;

  void Init(int w, int h) {
    color = Image3b::zeros(h, w);
    normal_cam = Image3f::zeros(h, w);
    normal_wld = Image3f::zeros(h, w);
    pos_cam = Image3f::zeros(h, w);
    pos_wld = Image3f::zeros(h, w);
    depth_01 = Image1f::zeros(h, w);
    stencil = Image1b::zeros(h, w);
    face_id = Image1i::zeros(h, w);
    uv = Image3f::zeros(h, w);
    bary = Image3f::zeros(h, w);
    geo_id = Image1i::zeros(h, w);
    shaded = Image3b::zeros(h, w);
  }

  void Reset() {
    color.setTo(0);
    normal_cam.setTo(0.0f);
    normal_wld.setTo(0.0f);
    pos_cam.setTo(0.0f);
    pos_wld.setTo(0.0f);
    depth_01.setTo(0.0f);
    stencil.setTo(0);
    face_id.setTo(0);
    uv.setTo(0.0f);
    bary.setTo(0.0f);
    geo_id.setTo(0);
    shaded.setTo(0);
  }
};

}  // namespace ugu
