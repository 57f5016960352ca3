// ugu/renderer/base.h (include/ugu/renderer/base.h:11-32 in UGU): renderer enums.
#pragma once

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

}  // namespace ugu
