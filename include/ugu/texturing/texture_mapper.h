// Subset of ugu/texturing/texture_mapper.h (UGU): TextureMapping with
// kSimpleProjection and kUseOriginalMeshUv, the faces rasterized into the
// texture on the GPU (ugu_hip_texture_map). The atlas generators
// (kGenerateSimpleTile / Triangles / Charts) are not provided.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "ugu/texturing/visibility_tester.h"

namespace ugu {

enum class TextureMappingType { kSimpleProjection = 0 };
enum class TexturingOutputUvType {
  kGenerateSimpleTile = 0,
  kUseOriginalMeshUv = 1,
  kGenerateSimpleTriangles = 2,
  kGenerateSimpleCharts = 3
};

// TextureMappingOption (texture_mapper.h:21-29)
struct TextureMappingOption {
  ViewSelectionCriteria criteria = ViewSelectionCriteria::kMaxArea;
  TextureMappingType type = TextureMappingType::kSimpleProjection;
  TexturingOutputUvType uv_type = TexturingOutputUvType::kGenerateSimpleTile;
  std::string texture_base_name = "ugutex";
  int tex_w = 1024;
  int tex_h = 1024;
  int padding_kernel = 3;
};

// TextureMapping (texture_mapper.cc:1009-1017). Keyframe ids must be their
// positions in `keyframes` (the reference indexes the list with them).
bool TextureMapping(const std::vector<std::shared_ptr<Keyframe>>& keyframes,
                    const VisibilityInfo& info, Mesh* mesh, const TextureMappingOption& option);

}  // namespace ugu
