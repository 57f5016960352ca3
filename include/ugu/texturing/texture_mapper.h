// ugu/texturing/texture_mapper.h (UGU): TextureMapping with kSimpleProjection
// for every uv type: kGenerateSimpleTile (keyframes tiled into one texture, uv
// from the projected triangles), kUseOriginalMeshUv (faces rasterized into the
// mesh's own uv layout on the GPU, ugu_hip_texture_map),
// kGenerateSimpleTriangles (each face rasterized into its half of a fixed
// rectangle grid on the GPU) and kGenerateSimpleCharts (charts of adjacent
// faces that chose one keyframe, their keyframe patches packed into one
// texture on the host).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "ugu/texturing/visibility_tester.h"

namespace ugu {

// Plain (unscoped) enums as in texture_mapper.h:12-19, so ugu::kUseOriginalMeshUv
// names the value directly.
enum TextureMappingType { kSimpleProjection = 0 };
enum TexturingOutputUvType {
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
