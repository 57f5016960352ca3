// ugu::Rasterizer (include/ugu/renderer/cpu/rasterizer.h:14-53 in UGU), same
// declaration; the implementation (ugu_amd/cpp/rasterizer.cc) renders on the
// GPU through the C ABI of include/ugu_hip.h.
#pragma once

#include <memory>
#include <vector>

#include "ugu/renderer/cpu/renderer.h"

// Feature tests for the extensions below (absent from UGU's header).
#define UGU_RASTERIZER_HAS_RENDER_BATCH 1
#define UGU_RASTERIZER_HAS_RENDER_GBUFFER 1

namespace ugu {

class Rasterizer : public RendererCpu {
  class Impl;
  std::unique_ptr<Impl> pimpl_;

 public:
  Rasterizer();
  ~Rasterizer() override;

  explicit Rasterizer(const RendererCpuOption& option);
  void set_option(const RendererCpuOption& option) override;
  void set_mesh(std::shared_ptr<const Mesh> mesh) override;
  bool PrepareMesh() override;
  void set_camera(std::shared_ptr<const Camera> camera) override;

  bool Render(Image3b* color, Image1f* depth, Image3f* normal, Image1b* mask,
              Image1i* face_id) const override;
  bool RenderColor(Image3b* color) const override;
  bool RenderDepth(Image1f* depth) const override;
  bool RenderNormal(Image3f* normal) const override;
  bool RenderMask(Image1b* mask) const override;
  bool RenderFaceId(Image1i* face_id) const override;
  bool RenderW(Image3b* color, Image1w* depth, Image3f* normal, Image1b* mask,
               Image1i* face_id) const override;
  bool RenderDepthW(Image1w* depth) const override;

  // Extension (no counterpart in UGU): renders several cameras in one batched
  // GPU pipeline; outputs[i] follow cameras[i]. Same results as a loop of
  // set_camera + Render.
  bool RenderBatch(const std::vector<std::shared_ptr<const Camera>>& cameras,
                   std::vector<Image3b>* color, std::vector<Image1f>* depth,
                   std::vector<Image3f>* normal, std::vector<Image1b>* mask,
                   std::vector<Image1i>* face_id) const;

  // Extension (no counterpart in ugu::Rasterizer; UGU fills GBuffer from
  // OpenGL): the current camera's G-buffer from the rasterizer's resolve.
  // color = shaded = the Render colour; normal_cam = the Render normal;
  // stencil = the Render mask; face_id = the Render face id (-1 where no
  // face); normal_wld = the shading normal in world coordinates; pos_wld,
  // pos_cam = the surface point (perspective-correct); bary = (u, v,
  // (1 - u) - v) and uv = (s, t, (1 - s) - t) with u, v the weights the
  // colourizers use; geo_id = 0 on covered pixels, -1 elsewhere; depth_01
  // (an OpenGL depth-buffer value) stays 0. Uncovered pixels are 0.
  bool RenderGBuffer(GBuffer* gbuf) const;
};

}  // namespace ugu
