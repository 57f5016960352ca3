// ugu::Raytracer (include/ugu/renderer/cpu/raytracer.h:14-53 in UGU), same
// declaration; the implementation (ugu_amd/cpp/raytracer.cc) traces rays on
// the GPU through the C ABI of include/ugu_hip.h.
#pragma once

#include <memory>
#include <vector>

#include "ugu/renderer/cpu/renderer.h"

// Feature test for the extension below (absent from UGU's header).
#define UGU_RAYTRACER_HAS_RENDER_BATCH 1

namespace ugu {

class Raytracer : public RendererCpu {
  class Impl;
  std::unique_ptr<Impl> pimpl_;

 public:
  Raytracer();
  ~Raytracer() override;

  explicit Raytracer(const RendererCpuOption& option);
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

  // Extension (no counterpart in UGU): traces several cameras in one launch;
  // outputs[i] follow cameras[i]. Same results as a loop of set_camera +
  // Render.
  bool RenderBatch(const std::vector<std::shared_ptr<const Camera>>& cameras,
                   std::vector<Image3b>* color, std::vector<Image1f>* depth,
                   std::vector<Image3f>* normal, std::vector<Image1b>* mask,
                   std::vector<Image1i>* face_id) const;
};

}  // namespace ugu
