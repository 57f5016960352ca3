// ugu::Raytracer on the gfx950 raytracer kernel.
//
// Same interface and error behaviour as UGU's CPU raytracer
// (src/renderer/cpu/raytracer.cc:57-305): set_* never fail, PrepareMesh
// checks the mesh (raytracer.cc:90-99), uploads it and builds the bounding
// volume hierarchy once (ugu_hip_prepare_raytracer, for raytracer.cc:101-104);
// Render* validate like ValidateAndInitBeforeRender and trace through
// ugu_hip_raytrace. Each pixel gets the closest hit over all faces (smallest
// float t, equal t -> lowest face id), which the reference's sort by t leaves
// open for equal t.

#include "ugu/renderer/cpu/raytracer.h"

#include "gpu_common.h"

namespace ugu {

using namespace gpu_detail;

class Raytracer::Impl {
  bool mesh_initialized_{false};
  std::shared_ptr<const Camera> camera_{nullptr};
  std::shared_ptr<const Mesh> mesh_{nullptr};
  RendererCpuOption option_;
  ugu_hip_ctx* ctx_{nullptr};

 public:
  Impl() { ctx_ = ugu_hip_create(device_ordinal()); }
  explicit Impl(const RendererCpuOption& option) : Impl() { set_option(option); }
  ~Impl() { ugu_hip_destroy(ctx_); }

  void set_option(const RendererCpuOption& option) { option.CopyTo(&option_); }

  // raytracer.cc:78-88
  void set_mesh(std::shared_ptr<const Mesh> mesh) {
    mesh_initialized_ = false;
    mesh_ = mesh;
    if (mesh_->face_normals().empty())
      LOGW("face normal is empty. culling and shading may not work\n");
    if (mesh_->normals().empty()) LOGW("vertex normal is empty. shading may not work\n");
  }

  bool PrepareMesh() {
    if (mesh_ == nullptr) {
      LOGE("mesh has not been set\n");
      return false;
    }
    if (mesh_->vertices().empty() || mesh_->vertex_indices().empty()) {
      LOGE("mesh is empty\n");
      return false;
    }
    if (ctx_ == nullptr) {
      LOGE("no HIP device available\n");
      return false;
    }
    if (!upload_mesh(ctx_, *mesh_)) return false;
    if (ugu_hip_prepare_raytracer(ctx_) != UGU_HIP_OK) {
      LOGE("%s", ugu_hip_last_error(ctx_));
      return false;
    }
    mesh_initialized_ = true;
    return true;
  }

  void set_camera(std::shared_ptr<const Camera> camera) { camera_ = camera; }

  bool TraceViews(const std::vector<const Camera*>& cams, Image3b* const* color,
                  Image1f* const* depth, Image3f* const* normal, Image1b* const* mask,
                  Image1i* const* face_id) const {
    const size_t n = cams.size();
    std::vector<UguCamera> abi(n);
    std::vector<UguOutputs> outs(n);
    for (size_t i = 0; i < n; i++) {
      if (!to_abi(*cams[i], &abi[i])) {
        LOGE("only PinholeCamera is supported by the GPU raytracer\n");
        return false;
      }
      outs[i] = init_outputs(cams[i]->width(), cams[i]->height(), color[i], depth[i], normal[i],
                             mask[i], face_id[i]);
    }
    const UguOptions opt = to_abi(option_);
    if (ugu_hip_raytrace(ctx_, abi.data(), static_cast<int>(n), &opt, outs.data(),
                         UGU_HIP_MEMORY_HOST) != UGU_HIP_OK) {
      LOGE("%s", ugu_hip_last_error(ctx_));
      return false;
    }
    return true;
  }

  bool Render(Image3b* color, Image1f* depth, Image3f* normal, Image1b* mask,
              Image1i* face_id) const {
    if (!Validate(mesh_initialized_, camera_, mesh_, option_,
                  color || depth || normal || mask || face_id))
      return false;
    std::vector<const Camera*> cams{camera_.get()};
    return TraceViews(cams, &color, &depth, &normal, &mask, &face_id);
  }

  bool RenderBatch(const std::vector<std::shared_ptr<const Camera>>& cameras,
                   std::vector<Image3b>* color, std::vector<Image1f>* depth,
                   std::vector<Image3f>* normal, std::vector<Image1b>* mask,
                   std::vector<Image1i>* face_id) const {
    if (cameras.empty()) {
      LOGE("camera has not been set\n");
      return false;
    }
    for (const auto& c : cameras)
      if (!Validate(mesh_initialized_, c, mesh_, option_,
                    color || depth || normal || mask || face_id))
        return false;
    const size_t n = cameras.size();
    if (color) color->resize(n);
    if (depth) depth->resize(n);
    if (normal) normal->resize(n);
    if (mask) mask->resize(n);
    if (face_id) face_id->resize(n);
    std::vector<const Camera*> cams(n);
    std::vector<Image3b*> pc(n, nullptr);
    std::vector<Image1f*> pd(n, nullptr);
    std::vector<Image3f*> pn(n, nullptr);
    std::vector<Image1b*> pm(n, nullptr);
    std::vector<Image1i*> pf(n, nullptr);
    for (size_t i = 0; i < n; i++) {
      cams[i] = cameras[i].get();
      if (color) pc[i] = &(*color)[i];
      if (depth) pd[i] = &(*depth)[i];
      if (normal) pn[i] = &(*normal)[i];
      if (mask) pm[i] = &(*mask)[i];
      if (face_id) pf[i] = &(*face_id)[i];
    }
    return TraceViews(cams, pc.data(), pd.data(), pn.data(), pm.data(), pf.data());
  }

  // raytracer.cc:246-262
  bool RenderW(Image3b* color, Image1w* depth, Image3f* normal, Image1b* mask,
               Image1i* face_id) const {
    if (depth == nullptr) {
      LOGE("depth is nullptr");
      return false;
    }
    Image1f f_depth;
    bool org_ret = Render(color, &f_depth, normal, mask, face_id);
    if (org_ret) ConvertTo(f_depth, depth);
    return org_ret;
  }
};

Raytracer::Raytracer() : pimpl_(std::unique_ptr<Impl>(new Impl)) {}
Raytracer::~Raytracer() {}
Raytracer::Raytracer(const RendererCpuOption& option)
    : pimpl_(std::unique_ptr<Impl>(new Impl(option))) {}
void Raytracer::set_option(const RendererCpuOption& option) { pimpl_->set_option(option); }
void Raytracer::set_mesh(std::shared_ptr<const Mesh> mesh) { pimpl_->set_mesh(mesh); }
bool Raytracer::PrepareMesh() { return pimpl_->PrepareMesh(); }
void Raytracer::set_camera(std::shared_ptr<const Camera> camera) { pimpl_->set_camera(camera); }
bool Raytracer::Render(Image3b* color, Image1f* depth, Image3f* normal, Image1b* mask,
                       Image1i* face_id) const {
  return pimpl_->Render(color, depth, normal, mask, face_id);
}
bool Raytracer::RenderColor(Image3b* color) const {
  return pimpl_->Render(color, nullptr, nullptr, nullptr, nullptr);
}
bool Raytracer::RenderDepth(Image1f* depth) const {
  return pimpl_->Render(nullptr, depth, nullptr, nullptr, nullptr);
}
bool Raytracer::RenderNormal(Image3f* normal) const {
  return pimpl_->Render(nullptr, nullptr, normal, nullptr, nullptr);
}
bool Raytracer::RenderMask(Image1b* mask) const {
  return pimpl_->Render(nullptr, nullptr, nullptr, mask, nullptr);
}
bool Raytracer::RenderFaceId(Image1i* face_id) const {
  return pimpl_->Render(nullptr, nullptr, nullptr, nullptr, face_id);
}
bool Raytracer::RenderW(Image3b* color, Image1w* depth, Image3f* normal, Image1b* mask,
                        Image1i* face_id) const {
  return pimpl_->RenderW(color, depth, normal, mask, face_id);
}
bool Raytracer::RenderDepthW(Image1w* depth) const {
  return pimpl_->RenderW(nullptr, depth, nullptr, nullptr, nullptr);
}
#ifdef UGU_RAYTRACER_HAS_RENDER_BATCH
bool Raytracer::RenderBatch(const std::vector<std::shared_ptr<const Camera>>& cameras,
                            std::vector<Image3b>* color, std::vector<Image1f>* depth,
                            std::vector<Image3f>* normal, std::vector<Image1b>* mask,
                            std::vector<Image1i>* face_id) const {
  return pimpl_->RenderBatch(cameras, color, depth, normal, mask, face_id);
}
#endif

}  // namespace ugu
