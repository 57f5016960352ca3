// ugu::Rasterizer on the gfx950 pipeline.
//
// Same interface and error behaviour as UGU's CPU rasterizer
// (src/renderer/cpu/rasterizer.cc:35-404): set_* never fail, PrepareMesh and
// Render* return false with the reference's LOGE text. PrepareMesh uploads
// the mesh once (allowed by the contract "don't modify mesh after
// PrepareMesh", renderer.h:52-53); Render reads the camera again on every call
// (callers move the camera between renders, examples/ex02_renderer.cc:142-143),
// reduces it to float on the host exactly as rasterizer.cc:118-119 and
// camera.h:322 do, and renders through ugu_hip_render.

#include "ugu/renderer/cpu/rasterizer.h"

#include <cstring>
#include <map>
#include <utility>

#include "gpu_common.h"

namespace ugu {

using namespace gpu_detail;

class Rasterizer::Impl {
  bool mesh_initialized_{false};
  std::shared_ptr<const Camera> camera_{nullptr};
  std::shared_ptr<const Mesh> mesh_{nullptr};
  RendererCpuOption option_;
  ugu_hip_ctx* ctx_{nullptr};
  mutable std::map<std::pair<int, int>, bool> tuned_;

 public:
  Impl() { ctx_ = ugu_hip_create(device_ordinal()); }
  explicit Impl(const RendererCpuOption& option) : Impl() { set_option(option); }
  ~Impl() { ugu_hip_destroy(ctx_); }

  void set_option(const RendererCpuOption& option) { option.CopyTo(&option_); }

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
    if (ctx_ == nullptr) {
      LOGE("no HIP device available\n");
      return false;
    }
    if (!upload_mesh(ctx_, *mesh_)) return false;
    tuned_.clear();
    mesh_initialized_ = true;
    return true;
  }

  void set_camera(std::shared_ptr<const Camera> camera) { camera_ = camera; }

  bool RenderViews(const std::vector<const Camera*>& cams, Image3b* const* color,
                   Image1f* const* depth, Image3f* const* normal, Image1b* const* mask,
                   Image1i* const* face_id) const {
    const size_t n = cams.size();
    std::vector<UguCamera> abi(n);
    std::vector<UguOutputs> outs(n);
    for (size_t i = 0; i < n; i++) {
      if (!to_abi(*cams[i], &abi[i])) {
        LOGE("only PinholeCamera is supported by the GPU rasterizer\n");
        return false;
      }
    }
    for (size_t i = 0; i < n; i++)
      outs[i] = init_outputs(cams[i]->width(), cams[i]->height(), color[i], depth[i], normal[i],
                             mask[i], face_id[i]);
    const UguOptions opt = to_abi(option_);
    // tile size and bin capacity are tuned once per view set (results do
    // not depend on them); the choice is kept in a file across processes, so
    // only the first process on a device tunes a given workload
    const auto key = std::make_pair(abi[0].width * 65536 + abi[0].height, static_cast<int>(n));
    if (!tuned_.count(key)) {
      UguTiling t;
      int cached = 0;
      const std::string path = tune_cache_path();
      if (ugu_hip_autotune_cached(ctx_, abi.data(), static_cast<int>(n), &opt, 2,
                                  path.c_str(), &t, &cached) == UGU_HIP_OK)
        tuned_[key] = true;
    }
    const int rc = n == 1 ? ugu_hip_render(ctx_, abi.data(), &opt, outs.data(), UGU_HIP_MEMORY_HOST)
                          : ugu_hip_render_batch(ctx_, abi.data(), static_cast<int>(n), &opt,
                                                 outs.data(), UGU_HIP_MEMORY_HOST);
    if (rc != UGU_HIP_OK) {
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
    return RenderViews(cams, &color, &depth, &normal, &mask, &face_id);
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
    return RenderViews(cams, pc.data(), pd.data(), pn.data(), pm.data(), pf.data());
  }

  bool RenderGBuffer(GBuffer* gbuf) const {
    if (gbuf == nullptr) {
      LOGE("gbuf is nullptr");
      return false;
    }
    if (!Validate(mesh_initialized_, camera_, mesh_, option_, true)) return false;
    UguCamera cam;
    if (!to_abi(*camera_, &cam)) {
      LOGE("only PinholeCamera is supported by the GPU rasterizer\n");
      return false;
    }
    gbuf->Init(cam.width, cam.height);
    UguOutputs out;
    out.color = gbuf->color.data;
    out.depth = nullptr;
    out.normal = reinterpret_cast<float*>(gbuf->normal_cam.data);
    out.mask = gbuf->stencil.data;
    out.face_id = reinterpret_cast<int32_t*>(gbuf->face_id.data);
    UguGBuffer g;
    g.bary = reinterpret_cast<float*>(gbuf->bary.data);
    g.uv = reinterpret_cast<float*>(gbuf->uv.data);
    g.pos_wld = reinterpret_cast<float*>(gbuf->pos_wld.data);
    g.pos_cam = reinterpret_cast<float*>(gbuf->pos_cam.data);
    g.normal_wld = reinterpret_cast<float*>(gbuf->normal_wld.data);
    g.geo_id = reinterpret_cast<int32_t*>(gbuf->geo_id.data);
    const UguOptions opt = to_abi(option_);
    if (ugu_hip_render_gbuffer(ctx_, &cam, 1, &opt, &out, &g, UGU_HIP_MEMORY_HOST) != UGU_HIP_OK) {
      LOGE("%s", ugu_hip_last_error(ctx_));
      return false;
    }
    std::memcpy(gbuf->shaded.data, gbuf->color.data, gbuf->color.total() * 3);
    return true;
  }

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

Rasterizer::Rasterizer() : pimpl_(std::unique_ptr<Impl>(new Impl)) {}
Rasterizer::~Rasterizer() {}
Rasterizer::Rasterizer(const RendererCpuOption& option)
    : pimpl_(std::unique_ptr<Impl>(new Impl(option))) {}
void Rasterizer::set_option(const RendererCpuOption& option) { pimpl_->set_option(option); }
void Rasterizer::set_mesh(std::shared_ptr<const Mesh> mesh) { pimpl_->set_mesh(mesh); }
bool Rasterizer::PrepareMesh() { return pimpl_->PrepareMesh(); }
void Rasterizer::set_camera(std::shared_ptr<const Camera> camera) { pimpl_->set_camera(camera); }
bool Rasterizer::Render(Image3b* color, Image1f* depth, Image3f* normal, Image1b* mask,
                        Image1i* face_id) const {
  return pimpl_->Render(color, depth, normal, mask, face_id);
}
bool Rasterizer::RenderColor(Image3b* color) const {
  return pimpl_->Render(color, nullptr, nullptr, nullptr, nullptr);
}
bool Rasterizer::RenderDepth(Image1f* depth) const {
  return pimpl_->Render(nullptr, depth, nullptr, nullptr, nullptr);
}
bool Rasterizer::RenderNormal(Image3f* normal) const {
  return pimpl_->Render(nullptr, nullptr, normal, nullptr, nullptr);
}
bool Rasterizer::RenderMask(Image1b* mask) const {
  return pimpl_->Render(nullptr, nullptr, nullptr, mask, nullptr);
}
bool Rasterizer::RenderFaceId(Image1i* face_id) const {
  return pimpl_->Render(nullptr, nullptr, nullptr, nullptr, face_id);
}
bool Rasterizer::RenderW(Image3b* color, Image1w* depth, Image3f* normal, Image1b* mask,
                         Image1i* face_id) const {
  return pimpl_->RenderW(color, depth, normal, mask, face_id);
}
bool Rasterizer::RenderDepthW(Image1w* depth) const {
  return pimpl_->RenderW(nullptr, depth, nullptr, nullptr, nullptr);
}
#ifdef UGU_RASTERIZER_HAS_RENDER_BATCH
bool Rasterizer::RenderBatch(const std::vector<std::shared_ptr<const Camera>>& cameras,
                             std::vector<Image3b>* color, std::vector<Image1f>* depth,
                             std::vector<Image3f>* normal, std::vector<Image1b>* mask,
                             std::vector<Image1i>* face_id) const {
  return pimpl_->RenderBatch(cameras, color, depth, normal, mask, face_id);
}
#endif

#ifdef UGU_RASTERIZER_HAS_RENDER_GBUFFER
bool Rasterizer::RenderGBuffer(GBuffer* gbuf) const { return pimpl_->RenderGBuffer(gbuf); }
#endif

}  // namespace ugu
