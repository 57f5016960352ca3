// Helpers shared by the GPU implementations of ugu::Rasterizer and
// ugu::Raytracer: float reductions of the camera and options for the C ABI,
// the ValidateAndInitBeforeRender checks and the mesh upload.
#pragma once

#include <cstdlib>
#include <filesystem>
#include <memory>
#include <string>
#include <system_error>
#include <vector>

#include "ugu/log.h"
#include "ugu/renderer/cpu/renderer.h"
#include "../../include/ugu_hip.h"

namespace ugu {
namespace gpu_detail {

inline int device_ordinal() {
  const char* s = std::getenv("UGU_HIP_DEVICE");
  return s ? std::atoi(s) : 0;
}

// File that keeps the autotuner's choices across processes:
// $UGU_HIP_TUNE_CACHE if set (empty: no file), else
// $XDG_CACHE_HOME/ugu_amd/tiling.txt or ~/.cache/ugu_amd/tiling.txt (the
// directory is created when missing). "" when there is nowhere to keep it.
inline std::string tune_cache_path() {
  if (const char* e = std::getenv("UGU_HIP_TUNE_CACHE")) return e;
  std::string dir;
  if (const char* x = std::getenv("XDG_CACHE_HOME"); x && *x) {
    dir = x;
  } else if (const char* h = std::getenv("HOME"); h && *h) {
    dir = std::string(h) + "/.cache";
  } else {
    return "";
  }
  dir += "/ugu_amd";
  std::error_code ec;
  std::filesystem::create_directories(dir, ec);
  if (ec) return "";
  return dir + "/tiling.txt";
}

inline UguOptions to_abi(const RendererCpuOption& o) {
  UguOptions r;
  r.diffuse_color = static_cast<int32_t>(o.diffuse_color);
  r.interp = static_cast<int32_t>(o.interp);
  r.shading_normal = static_cast<int32_t>(o.shading_normal);
  r.diffuse_shading = static_cast<int32_t>(o.diffuse_shading);
  r.depth_scale = o.depth_scale;
  r.backface_culling = o.backface_culling ? 1 : 0;
  r.oren_nayar_sigma = o.oren_nayar_sigma;
  return r;
}

// rasterizer.cc:118-119 and camera.h:322: float reductions of the pose
inline bool to_abi(const Camera& camera, UguCamera* c) {
  const PinholeCamera* p = dynamic_cast<const PinholeCamera*>(&camera);
  if (p == nullptr) return false;
  c->width = p->width();
  c->height = p->height();
  c->fx = p->focal_length()[0];
  c->fy = p->focal_length()[1];
  c->cx = p->principal_point()[0];
  c->cy = p->principal_point()[1];
  const Eigen::Matrix3f w2c_R = p->w2c().rotation().cast<float>();
  const Eigen::Vector3f w2c_t = p->w2c().translation().cast<float>();
  const Eigen::Matrix3f c2w_R = p->c2w().matrix().block<3, 3>(0, 0).cast<float>();
  for (int r = 0; r < 3; r++) {
    for (int k = 0; k < 3; k++) {
      c->w2c_R[r * 3 + k] = w2c_R(r, k);
      c->c2w_R[r * 3 + k] = c2w_R(r, k);
    }
    c->w2c_t[r] = w2c_t[r];
  }
  const Eigen::Vector3f c2w_t = p->c2w().translation().cast<float>();
  for (int r = 0; r < 3; r++) c->c2w_t[r] = c2w_t[r];
  return true;
}

// ValidateAndInitBeforeRender checks (src/renderer/cpu/util_private.cc:19-66)
inline bool Validate(bool mesh_initialized, const std::shared_ptr<const Camera>& camera,
              const std::shared_ptr<const Mesh>& mesh, const RendererCpuOption& option,
              bool any_output) {
  if (camera == nullptr) {
    LOGE("camera has not been set\n");
    return false;
  }
  if (!mesh_initialized) {
    LOGE("mesh has not been initialized\n");
    return false;
  }
  if (option.backface_culling && mesh->face_normals().empty()) {
    LOGE("specified back-face culling but face normal is empty.\n");
    return false;
  }
  if (option.diffuse_color == DiffuseColor::kTexture && mesh->materials().empty()) {
    LOGE("specified texture as diffuse color but texture is empty.\n");
    return false;
  }
  if (option.diffuse_color == DiffuseColor::kTexture) {
    for (int i = 0; i < static_cast<int>(mesh->materials().size()); i++) {
      if (mesh->materials()[i].diffuse_tex.empty()) {
        LOGE("specified texture as diffuse color but %d th texture is empty.\n", i);
        return false;
      }
    }
  }
  if (option.diffuse_color == DiffuseColor::kVertex && mesh->vertex_colors().empty()) {
    LOGE("specified vertex color as diffuse color but vertex color is empty.\n");
    return false;
  }
  if (option.shading_normal == ShadingNormal::kFace && mesh->face_normals().empty()) {
    LOGE("specified face normal as shading normal but face normal is empty.\n");
    return false;
  }
  if (option.shading_normal == ShadingNormal::kVertex && mesh->normals().empty()) {
    LOGE("specified vertex normal as shading normal but vertex normal is empty.\n");
    return false;
  }
  if (!any_output) {
    LOGE("all arguments are nullptr. nothing to do\n");
    return false;
  }
  return true;
}

template <typename V>
inline const void* data_or_null(const std::vector<V>& v) {
  return v.empty() ? nullptr : static_cast<const void*>(v.data());
}

// Uploads the mesh with ugu_hip_prepare_mesh (the device copy the renderers
// draw from); logs the library's message and returns false on failure.
inline bool upload_mesh(ugu_hip_ctx* ctx, const Mesh& m) {
  std::vector<UguTexture> tex(m.materials().size());
  for (size_t i = 0; i < tex.size(); i++) {
    const Image3b& t = m.materials()[i].diffuse_tex;
    tex[i].width = t.empty() ? 0 : t.cols;
    tex[i].height = t.empty() ? 0 : t.rows;
    tex[i].data = t.empty() ? nullptr : t.data;
  }
  UguMeshDesc d;
  d.num_vertices = static_cast<int32_t>(m.vertices().size());
  d.vertices = static_cast<const float*>(data_or_null(m.vertices()));
  d.num_faces = static_cast<int32_t>(m.vertex_indices().size());
  d.faces = static_cast<const int32_t*>(data_or_null(m.vertex_indices()));
  d.face_normals = m.face_normals().size() >= m.vertex_indices().size()
                       ? static_cast<const float*>(data_or_null(m.face_normals()))
                       : nullptr;
  d.num_normals = static_cast<int32_t>(m.normals().size());
  d.normals = static_cast<const float*>(data_or_null(m.normals()));
  d.normal_indices = m.normal_indices().size() >= m.vertex_indices().size()
                         ? static_cast<const int32_t*>(data_or_null(m.normal_indices()))
                         : nullptr;
  d.num_vertex_colors = static_cast<int32_t>(m.vertex_colors().size());
  d.vertex_colors = static_cast<const float*>(data_or_null(m.vertex_colors()));
  d.num_uv = static_cast<int32_t>(m.uv().size());
  d.uv = static_cast<const float*>(data_or_null(m.uv()));
  d.uv_indices = m.uv_indices().size() >= m.vertex_indices().size()
                     ? static_cast<const int32_t*>(data_or_null(m.uv_indices()))
                     : nullptr;
  d.material_ids = m.material_ids().size() >= m.vertex_indices().size()
                       ? static_cast<const int32_t*>(data_or_null(m.material_ids()))
                       : nullptr;
  d.num_materials = static_cast<int32_t>(tex.size());
  d.textures = tex.empty() ? nullptr : tex.data();
  if (ugu_hip_prepare_mesh(ctx, &d) != UGU_HIP_OK) {
    LOGE("%s", ugu_hip_last_error(ctx));
    return false;
  }
  return true;
}

// ValidateAndInitBeforeRender's output initialisation (util_private.cc:68-87)
// and the matching C-ABI output pointers.
inline UguOutputs init_outputs(int w, int h, Image3b* color, Image1f* depth, Image3f* normal,
                               Image1b* mask, Image1i* face_id) {
  if (color) Init(color, w, h, static_cast<unsigned char>(0));
  if (depth) Init(depth, w, h, 0.0f);
  if (normal) Init(normal, w, h, 0.0f);
  if (mask) Init(mask, w, h, static_cast<unsigned char>(0));
  if (face_id) Init(face_id, w, h, -1);
  UguOutputs o;
  o.color = color ? color->data : nullptr;
  o.depth = depth ? reinterpret_cast<float*>(depth->data) : nullptr;
  o.normal = normal ? reinterpret_cast<float*>(normal->data) : nullptr;
  o.mask = mask ? mask->data : nullptr;
  o.face_id = face_id ? reinterpret_cast<int32_t*>(face_id->data) : nullptr;
  return o;
}

}  // namespace gpu_detail
}  // namespace ugu
