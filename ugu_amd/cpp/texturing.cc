// ugu::VisibilityTester and ugu::TextureMapping (kUseOriginalMeshUv) for the
// C++ API: the per-vertex test and the face rasterization into the texture
// run on the GPU (ugu_hip_visibility_test, ugu_hip_texture_map); the loops
// that assemble VisibilityInfo and the statistics follow
// src/texturing/visibility_tester.cc and include/ugu/util/math_util.h, with
// the same standard-library algorithms (std::nth_element, std::sort,
// std::min_element), so a libstdc++ build reproduces the reference's choices.
// The face-wise test records a vertex from the lowest face id that passes it
// (the reference's OpenMP loop leaves that order unspecified).

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "gpu_common.h"
#include "ugu/texturing/texture_mapper.h"
#include "ugu/texturing/visibility_tester.h"

namespace ugu {

namespace {

using gpu_detail::device_ordinal;
using gpu_detail::to_abi;

// NormalizeWeights / WeightedAverage / WeightedMedian / Median / MedianColor
// (include/ugu/util/math_util.h:29-126, src/util/math_util.cc:44-56)
void NormalizeWeights(const std::vector<float>& weights, std::vector<float>* out) {
  *out = weights;
  const float sum = std::accumulate(out->begin(), out->end(), 0.0f);
  if (sum > 0.000001) {
    for (float& n : *out) n /= sum;
  } else {
    const float val = 1.0f / static_cast<float>(out->size());
    std::fill(out->begin(), out->end(), val);
  }
}

Eigen::Vector3f WeightedAverage(const std::vector<Eigen::Vector3f>& data,
                                const std::vector<float>& weights) {
  std::vector<float> nw;
  NormalizeWeights(weights, &nw);
  double acc[3] = {0.0, 0.0, 0.0};
  for (size_t i = 0; i < data.size(); i++)
    for (int k = 0; k < 3; k++) acc[k] += (data[i][k] * nw[i]);
  return Eigen::Vector3f(static_cast<float>(acc[0]), static_cast<float>(acc[1]),
                         static_cast<float>(acc[2]));
}

Eigen::Vector3f WeightedMedian(const std::vector<Eigen::Vector3f>& data,
                               const std::vector<float>& weights) {
  std::vector<float> nw;
  NormalizeWeights(weights, &nw);
  std::vector<std::pair<float, Eigen::Vector3f>> dw;
  for (size_t i = 0; i < data.size(); i++) dw.push_back(std::make_pair(nw[i], data[i]));
  std::sort(dw.begin(), dw.end(),
            [](const std::pair<float, Eigen::Vector3f>& a,
               const std::pair<float, Eigen::Vector3f>& b) { return a.first < b.first; });
  float sum = 0.0f;
  size_t index = 0;
  for (size_t i = 0; i < dw.size(); i++) {
    sum += dw[i].first;
    if (sum > 0.5f) {
      index = i;
      break;
    }
  }
  return dw[index].second;
}

float Median(const std::vector<float>& data, bool force_org_val = false) {
  if (data.size() == 1) return data[0];
  if (data.size() == 2) return force_org_val ? data[0] : (data[0] + data[1]) * 0.5f;
  std::vector<float> d(data);
  const size_t n = d.size() / 2;
  if (force_org_val || d.size() % 2 == 0) {
    std::nth_element(d.begin(), d.begin() + long(n), d.end());
    return d[n];
  }
  std::nth_element(d.begin(), d.begin() + long(n) + 1, d.end());
  return (d[n] + d[n + 1]) * 0.5f;
}


// EdgeFunction of two-vectors (texture_mapper.cc:143-146)
float Edge(const Eigen::Vector2f& a, const Eigen::Vector2f& b, const Eigen::Vector2f& c) {
  return (c[0] - a[0]) * (b[1] - a[1]) - (c[1] - a[1]) * (b[0] - a[0]);
}

// The keyframe's camera for the C ABI; false for a camera that is not a
// PinholeCamera (the only model the GPU path implements).
bool keyframe_camera(const Camera& cam, UguCamera* c) {
  if (!to_abi(cam, c)) return false;
  // TestVertex multiplies by w2c().cast<float>(): its linear part, not rotation()
  for (int r = 0; r < 3; r++)
    for (int k = 0; k < 3; k++) c->w2c_R[r * 3 + k] = static_cast<float>(cam.w2c().matrix()(r, k));
  return true;
}

}  // namespace

void VertexInfo::CalcStat() {
  if (visible_keyframes.empty()) return;
  mean_color = Eigen::Vector3f(0, 0, 0);
  mean_distance = 0;
  mean_viewing_angle = 0;
  for (auto& kf : visible_keyframes) {
    mean_color += kf.color;
    mean_distance += kf.distance;
    mean_viewing_angle += kf.viewing_angle;
  }
  mean_color /= static_cast<float>(visible_keyframes.size());
  mean_distance /= visible_keyframes.size();
  mean_viewing_angle /= visible_keyframes.size();
  std::vector<Eigen::Vector3f> colors;
  std::vector<float> distances, viewing_angles, intensities, inv_distances, inv_viewing_angles;
  for (auto& kf : visible_keyframes) {
    colors.push_back(kf.color);
    distances.push_back(kf.distance);
    viewing_angles.push_back(kf.viewing_angle);
    intensities.push_back(kf.intensity);
    inv_distances.push_back(1.0f / kf.distance);
    inv_viewing_angles.push_back(1.0f / kf.viewing_angle);
  }
  min_viewing_angle_index = static_cast<int>(
      std::min_element(viewing_angles.begin(), viewing_angles.end()) - viewing_angles.begin());
  min_viewing_angle_id = visible_keyframes[size_t(min_viewing_angle_index)].kf_id;
  min_viewing_angle_color = visible_keyframes[size_t(min_viewing_angle_index)].color;
  min_distance_index =
      static_cast<int>(std::min_element(distances.begin(), distances.end()) - distances.begin());
  min_distance_id = visible_keyframes[size_t(min_distance_index)].kf_id;
  min_distance_color = visible_keyframes[size_t(min_distance_index)].color;
  median_viewing_angle = Median(viewing_angles);
  median_distance = Median(distances);
  for (int k = 0; k < 3; k++) {
    std::vector<float> ch;
    for (const auto& c : colors) ch.push_back(c[k]);
    median_color[k] = Median(ch);
  }
  mean_viewing_angle_color = WeightedAverage(colors, inv_viewing_angles);
  mean_distance_color = WeightedAverage(colors, inv_distances);
  median_viewing_angle_color = WeightedMedian(colors, inv_viewing_angles);
  median_distance_color = WeightedMedian(colors, inv_distances);
  const auto min_it = std::min_element(intensities.begin(), intensities.end());
  min_intensity = *min_it;
  min_intensity_color = colors[size_t(min_it - intensities.begin())];
  median_intensity = Median(intensities, true);
  const auto med_it = std::find(intensities.begin(), intensities.end(), median_intensity);
  median_intensity_color = colors[size_t(med_it - intensities.begin())];
}

int VertexInfo::VisibleFrom(int kf_id) const {
  for (int i = 0; i < static_cast<int>(visible_keyframes.size()); i++)
    if (visible_keyframes[size_t(i)].kf_id == kf_id) return i;
  return -1;
}

void FaceInfo::CalcStat() {
  if (visible_keyframes.empty()) return;
  std::vector<float> distances, viewing_angles, areas;
  for (auto& kf : visible_keyframes) {
    distances.push_back(kf.distance);
    viewing_angles.push_back(kf.viewing_angle);
    areas.push_back(kf.area);
  }
  max_area_index = static_cast<int>(std::max_element(areas.begin(), areas.end()) - areas.begin());
  max_area_id = visible_keyframes[size_t(max_area_index)].kf_id;
  max_area = areas[size_t(max_area_index)];
  min_viewing_angle_index = static_cast<int>(
      std::min_element(viewing_angles.begin(), viewing_angles.end()) - viewing_angles.begin());
  min_viewing_angle_id = visible_keyframes[size_t(min_viewing_angle_index)].kf_id;
  min_viewing_angle = viewing_angles[size_t(min_viewing_angle_index)];
  min_distance_index =
      static_cast<int>(std::min_element(distances.begin(), distances.end()) - distances.begin());
  min_distance_id = visible_keyframes[size_t(min_distance_index)].kf_id;
  min_distance = distances[size_t(min_distance_index)];
}

void VisibilityInfo::CalcStatVertexInfo() {
  for (auto& v : vertex_info_list) v.CalcStat();
  has_vertex_stat = true;
}

void VisibilityInfo::CalcStatFaceInfo() {
  for (auto& f : face_info_list) f.CalcStat();
  has_face_stat = true;
}

class VisibilityTester::Impl {
 public:
  bool mesh_initialized_{false};
  std::shared_ptr<const Keyframe> keyframe_{nullptr};
  VisibilityTesterOption option_;
  std::vector<Eigen::Vector3f> vertices_, normals_, face_centers_;
  std::vector<Eigen::Vector3i> indices_;
  ugu_hip_ctx* ctx_{nullptr};

  Impl() { ctx_ = ugu_hip_create(device_ordinal()); }
  ~Impl() { ugu_hip_destroy(ctx_); }

  bool PrepareData() {
    if (vertices_.empty() || normals_.empty() || indices_.empty()) {
      LOGE("mesh is empty\n");
      return false;
    }
    if (ctx_ == nullptr) {
      LOGE("no HIP device available\n");
      return false;
    }
    UguMeshDesc d;
    std::memset(&d, 0, sizeof(d));
    d.num_vertices = static_cast<int32_t>(vertices_.size());
    d.vertices = vertices_[0].data();
    d.num_faces = static_cast<int32_t>(indices_.size());
    d.faces = indices_[0].data();
    if (ugu_hip_prepare_mesh(ctx_, &d) != UGU_HIP_OK || ugu_hip_prepare_raytracer(ctx_) != UGU_HIP_OK) {
      LOGE("%s", ugu_hip_last_error(ctx_));
      return false;
    }
    face_centers_.resize(indices_.size());
    for (size_t f = 0; f < indices_.size(); f++) {
      face_centers_[f].setZero();
      for (int i = 0; i < 3; i++) face_centers_[f] += vertices_[size_t(indices_[f][i])];
      face_centers_[f] /= 3;
    }
    mesh_initialized_ = true;
    return true;
  }

  // ValidateAndInitBeforeTest (visibility_tester.cc:331-389)
  bool Validate(VisibilityInfo* info) const {
    if (info == nullptr) {
      LOGE("info is nullptr\n");
      return false;
    }
    if (keyframe_ == nullptr) {
      LOGE("keyframe has not been set\n");
      return false;
    }
    if (!mesh_initialized_) {
      LOGE("mesh has not been initialized\n");
      return false;
    }
    if (keyframe_->color.empty()) {
      LOGE("keyframe color is empty\n");
      return false;
    }
    if (option_.use_mask && keyframe_->mask.empty()) {
      LOGE("use_mask is true but keyframe mask is empty\n");
      return false;
    }
    if (option_.use_depth && keyframe_->depth.empty()) {
      LOGE("use_depth is true but keyframe depth is empty\n");
      return false;
    }
    if (info->vertex_info_list.size() != vertices_.size() ||
        info->vertex_info_list.size() != normals_.size() ||
        info->face_info_list.size() != indices_.size()) {
      LOGE("info size is different from mesh\n");
      return false;
    }
    for (auto& ii : info->vertex_info_list) {
      ii.min_viewing_angle_color = option_.invalid_color;
      ii.min_distance_color = option_.invalid_color;
      ii.mean_color = option_.invalid_color;
      ii.median_color = option_.invalid_color;
      ii.mean_viewing_angle_color = option_.invalid_color;
      ii.median_viewing_angle_color = option_.invalid_color;
      ii.mean_distance_color = option_.invalid_color;
      ii.median_distance_color = option_.invalid_color;
    }
    return true;
  }

  bool TestItems(const std::vector<Eigen::Vector3f>& pos, const std::vector<Eigen::Vector3f>& vert,
                 const std::vector<Eigen::Vector3f>& nrm, std::vector<UguVertexVisibility>* out) const {
    UguKeyframe k;
    if (keyframe_->camera == nullptr || !keyframe_camera(*keyframe_->camera, &k.camera)) {
      LOGE("keyframe camera must be a PinholeCamera\n");
      return false;
    }
    k.width = keyframe_->color.cols;
    k.height = keyframe_->color.rows;
    k.color = keyframe_->color.data;
    k.depth = keyframe_->depth.empty() ? nullptr : reinterpret_cast<const float*>(keyframe_->depth.data);
    k.mask = keyframe_->mask.empty() ? nullptr : keyframe_->mask.data;
    UguVisibilityOptions o;
    o.backface_culling_angle_rad_th = option_.backface_culling_angle_rad_th;
    o.use_mask = option_.use_mask ? 1 : 0;
    o.use_depth = option_.use_depth ? 1 : 0;
    o.interp = static_cast<int32_t>(option_.interp);
    o.max_viewing_angle = option_.max_viewing_angle;
    o.max_distance_from_camera_to_vertex = option_.max_distance_from_camera_to_vertex;
    o.max_depth_difference = option_.max_depth_difference;
    o.eps = option_.eps;
    out->assign(pos.size(), UguVertexVisibility{});
    if (pos.empty()) return true;
    if (ugu_hip_visibility_test(ctx_, &k, &o, static_cast<int>(pos.size()), pos[0].data(),
                                vert[0].data(), nrm[0].data(), out->data()) != UGU_HIP_OK) {
      LOGE("%s", ugu_hip_last_error(ctx_));
      return false;
    }
    return true;
  }

  static VertexInfoPerKeyframe Record(int kf_id, const UguVertexVisibility& r) {
    VertexInfoPerKeyframe v;
    v.kf_id = kf_id;
    v.color = Eigen::Vector3f(r.color[0], r.color[1], r.color[2]);
    v.intensity = r.intensity;
    v.projected_pos = Eigen::Vector2f(r.projected[0], r.projected[1]);
    v.viewing_angle = r.viewing_angle;
    v.distance = r.distance;
    return v;
  }

  static FaceInfoPerKeyframe FaceRecord(int kf_id, int f, const VertexInfoPerKeyframe* c[3]) {
    FaceInfoPerKeyframe fi;
    fi.kf_id = kf_id;
    fi.face_id = f;
    fi.viewing_angle = 0.0f;
    fi.distance = 0.0f;
    for (int j = 0; j < 3; j++) {
      fi.viewing_angle += c[j]->viewing_angle;
      fi.distance += c[j]->distance;
      fi.projected_tri[size_t(j)] = c[j]->projected_pos;
    }
    fi.viewing_angle /= 3;
    fi.distance /= 3;
    fi.area = 0.5f * std::abs(Edge(fi.projected_tri[0], fi.projected_tri[1], fi.projected_tri[2]));
    return fi;
  }

  bool Test(VisibilityInfo* info, bool facewise) const {
    if (!Validate(info)) return false;
    const int kf_id = keyframe_->id;
    if (facewise) {
      // TestFacewiseVertices (visibility_tester.cc:573-641)
      const size_t F = indices_.size();
      std::vector<Eigen::Vector3f> pos(F * 3), vert(F * 3), nrm(F * 3);
      for (size_t f = 0; f < F; f++)
        for (int j = 0; j < 3; j++) {
          const size_t vid = size_t(indices_[f][j]);
          const Eigen::Vector3f diff = face_centers_[f] - vertices_[vid];
          const Eigen::Vector3f off = diff * 1e-3f;
          pos[f * 3 + size_t(j)] = vertices_[vid] + off;
          vert[f * 3 + size_t(j)] = vertices_[vid];
          nrm[f * 3 + size_t(j)] = normals_[vid];
        }
      std::vector<UguVertexVisibility> rec;
      if (!TestItems(pos, vert, nrm, &rec)) return false;
      std::vector<bool> done(vertices_.size(), false);
      for (size_t f = 0; f < F; f++) {
        VertexInfoPerKeyframe corner[3];
        const VertexInfoPerKeyframe* cp[3];
        int ok = 0;
        for (int j = 0; j < 3; j++) {
          const UguVertexVisibility& r = rec[f * 3 + size_t(j)];
          if (!r.visible) break;
          corner[j] = Record(kf_id, r);
          cp[j] = &corner[j];
          const size_t vid = size_t(indices_[f][j]);
          if (!done[vid]) {
            done[vid] = true;
            info->Update(static_cast<int>(vid), corner[j]);
          }
          ok++;
        }
        if (ok == 3) info->Update(static_cast<int>(f), FaceRecord(kf_id, static_cast<int>(f), cp));
      }
      if (option_.calc_stat_vertex_info) info->CalcStatVertexInfo();
      if (option_.calc_stat_face_info) info->CalcStatFaceInfo();
      return true;
    }
    // TestVertices + TestFaces (visibility_tester.cc:532-571)
    std::vector<UguVertexVisibility> rec;
    if (!TestItems(vertices_, vertices_, normals_, &rec)) return false;
    for (size_t i = 0; i < rec.size(); i++)
      if (rec[i].visible) info->Update(static_cast<int>(i), Record(kf_id, rec[i]));
    if (option_.calc_stat_vertex_info) info->CalcStatVertexInfo();
    if (option_.collect_face_info) {
      for (size_t f = 0; f < indices_.size(); f++) {
        const VertexInfoPerKeyframe* cp[3];
        bool vis = true;
        for (int j = 0; j < 3 && vis; j++) {
          const VertexInfo& vi = info->vertex_info_list[size_t(indices_[f][j])];
          const int idx = vi.VisibleFrom(kf_id);
          vis = idx >= 0;
          if (vis) cp[j] = &vi.visible_keyframes[size_t(idx)];
        }
        if (vis) info->Update(static_cast<int>(f), FaceRecord(kf_id, static_cast<int>(f), cp));
      }
      if (option_.calc_stat_face_info) info->CalcStatFaceInfo();
    }
    return true;
  }
};

VisibilityTester::VisibilityTester() : pimpl_(new Impl) {}
VisibilityTester::~VisibilityTester() {}
VisibilityTester::VisibilityTester(const VisibilityTesterOption& option) : pimpl_(new Impl) {
  set_option(option);
}
void VisibilityTester::set_option(const VisibilityTesterOption& option) {
  option.CopyTo(&pimpl_->option_);
}
void VisibilityTester::set_data(const std::vector<Eigen::Vector3f>& vertices,
                                const std::vector<Eigen::Vector3f>& normals,
                                const std::vector<Eigen::Vector3i>& indices) {
  pimpl_->mesh_initialized_ = false;
  pimpl_->vertices_ = vertices;
  pimpl_->normals_ = normals;
  pimpl_->indices_ = indices;
}
bool VisibilityTester::PrepareData() { return pimpl_->PrepareData(); }
void VisibilityTester::set_keyframe(std::shared_ptr<Keyframe> keyframe) {
  pimpl_->keyframe_ = keyframe;
}
bool VisibilityTester::Test(VisibilityInfo* info, bool facewise) const {
  return pimpl_->Test(info, facewise);
}
bool VisibilityTester::Test(std::vector<std::shared_ptr<Keyframe>> keyframes, VisibilityInfo* info,
                            bool facewise) {
  // visibility_tester.cc:693-727: statistics only after the last keyframe
  if (keyframes.empty()) return false;
  bool ret = true;
  VisibilityTesterOption org;
  pimpl_->option_.CopyTo(&org);
  pimpl_->option_.calc_stat_face_info = false;
  pimpl_->option_.calc_stat_vertex_info = false;
  for (size_t i = 0; i + 1 < keyframes.size(); i++) {
    set_keyframe(keyframes[i]);
    if (!Test(info, facewise)) {
      ret = false;
      break;
    }
  }
  set_option(org);
  set_keyframe(keyframes.back());
  if (!Test(info, facewise)) ret = false;
  return ret;
}

bool TextureMapping(const std::vector<std::shared_ptr<Keyframe>>& keyframes,
                    const VisibilityInfo& info, Mesh* mesh, const TextureMappingOption& option) {
  if (option.type != TextureMappingType::kSimpleProjection) {
    LOGE("TextureMappingType %d is not implemented", static_cast<int>(option.type));
    return false;
  }
  if (option.uv_type != TexturingOutputUvType::kUseOriginalMeshUv) {
    LOGE("OutputUvType %d is not implemented\n", static_cast<int>(option.uv_type));
    return false;
  }
  if (mesh->uv().empty() || mesh->uv_indices().size() != mesh->vertex_indices().size()) {
    LOGE("OutputUvType kUseOriginalMeshUv is specified but UV on mesh is invalid\n");
    return false;
  }
  const size_t F = info.face_info_list.size();
  std::vector<UguTexFace> faces(F);
  for (size_t i = 0; i < F; i++) {
    const FaceInfo& fi = info.face_info_list[i];
    UguTexFace& r = faces[i];
    if (fi.visible_keyframes.empty()) {
      r.keyframe = -1;
      continue;
    }
    int idx;
    if (option.criteria == ViewSelectionCriteria::kMinViewingAngle) {
      idx = fi.min_viewing_angle_index;
    } else if (option.criteria == ViewSelectionCriteria::kMinDistance) {
      idx = fi.min_distance_index;
    } else if (option.criteria == ViewSelectionCriteria::kMaxArea) {
      idx = fi.max_area_index;
    } else {
      LOGE("ViewSelectionCriteria %d is not implemented\n", static_cast<int>(option.criteria));
      return false;
    }
    const FaceInfoPerKeyframe& best = fi.visible_keyframes[size_t(idx)];
    if (best.kf_id < 0 || best.kf_id >= static_cast<int>(keyframes.size())) {
      LOGE("keyframe ids must index the keyframe list\n");
      return false;
    }
    r.keyframe = best.kf_id;
    for (int j = 0; j < 3; j++) {
      r.src[2 * j] = best.projected_tri[size_t(j)].x();
      r.src[2 * j + 1] = best.projected_tri[size_t(j)].y();
      const Eigen::Vector2f& uv = mesh->uv()[size_t(mesh->uv_indices()[i][j])];
      r.dst[2 * j] = uv.x() * option.tex_w - 0.5f;
      r.dst[2 * j + 1] = (1.0f - uv.y()) * option.tex_h - 0.5f;
    }
  }
  std::vector<UguTexture> images(keyframes.size());
  for (size_t i = 0; i < keyframes.size(); i++)
    images[i] = UguTexture{keyframes[i]->color.cols, keyframes[i]->color.rows,
                           keyframes[i]->color.data};
  Image3b texture = Image3b::zeros(option.tex_h, option.tex_w);
  Image1b mask = Image1b::zeros(option.tex_h, option.tex_w);
  ugu_hip_ctx* ctx = ugu_hip_create(device_ordinal());
  if (ctx == nullptr) {
    LOGE("no HIP device available\n");
    return false;
  }
  const int rc = ugu_hip_texture_map(ctx, images.data(), static_cast<int>(images.size()),
                                     faces.data(), static_cast<int>(F), option.tex_w, option.tex_h,
                                     option.padding_kernel, texture.data, mask.data);
  if (rc != UGU_HIP_OK) LOGE("%s", ugu_hip_last_error(ctx));
  ugu_hip_destroy(ctx);
  if (rc != UGU_HIP_OK) return false;
  std::vector<ObjMaterial> materials(1);
  materials[0].name = option.texture_base_name;
  materials[0].diffuse_tex = texture;
  mesh->set_materials(materials);
  mesh->set_material_ids(std::vector<int>(mesh->vertex_indices().size(), 0));
  return true;
}

}  // namespace ugu
