// ugu::VisibilityTester and ugu::TextureMapping (every uv type) for the C++
// API: the per-vertex test and the face rasterization into the texture run on
// the GPU (ugu_hip_visibility_test, ugu_hip_texture_map); the chart atlas is
// host work (a walk over face adjacency, patch copies, a rectangle packer);
// the loops that assemble VisibilityInfo and the statistics follow
// src/texturing/visibility_tester.cc and include/ugu/util/math_util.h, with
// the same standard-library algorithms (std::nth_element, std::sort,
// std::min_element), so a libstdc++ build reproduces the reference's choices.
// The face-wise test records a vertex from the lowest face id that passes it
// (the reference's OpenMP loop leaves that order unspecified).

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <numeric>
#include <sstream>
#include <unordered_map>

#include "gpu_common.h"
#include "ugu/texturing/texture_mapper.h"
#include "ugu/texturing/visibility_tester.h"

namespace ugu {

namespace {

using gpu_detail::device_ordinal;
using gpu_detail::to_abi;

// The statistics of VertexInfo::CalcStat (visibility_tester.cc:68-153 with
// math_util.h:29-181 and math_util.cc:44-56), from their definitions:
//   weights      w_i + 1e-6f, divided by their float sum taken in order; all
//                1/n when that sum is not above 1e-6
//   weighted     average: per channel the double sum of float(c_i * w_i),
//                rounded to float; median: the entries ordered by weight
//                (std::sort), the first whose running float sum of weights
//                exceeds 0.5 (the lightest when none does)
//   median       n = size/2; even sizes (or force_org_val) the n-th smallest;
//                odd sizes the mean of the (n+1)-th smallest and whatever
//                std::nth_element leaves at position n (so a libstdc++ build
//                picks the reference's element); sizes 1 and 2 directly
//   mode         the first value, in input order, whose running (weighted)
//                count becomes the largest seen
std::vector<float> normalized(const std::vector<float>& w) {
  const float eps = 0.000001f;
  std::vector<float> out(w.size());
  float total = 0.0f;
  for (size_t i = 0; i < w.size(); i++) {
    out[i] = w[i] + eps;
    total += out[i];
  }
  const bool usable = total > 0.000001;
  for (float& x : out) x = usable ? x / total : 1.0f / static_cast<float>(w.size());
  return out;
}

Eigen::Vector3f weighted_mean(const std::vector<Eigen::Vector3f>& c, const std::vector<float>& w) {
  const std::vector<float> nw = normalized(w);
  double acc[3] = {0.0, 0.0, 0.0};
  for (size_t i = 0; i < c.size(); i++)
    for (int k = 0; k < 3; k++) acc[k] += static_cast<float>(c[i][k] * nw[i]);
  return Eigen::Vector3f(static_cast<float>(acc[0]), static_cast<float>(acc[1]),
                         static_cast<float>(acc[2]));
}

Eigen::Vector3f weighted_median(const std::vector<Eigen::Vector3f>& c,
                                const std::vector<float>& w) {
  const std::vector<float> nw = normalized(w);
  std::vector<std::pair<float, Eigen::Vector3f>> by_weight;
  by_weight.reserve(c.size());
  for (size_t i = 0; i < c.size(); i++) by_weight.emplace_back(nw[i], c[i]);
  std::sort(by_weight.begin(), by_weight.end(),
            [](const std::pair<float, Eigen::Vector3f>& a,
               const std::pair<float, Eigen::Vector3f>& b) { return a.first < b.first; });
  float running = 0.0f;
  for (const auto& e : by_weight) {
    running += e.first;
    if (running > 0.5f) return e.second;
  }
  return by_weight.front().second;
}

float median_of(std::vector<float> d, bool force_org_val = false) {
  if (d.size() == 1) return d[0];
  if (d.size() == 2) return force_org_val ? d[0] : (d[0] + d[1]) * 0.5f;
  const size_t n = d.size() / 2;
  const bool single = force_org_val || d.size() % 2 == 0;
  std::nth_element(d.begin(), d.begin() + long(single ? n : n + 1), d.end());
  return single ? d[n] : (d[n] + d[n + 1]) * 0.5f;
}

template <typename Count>
void mode_of(const std::vector<Eigen::Vector3f>& c, const std::vector<Count>& weight,
             Eigen::Vector3f* mode, Count* freq,
             std::unordered_map<Eigen::Vector3f, Count>* occurrence) {
  *freq = 0;
  occurrence->clear();
  for (size_t i = 0; i < c.size(); i++) {
    Count& n = (*occurrence)[c[i]];
    n = n + weight[i];
    if (*freq < n) {
      *freq = n;
      *mode = c[i];
    }
  }
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

void VertexInfo::CalcStat(std::function<void(VertexInfo&)> vert_custom_func) {
  const size_t n = visible_keyframes.size();
  if (n == 0) return;
  std::vector<Eigen::Vector3f> colors(n);
  std::vector<float> dist(n), angle(n), inten(n), inv_dist(n), inv_angle(n);
  mean_color = Eigen::Vector3f(0, 0, 0);
  mean_distance = 0;
  mean_viewing_angle = 0;
  for (size_t i = 0; i < n; i++) {
    const VertexInfoPerKeyframe& kf = visible_keyframes[i];
    colors[i] = kf.color;
    dist[i] = kf.distance;
    angle[i] = kf.viewing_angle;
    inten[i] = kf.intensity;
    inv_dist[i] = 1.0f / kf.distance;
    inv_angle[i] = 1.0f / kf.viewing_angle;
    mean_color += kf.color;
    mean_distance += kf.distance;
    mean_viewing_angle += kf.viewing_angle;
  }
  mean_color /= static_cast<float>(n);
  mean_distance /= n;
  mean_viewing_angle /= n;

  // first minimum, as std::min_element
  auto argmin = [](const std::vector<float>& v) {
    return static_cast<int>(std::min_element(v.begin(), v.end()) - v.begin());
  };
  min_viewing_angle_index = argmin(angle);
  min_viewing_angle_id = visible_keyframes[size_t(min_viewing_angle_index)].kf_id;
  min_viewing_angle_color = colors[size_t(min_viewing_angle_index)];
  min_distance_index = argmin(dist);
  min_distance_id = visible_keyframes[size_t(min_distance_index)].kf_id;
  min_distance_color = colors[size_t(min_distance_index)];

  median_viewing_angle = median_of(angle);
  median_distance = median_of(dist);
  for (int k = 0; k < 3; k++) {
    std::vector<float> channel(n);
    for (size_t i = 0; i < n; i++) channel[i] = colors[i][k];
    median_color[k] = median_of(channel);
  }

  mean_viewing_angle_color = weighted_mean(colors, inv_angle);
  mean_distance_color = weighted_mean(colors, inv_dist);
  median_viewing_angle_color = weighted_median(colors, inv_angle);
  median_distance_color = weighted_median(colors, inv_dist);

  const int imin = argmin(inten);
  min_intensity = inten[size_t(imin)];
  min_intensity_color = colors[size_t(imin)];
  median_intensity = median_of(inten, true);
  // the colour of the first keyframe with that intensity
  median_intensity_color =
      colors[size_t(std::find(inten.begin(), inten.end(), median_intensity) - inten.begin())];

  mode_of(colors, std::vector<int>(n, 1), &mode, &mode_frequency, &occurrence);
  mode_of(colors, inv_angle, &mode_viewing_angle_color, &mode_frequency_viewing_angle_color,
          &occurrence_viewing_angle);

  if (vert_custom_func != nullptr) vert_custom_func(*this);
}

int VertexInfo::VisibleFrom(int kf_id) const {
  for (int i = 0; i < static_cast<int>(visible_keyframes.size()); i++)
    if (visible_keyframes[size_t(i)].kf_id == kf_id) return i;
  return -1;
}

// FaceInfo::CalcStat (visibility_tester.cc:185-224) sets the indices and the
// keyframe ids only; max_area, min_viewing_angle and min_distance keep their
// initial values there, and so they do here.
void FaceInfo::CalcStat() {
  const size_t n = visible_keyframes.size();
  if (n == 0) return;
  size_t imax_area = 0, imin_angle = 0, imin_dist = 0;
  for (size_t i = 1; i < n; i++) {
    const FaceInfoPerKeyframe& k = visible_keyframes[i];
    if (visible_keyframes[imax_area].area < k.area) imax_area = i;
    if (k.viewing_angle < visible_keyframes[imin_angle].viewing_angle) imin_angle = i;
    if (k.distance < visible_keyframes[imin_dist].distance) imin_dist = i;
  }
  max_area_index = static_cast<int>(imax_area);
  max_area_id = visible_keyframes[imax_area].kf_id;
  min_viewing_angle_index = static_cast<int>(imin_angle);
  min_viewing_angle_id = visible_keyframes[imin_angle].kf_id;
  min_distance_index = static_cast<int>(imin_dist);
  min_distance_id = visible_keyframes[imin_dist].kf_id;
}

void VisibilityInfo::CalcStatVertexInfo(std::function<void(VertexInfo&)> vert_custom_func) {
  for (auto& v : vertex_info_list) v.CalcStat(vert_custom_func);
  has_vertex_stat = true;
}

void VisibilityInfo::CalcStatFaceInfo() {
  for (auto& f : face_info_list) f.CalcStat();
  has_face_stat = true;
}

std::ostream& operator<<(std::ostream& os, const VertexInfoPerKeyframe& vi) {
  os << "{ \"kf_id\": " << vi.kf_id << ", \"projected_pos\":[" << vi.projected_pos[0] << ","
     << vi.projected_pos[1] << "]," << "\"viewing_angle\":" << vi.viewing_angle << ","
     << "\"distance\":" << vi.distance << "}";
  return os;
}

std::ostream& operator<<(std::ostream& os, const VertexInfo& vi) {
  os << "{\"visible_keyframes\":[";
  for (size_t i = 0; i < vi.visible_keyframes.size(); i++)
    os << (i ? "," : "") << vi.visible_keyframes[i];
  os << "]}";
  return os;
}

std::string VisibilityInfo::SerializeAsJson() const {
  std::stringstream ss;
  ss << "{\"vertex_info_list\":[";
  for (size_t i = 0; i < vertex_info_list.size(); i++) ss << (i ? "," : "") << vertex_info_list[i];
  ss << "]}";
  return ss.str();
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

  bool Test(VisibilityInfo* info, bool facewise,
            const std::function<void(VertexInfo&)>& custom) const {
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
      if (option_.calc_stat_vertex_info) info->CalcStatVertexInfo(custom);
      if (option_.calc_stat_face_info) info->CalcStatFaceInfo();
      return true;
    }
    // TestVertices + TestFaces (visibility_tester.cc:532-571)
    std::vector<UguVertexVisibility> rec;
    if (!TestItems(vertices_, vertices_, normals_, &rec)) return false;
    for (size_t i = 0; i < rec.size(); i++)
      if (rec[i].visible) info->Update(static_cast<int>(i), Record(kf_id, rec[i]));
    if (option_.calc_stat_vertex_info) info->CalcStatVertexInfo(custom);
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
bool VisibilityTester::Test(VisibilityInfo* info, bool facewise,
                            std::function<void(VertexInfo&)> vert_custom_func) const {
  return pimpl_->Test(info, facewise, vert_custom_func);
}
bool VisibilityTester::Test(std::vector<std::shared_ptr<Keyframe>> keyframes, VisibilityInfo* info,
                            bool facewise, std::function<void(VertexInfo&)> vert_custom_func) {
  // visibility_tester.cc:693-727: statistics only after the last keyframe
  if (keyframes.empty()) return false;
  bool ret = true;
  VisibilityTesterOption org;
  pimpl_->option_.CopyTo(&org);
  pimpl_->option_.calc_stat_face_info = false;
  pimpl_->option_.calc_stat_vertex_info = false;
  for (size_t i = 0; i + 1 < keyframes.size(); i++) {
    set_keyframe(keyframes[i]);
    if (!Test(info, facewise, vert_custom_func)) {
      ret = false;
      break;
    }
  }
  set_option(org);
  set_keyframe(keyframes.back());
  if (!Test(info, facewise, vert_custom_func)) ret = false;
  return ret;
}

namespace {

// The mesh's single material: `texture`, material id 0 for every face.
void SetTexture(const Image3b& texture, const TextureMappingOption& option, Mesh* mesh) {
  std::vector<ObjMaterial> materials(1);
  materials[0].name = option.texture_base_name;
  materials[0].diffuse_tex = texture;
  mesh->set_materials(materials);
  mesh->set_material_ids(std::vector<int>(mesh->vertex_indices().size(), 0));
}

// kGenerateSimpleTile (texture_mapper.cc:30-136): the keyframe images laid out
// row-major on a grid of floor(sqrt(n)) columns, and every face's uv the
// chosen keyframe's projected corners moved into its tile, at texel centres
// ((x + 0.5) / W, 1 - (y + 0.5) / H); faces seen by no keyframe get uv 0.
// Departures, both where the reference's memory accesses are undefined: the
// grid has ceil(n / columns) rows (the reference's n / columns rows leave the
// last keyframes outside the image when n is not a multiple of the column
// count), and keyframes must all have the first one's size.
bool SimpleTile(const std::vector<std::shared_ptr<Keyframe>>& keyframes,
                const std::vector<const FaceInfoPerKeyframe*>& best, Mesh* mesh,
                const TextureMappingOption& option) {
  const int n = static_cast<int>(keyframes.size());
  const int w0 = keyframes[0]->color.cols, h0 = keyframes[0]->color.rows;
  for (const auto& kf : keyframes)
    if (kf->color.cols != w0 || kf->color.rows != h0) {
      LOGE("kGenerateSimpleTile needs keyframes of one size\n");
      return false;
    }
  const int cols = std::max(1, static_cast<int>(std::sqrt(keyframes.size())));
  const int rows = (n + cols - 1) / cols;
  const int tex_w = cols * w0, tex_h = rows * h0;
  Image3b texture = Image3b::zeros(tex_h, tex_w);
  const size_t row_bytes = size_t(w0) * 3;
  for (int i = 0; i < n; i++) {
    const size_t x0 = size_t(i % cols) * size_t(w0), y0 = size_t(i / cols) * size_t(h0);
    for (int y = 0; y < h0; y++)
      std::memcpy(texture.data + ((y0 + size_t(y)) * size_t(tex_w) + x0) * 3,
                  keyframes[size_t(i)]->color.data + size_t(y) * row_bytes, row_bytes);
  }
  // keyframe id -> list position (the first keyframe with that id); an id
  // missing from the list lands on tile 0, as the reference's map lookup does
  std::unordered_map<int, int> position;
  for (int i = 0; i < n; i++) position.emplace(keyframes[size_t(i)]->id, i);

  const size_t F = best.size();
  std::vector<Eigen::Vector2f> uv(F * 3, Eigen::Vector2f(0.0f, 0.0f));
  std::vector<Eigen::Vector3i> uv_indices(F);
  for (size_t f = 0; f < F; f++) {
    const int base = static_cast<int>(f * 3);
    uv_indices[f] = Eigen::Vector3i(base, base + 1, base + 2);
    const FaceInfoPerKeyframe* b = best[f];
    if (b == nullptr || b->kf_id < 0) continue;
    const auto it = position.find(b->kf_id);
    const int slot = it == position.end() ? 0 : it->second;
    const int ox = (slot % cols) * w0, oy = (slot / cols) * h0;
    for (int j = 0; j < 3; j++) {
      const float tx = b->projected_tri[size_t(j)].x() + ox;
      const float ty = b->projected_tri[size_t(j)].y() + oy;
      uv[f * 3 + size_t(j)] = Eigen::Vector2f((tx + 0.5f) / tex_w, 1.0f - ((ty + 0.5f) / tex_h));
    }
  }
  mesh->set_uv(uv);
  mesh->set_uv_indices(uv_indices);
  SetTexture(texture, option, mesh);
  return true;
}

// RasterizeTriangle for every face record (texture_mapper.cc:234-292; the
// face with the larger index wins where target triangles overlap), then
// PaddingSimple with `padding_kernel` (:148-201), on the GPU.
bool RasterizeFaces(const std::vector<std::shared_ptr<Keyframe>>& keyframes,
                    const std::vector<UguTexFace>& faces, int tex_w, int tex_h,
                    int padding_kernel, Image3b* texture) {
  std::vector<UguTexture> images(keyframes.size());
  for (size_t i = 0; i < keyframes.size(); i++)
    images[i] = UguTexture{keyframes[i]->color.cols, keyframes[i]->color.rows,
                           keyframes[i]->color.data};
  *texture = Image3b::zeros(tex_h, tex_w);
  Image1b mask = Image1b::zeros(tex_h, tex_w);
  ugu_hip_ctx* ctx = ugu_hip_create(device_ordinal());
  if (ctx == nullptr) {
    LOGE("no HIP device available\n");
    return false;
  }
  const int rc = ugu_hip_texture_map(ctx, images.data(), static_cast<int>(images.size()),
                                     faces.data(), static_cast<int>(faces.size()), tex_w, tex_h,
                                     padding_kernel, texture->data, mask.data);
  if (rc != UGU_HIP_OK) LOGE("%s", ugu_hip_last_error(ctx));
  ugu_hip_destroy(ctx);
  return rc == UGU_HIP_OK;
}

// kUseOriginalMeshUv (texture_mapper.cc:295-345): every face's uv triangle
// rasterized into the texture on the GPU, then PaddingSimple.
bool OriginalUv(const std::vector<std::shared_ptr<Keyframe>>& keyframes,
                const std::vector<const FaceInfoPerKeyframe*>& best, Mesh* mesh,
                const TextureMappingOption& option) {
  if (mesh->uv().empty() || mesh->uv_indices().size() != mesh->vertex_indices().size()) {
    LOGE("OutputUvType kUseOriginalMeshUv is specified but UV on mesh is invalid\n");
    return false;
  }
  const size_t F = best.size();
  std::vector<UguTexFace> faces(F);
  for (size_t i = 0; i < F; i++) {
    UguTexFace& r = faces[i];
    const FaceInfoPerKeyframe* b = best[i];
    if (b == nullptr || b->kf_id < 0) {
      r.keyframe = -1;
      continue;
    }
    if (b->kf_id >= static_cast<int>(keyframes.size())) {
      LOGE("keyframe ids must index the keyframe list\n");
      return false;
    }
    r.keyframe = b->kf_id;
    for (int j = 0; j < 3; j++) {
      r.src[2 * j] = b->projected_tri[size_t(j)].x();
      r.src[2 * j + 1] = b->projected_tri[size_t(j)].y();
      const Eigen::Vector2f& uv = mesh->uv()[size_t(mesh->uv_indices()[i][j])];
      r.dst[2 * j] = uv.x() * option.tex_w - 0.5f;
      r.dst[2 * j + 1] = (1.0f - uv.y()) * option.tex_h - 0.5f;
    }
  }
  Image3b texture;
  if (!RasterizeFaces(keyframes, faces, option.tex_w, option.tex_h, option.padding_kernel,
                      &texture))
    return false;
  SetTexture(texture, option, mesh);
  return true;
}

// kGenerateSimpleTriangles (texture_mapper.cc:347-452). Face i takes one half
// of rectangle i / 2 of a row-major grid. A rectangle holds (s + 2) x (s + 1)
// texels and is followed by a 2-texel gap, where s is the largest side not
// above min(floor(sqrt(tex_w * tex_h / rectangles)), 100) for which every
// rectangle fits. Even faces take the upper-right triangle of their rectangle,
// odd faces the lower-left one (corners below). The chosen keyframe's
// projected triangle is rasterized into that half on the GPU, then
// PaddingSimple runs with a 3 x 3 kernel whatever option.padding_kernel is.
// Faces seen by no keyframe leave their half empty and get uv 0. As in the
// reference, the layout fails when a rectangle would hold fewer than 6 texels
// or s would drop below 3; it also fails for a mesh without faces (where the
// reference divides by zero).
struct TriangleGrid {
  int side = 0;     // s
  int pitch_x = 0;  // s + 2 texels and the gap
  int pitch_y = 0;  // s + 1 texels and the gap
  int per_row = 0;  // rectangles per grid row
};

bool triangle_grid(size_t n_faces, int tex_w, int tex_h, TriangleGrid* g) {
  constexpr int kGap = 2;
  const int64_t rects = static_cast<int64_t>((n_faces + 1) / 2);
  if (rects < 1 || tex_w < 1 || tex_h < 1) return false;
  const int64_t texels = int64_t(tex_h) * int64_t(tex_w) / rects;
  if (texels < 6) return false;
  int side = std::min(static_cast<int>(std::sqrt(static_cast<double>(texels))), 100);
  auto fits = [&](int sd) {
    return int64_t(tex_w / (sd + 2 + kGap)) * int64_t(tex_h / (sd + 1 + kGap)) >= rects;
  };
  while (!fits(side)) {
    side--;
    if (side < 3) return false;
  }
  g->side = side;
  g->pitch_x = side + 2 + kGap;
  g->pitch_y = side + 1 + kGap;
  g->per_row = tex_w / g->pitch_x;
  return true;
}

// Texel corners of face i's triangle: with (x0, y0) the rectangle's corner,
// even i (x0 + 2, y0), (x0 + s + 1, y0), (x0 + s + 1, y0 + s - 1); odd i
// (x0, y0 + 1), (x0, y0 + s), (x0 + s - 1, y0 + s).
void triangle_corners(const TriangleGrid& g, int i, float xy[6]) {
  const int r = i / 2, rx = r % g.per_row, ry = r / g.per_row, s = g.side;
  int x0, y0;
  if (i % 2 == 0) {
    x0 = g.pitch_x * rx + 2;
    y0 = g.pitch_y * ry;
    const int c[6] = {x0, y0, x0 + s - 1, y0, x0 + s - 1, y0 + s - 1};
    for (int k = 0; k < 6; k++) xy[k] = static_cast<float>(c[k]);
  } else {
    x0 = g.pitch_x * rx;
    y0 = g.pitch_y * ry + 1;
    const int c[6] = {x0, y0, x0, y0 + s - 1, x0 + s - 1, y0 + s - 1};
    for (int k = 0; k < 6; k++) xy[k] = static_cast<float>(c[k]);
  }
}

bool SimpleTriangles(const std::vector<std::shared_ptr<Keyframe>>& keyframes,
                     const std::vector<const FaceInfoPerKeyframe*>& best, Mesh* mesh,
                     const TextureMappingOption& option) {
  TriangleGrid g;
  if (!triangle_grid(mesh->vertex_indices().size(), option.tex_w, option.tex_h, &g)) {
    LOGE("kGenerateSimpleTriangles: a %d x %d texture cannot hold %zu faces\n", option.tex_w,
         option.tex_h, mesh->vertex_indices().size());
    return false;
  }
  const size_t F = best.size();
  std::vector<UguTexFace> faces(F);
  std::vector<Eigen::Vector2f> uv(F * 3, Eigen::Vector2f(0.0f, 0.0f));
  std::vector<Eigen::Vector3i> uv_indices(F);
  const float fw = static_cast<float>(option.tex_w), fh = static_cast<float>(option.tex_h);
  for (size_t i = 0; i < F; i++) {
    const int base = static_cast<int>(i * 3);
    uv_indices[i] = Eigen::Vector3i(base, base + 1, base + 2);
    UguTexFace& r = faces[i];
    const FaceInfoPerKeyframe* b = best[i];
    if (b == nullptr || b->kf_id < 0) {
      r.keyframe = -1;
      continue;
    }
    if (b->kf_id >= static_cast<int>(keyframes.size())) {
      LOGE("keyframe ids must index the keyframe list\n");
      return false;
    }
    r.keyframe = b->kf_id;
    float xy[6];
    triangle_corners(g, static_cast<int>(i), xy);
    for (int j = 0; j < 3; j++) {
      r.src[2 * j] = b->projected_tri[size_t(j)].x();
      r.src[2 * j + 1] = b->projected_tri[size_t(j)].y();
      r.dst[2 * j] = xy[2 * j];
      r.dst[2 * j + 1] = xy[2 * j + 1];
      uv[i * 3 + size_t(j)] =
          Eigen::Vector2f((xy[2 * j] + 0.5f) / fw, 1.0f - ((xy[2 * j + 1] + 0.5f) / fh));
    }
  }
  Image3b texture;
  if (!RasterizeFaces(keyframes, faces, option.tex_w, option.tex_h, 3, &texture)) return false;
  mesh->set_uv(uv);
  mesh->set_uv_indices(uv_indices);
  SetTexture(texture, option, mesh);
  return true;
}

// Face adjacency of the chart growth (face_adjacency.h:34-263, the sparse
// version): the directed edge (a, b) of face i -- its corners taken in order
// -- holds i + 1, summed over the faces that share that directed edge (as
// the reference's setFromTriplets sums duplicates). The neighbours of a face
// are the non-zero entries of its three reversed edges, in edge order, minus
// one; removing a face clears its own three entries.
class EdgeFaces {
 public:
  explicit EdgeFaces(const std::vector<Eigen::Vector3i>& f) : f_(f) {
    m_.reserve(f.size() * 3);
    for (size_t i = 0; i < f.size(); i++)
      for (int k = 0; k < 3; k++) m_[key(f[i][k], f[i][(k + 1) % 3])] += int64_t(i) + 1;
  }
  void neighbours(size_t face, std::vector<int64_t>* out) const {
    out->clear();
    const Eigen::Vector3i& v = f_[face];
    for (int k = 0; k < 3; k++) {
      const auto it = m_.find(key(v[(k + 1) % 3], v[k]));
      if (it != m_.end() && it->second > 0) out->push_back(it->second - 1);
    }
  }
  void remove(size_t face) {
    const Eigen::Vector3i& v = f_[face];
    for (int k = 0; k < 3; k++) {
      const auto it = m_.find(key(v[k], v[(k + 1) % 3]));
      if (it != m_.end()) it->second = 0;
    }
  }

 private:
  static uint64_t key(int a, int b) {
    return (uint64_t(uint32_t(a)) << 32) | uint64_t(uint32_t(b));
  }
  const std::vector<Eigen::Vector3i>& f_;
  std::unordered_map<uint64_t, int64_t> m_;
};

// One chart (texture_mapper.cc:454-619): faces grown from a seed through
// shared edges while the neighbour chose the same keyframe.
struct Chart {
  int kf = -1;
  std::vector<size_t> members;          // face positions, in growth order
  Image3b patch;                        // keyframe pixels of the chart's bbox
  std::vector<Eigen::Vector2f> pos;     // patch position per distinct vertex id
  std::vector<Eigen::Vector3i> corner;  // per member: its corners' entries of pos
};

// Chart::Finalize for a chart with a keyframe (texture_mapper.cc:469-617):
// the patch is the keyframe's pixels over the integer bbox (floor of the
// smallest, ceil of the largest projected corner) with a 1-pixel border
// that repeats the edge rows, then the edge columns (PaddingBorderSimple,
// :203-232). Every corner moves to (x - xmin + 1, y - ymin + 1) in the
// patch; vertices are numbered in increasing vertex id, one position each
// (that of the first member that names the vertex). The reference copies the
// bbox rows without bounds checks, so a corner outside [0, cols - 1] x
// [0, rows - 1] of the keyframe image (or not a number) is an error here;
// the visibility tester projects only corners inside that range.
bool finalize_chart(Chart* c, const Image3b& color,
                    const std::vector<const FaceInfoPerKeyframe*>& best,
                    const std::vector<Eigen::Vector3i>& vertex_indices) {
  constexpr int kPad = 1;
  const float xlim = static_cast<float>(color.cols - 1), ylim = static_cast<float>(color.rows - 1);
  float xlo = 0.0f, xhi = 0.0f, ylo = 0.0f, yhi = 0.0f;
  bool first = true;
  for (size_t m : c->members)
    for (const Eigen::Vector2f& p : best[m]->projected_tri) {
      if (!(p.x() >= 0.0f && p.x() <= xlim && p.y() >= 0.0f && p.y() <= ylim)) {
        LOGE("face %d: projected corner (%g, %g) outside the %d x %d keyframe image\n",
             best[m]->face_id, static_cast<double>(p.x()), static_cast<double>(p.y()), color.cols,
             color.rows);
        return false;
      }
      xlo = first ? p.x() : std::min(xlo, p.x());
      xhi = first ? p.x() : std::max(xhi, p.x());
      ylo = first ? p.y() : std::min(ylo, p.y());
      yhi = first ? p.y() : std::max(yhi, p.y());
      first = false;
    }
  const int xmin = static_cast<int>(std::floor(xlo)), xmax = static_cast<int>(std::ceil(xhi));
  const int ymin = static_cast<int>(std::floor(ylo)), ymax = static_cast<int>(std::ceil(yhi));
  Image3b& pt = c->patch;
  pt = Image3b::zeros(ymax - ymin + 1 + 2 * kPad, xmax - xmin + 1 + 2 * kPad);
  for (int y = ymin; y <= ymax; y++)
    std::memcpy(&pt.at<Vec3b>(y - ymin + kPad, kPad), &color.at<Vec3b>(y, xmin),
                size_t(xmax - xmin + 1) * 3);
  for (int y = 0; y < kPad; y++)
    for (int x = 0; x < pt.cols; x++) pt.at<Vec3b>(y, x) = pt.at<Vec3b>(kPad, x);
  for (int y = pt.rows - kPad; y < pt.rows; y++)
    for (int x = 0; x < pt.cols; x++) pt.at<Vec3b>(y, x) = pt.at<Vec3b>(pt.rows - kPad - 1, x);
  for (int y = 0; y < pt.rows; y++) {
    for (int x = 0; x < kPad; x++) pt.at<Vec3b>(y, x) = pt.at<Vec3b>(y, kPad);
    for (int x = pt.cols - kPad; x < pt.cols; x++)
      pt.at<Vec3b>(y, x) = pt.at<Vec3b>(y, pt.cols - kPad - 1);
  }
  std::map<int, Eigen::Vector2f> by_vertex;
  for (size_t m : c->members) {
    const Eigen::Vector3i& vi = vertex_indices[size_t(best[m]->face_id)];
    for (int j = 0; j < 3; j++) {
      const Eigen::Vector2f& p = best[m]->projected_tri[size_t(j)];
      by_vertex.emplace(vi[j], Eigen::Vector2f(p.x() - xmin + kPad, p.y() - ymin + kPad));
    }
  }
  std::unordered_map<int, int> slot;
  c->pos.clear();
  for (const auto& kv : by_vertex) {
    slot.emplace(kv.first, static_cast<int>(c->pos.size()));
    c->pos.push_back(kv.second);
  }
  c->corner.clear();
  for (size_t m : c->members) {
    const Eigen::Vector3i& vi = vertex_indices[size_t(best[m]->face_id)];
    c->corner.emplace_back(slot[vi[0]], slot[vi[1]], slot[vi[2]]);
  }
  return true;
}

struct PackRect {
  int x, y, w, h;
  int area() const { return w * h; }
};

// BinPacking2D (bin_packer_2d.cc:14-115), guillotine best fit: rectangles in
// the given order; each goes to the free rectangle it fits with the smallest
// spare area (the first such in the free list); a free rectangle larger than
// the target is cut along the axis with the larger leftover -- keeping a
// strip of the target's height (full width) when the height leftover is
// larger, else a strip of its width (full height) -- the strip and the rest
// replace it at the end of the list (rest first), and the search repeats
// until a free rectangle matches the target exactly. The free list starts as
// one rectangle at (x0, y0) of size w x h.
bool pack_rects(const std::vector<PackRect>& rects, int x0, int y0, int w, int h,
                std::vector<PackRect>* packed, std::vector<PackRect>* free_rects) {
  free_rects->assign(1, PackRect{x0, y0, w, h});
  packed->clear();
  for (const PackRect& t : rects) {
    if (free_rects->empty()) return false;
    for (;;) {
      int best_i = -1, best_diff = std::numeric_limits<int>::max();
      for (int i = 0; i < static_cast<int>(free_rects->size()); i++) {
        const PackRect& f = (*free_rects)[size_t(i)];
        if (t.h <= f.h && t.w <= f.w && f.area() - t.area() < best_diff) {
          best_diff = f.area() - t.area();
          best_i = i;
        }
      }
      if (best_i < 0) return false;
      const PackRect f = (*free_rects)[size_t(best_i)];
      free_rects->erase(free_rects->begin() + best_i);
      if (f.w == t.w && f.h == t.h) {
        packed->push_back(f);
        break;
      }
      if (f.h - t.h > f.w - t.w) {
        free_rects->push_back(PackRect{f.x, f.y + t.h, f.w, f.h - t.h});
        free_rects->push_back(PackRect{f.x, f.y, f.w, t.h});
      } else {
        free_rects->push_back(PackRect{f.x + t.w, f.y, f.w - t.w, f.h});
        free_rects->push_back(PackRect{f.x, f.y, t.w, f.h});
      }
    }
  }
  return true;
}

// kGenerateSimpleCharts (texture_mapper.cc:816-922 with GenerateAtlas,
// :658-814). Charts grow breadth-first from the lowest face id not yet in a
// chart, through the face adjacency above, over neighbours that chose the
// same keyframe (faces seen by no keyframe form charts of their own). The
// charts with a keyframe are ordered by patch area, largest first
// (std::sort, as the reference, so a libstdc++ build breaks ties the same
// way), and packed into a texture of option.tex_w x option.tex_h with a
// 1-texel margin at the top and left. When they do not fit, both sides grow
// to floor(tex * sqrt(2)^k) for k = 1, 2, ... up to 8192. Each patch is
// copied to its place; a chart vertex gets uv ((x + X + 0.5) / W,
// 1 - (y + Y + 0.5) / H) for its patch position (x, y) and the patch's
// place (X, Y). Faces without a keyframe share one uv triple at the first
// free rectangle's corner (or 0).
bool SimpleCharts(const std::vector<std::shared_ptr<Keyframe>>& keyframes,
                  const std::vector<const FaceInfoPerKeyframe*>& best, Mesh* mesh,
                  const TextureMappingOption& option) {
  const std::vector<Eigen::Vector3i>& vi = mesh->vertex_indices();
  const size_t F = vi.size();
  if (best.size() != F) {
    LOGE("VisibilityInfo has %zu faces, the mesh %zu\n", best.size(), F);
    return false;
  }
  // faces seen by no keyframe: records with keyframe -1 and their own id
  std::vector<FaceInfoPerKeyframe> unseen(F);
  std::vector<const FaceInfoPerKeyframe*> rec(F);
  for (size_t i = 0; i < F; i++) {
    if (best[i] != nullptr) {
      rec[i] = best[i];
    } else {
      unseen[i].kf_id = -1;
      unseen[i].face_id = static_cast<int>(i);
      rec[i] = &unseen[i];
    }
    if (rec[i]->face_id < 0 || rec[i]->face_id >= static_cast<int>(F) ||
        rec[i]->kf_id >= static_cast<int>(keyframes.size())) {
      LOGE("face %zu: face id %d or keyframe id %d out of range\n", i, rec[i]->face_id,
           rec[i]->kf_id);
      return false;
    }
  }

  EdgeFaces adjacency(vi);
  std::vector<char> taken(F, 0);
  std::vector<Chart> valid, invalid;
  std::vector<int64_t> near;
  std::deque<size_t> queue;
  for (size_t seed = 0; seed < F; seed++) {
    if (taken[seed]) continue;
    Chart c;
    c.kf = rec[seed]->kf_id;
    c.members.push_back(seed);
    taken[seed] = 1;
    queue.assign(1, seed);
    while (!queue.empty()) {
      const size_t f = queue.front();
      queue.pop_front();
      adjacency.neighbours(f, &near);
      adjacency.remove(f);
      // an entry summed from several faces may name no face (the reference
      // reads past its arrays there); it is skipped
      size_t n_add = 0;
      int64_t add[3];
      for (int64_t a : near)
        if (a < int64_t(F) && rec[size_t(a)]->kf_id == c.kf && !taken[size_t(a)]) add[n_add++] = a;
      for (size_t k = 0; k < n_add; k++) {
        queue.push_back(size_t(add[k]));
        c.members.push_back(size_t(add[k]));
        taken[size_t(add[k])] = 1;
      }
    }
    if (c.kf < 0) {
      c.patch = Image3b::zeros(1, 1);
      invalid.push_back(std::move(c));
    } else {
      if (!finalize_chart(&c, keyframes[size_t(c.kf)]->color, rec, vi)) return false;
      valid.push_back(std::move(c));
    }
  }

  std::vector<size_t> order(valid.size());
  std::iota(order.begin(), order.end(), size_t(0));
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return valid[a].patch.rows * valid[a].patch.cols > valid[b].patch.rows * valid[b].patch.cols;
  });
  std::vector<PackRect> rects;
  rects.reserve(order.size());
  for (size_t k : order) rects.push_back(PackRect{0, 0, valid[k].patch.cols, valid[k].patch.rows});

  constexpr int kMaxSide = 8192, kEdge = 1;
  const float ratio = std::sqrt(2.0f);
  int tw = option.tex_w, th = option.tex_h, tries = 0;
  std::vector<PackRect> packed, free_rects;
  while (!pack_rects(rects, kEdge, kEdge, tw - kEdge, th - kEdge, &packed, &free_rects)) {
    tries++;
    const double grow = std::pow(static_cast<double>(ratio), static_cast<double>(tries));
    tw = static_cast<int>(option.tex_w * grow);
    th = static_cast<int>(option.tex_h * grow);
    if (kMaxSide < tw || kMaxSide < th) {
      LOGE("Failed to pack in one texture\n");
      return false;
    }
  }
  if (tries > 0)
    LOGW("Texture size is adjusted (%d, %d) -> (%d, %d)\n", option.tex_w, option.tex_h, tw, th);

  Image3b texture = Image3b::zeros(th, tw);
  std::vector<Eigen::Vector2f> uv;
  std::vector<Eigen::Vector3i> uv_indices(F, Eigen::Vector3i(0, 0, 0));
  const float fw = static_cast<float>(tw), fh = static_cast<float>(th);
  for (size_t k = 0; k < order.size(); k++) {
    const Chart& c = valid[order[k]];
    const PackRect& at = packed[k];
    for (int y = 0; y < c.patch.rows; y++)
      std::memcpy(&texture.at<Vec3b>(y + at.y, at.x), &c.patch.at<Vec3b>(y, 0),
                  size_t(c.patch.cols) * 3);
    const int offset = static_cast<int>(uv.size());
    for (const Eigen::Vector2f& p : c.pos)
      uv.emplace_back((p.x() + at.x + 0.5f) / fw, 1.0f - ((p.y() + at.y + 0.5f) / fh));
    for (size_t m = 0; m < c.members.size(); m++) {
      const Eigen::Vector3i& k3 = c.corner[m];
      uv_indices[size_t(rec[c.members[m]]->face_id)] =
          Eigen::Vector3i(offset + k3[0], offset + k3[1], offset + k3[2]);
    }
  }
  Eigen::Vector2f none(0.0f, 0.0f);
  if (!free_rects.empty())
    none = Eigen::Vector2f((free_rects[0].x + 0.5f) / fw, 1.0f - ((free_rects[0].y + 0.5f) / fh));
  const int n0 = static_cast<int>(uv.size());
  uv.insert(uv.end(), 3, none);
  for (const Chart& c : invalid)
    for (size_t m : c.members) uv_indices[size_t(rec[m]->face_id)] = Eigen::Vector3i(n0, n0 + 1, n0 + 2);
  mesh->set_uv(uv);
  mesh->set_uv_indices(uv_indices);
  SetTexture(texture, option, mesh);
  return true;
}

}  // namespace

// TextureMapping (texture_mapper.cc:894-1017): choose each face's keyframe by
// the criterion, then build the texture for the requested uv type.
bool TextureMapping(const std::vector<std::shared_ptr<Keyframe>>& keyframes,
                    const VisibilityInfo& info, Mesh* mesh, const TextureMappingOption& option) {
  if (option.type != TextureMappingType::kSimpleProjection) {
    LOGE("TextureMappingType %d is not implemented", static_cast<int>(option.type));
    return false;
  }
  if (option.criteria != ViewSelectionCriteria::kMinViewingAngle &&
      option.criteria != ViewSelectionCriteria::kMinDistance &&
      option.criteria != ViewSelectionCriteria::kMaxArea) {
    LOGE("ViewSelectionCriteria %d is not implemented\n", static_cast<int>(option.criteria));
    return false;
  }
  if (keyframes.empty()) {
    LOGE("no keyframe\n");
    return false;
  }
  const size_t F = info.face_info_list.size();
  std::vector<FaceInfo> restat;  // faces whose statistics were never computed
  restat.reserve(F);
  std::vector<const FaceInfoPerKeyframe*> best(F, nullptr);
  for (size_t i = 0; i < F; i++) {
    const FaceInfo* fi = &info.face_info_list[i];
    if (fi->visible_keyframes.empty()) continue;
    if (fi->max_area_index < 0) {
      restat.push_back(*fi);
      restat.back().CalcStat();
      fi = &restat.back();
    }
    int idx = fi->max_area_index;
    if (option.criteria == ViewSelectionCriteria::kMinViewingAngle)
      idx = fi->min_viewing_angle_index;
    else if (option.criteria == ViewSelectionCriteria::kMinDistance)
      idx = fi->min_distance_index;
    best[i] = &info.face_info_list[i].visible_keyframes[size_t(idx)];
  }
  switch (option.uv_type) {
    case kGenerateSimpleTile:
      return SimpleTile(keyframes, best, mesh, option);
    case kUseOriginalMeshUv:
      return OriginalUv(keyframes, best, mesh, option);
    case kGenerateSimpleTriangles:
      return SimpleTriangles(keyframes, best, mesh, option);
    case kGenerateSimpleCharts:
      return SimpleCharts(keyframes, best, mesh, option);
    default:
      LOGE("OutputUvType %d is not implemented\n", static_cast<int>(option.uv_type));
      return false;
  }
}

}  // namespace ugu
