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
  std::vector<ObjMaterial> materials(1);
  materials[0].name = option.texture_base_name;
  materials[0].diffuse_tex = texture;
  mesh->set_materials(materials);
  mesh->set_material_ids(std::vector<int>(mesh->vertex_indices().size(), 0));
  return true;
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
    default:
      LOGE("OutputUvType %d is not implemented\n", static_cast<int>(option.uv_type));
      return false;
  }
}

}  // namespace ugu
