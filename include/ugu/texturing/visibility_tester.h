// ugu/texturing/visibility_tester.h (UGU): the visibility tester and its
// records, with the per-vertex test on the GPU (ugu_amd/cpp/texturing.cc over
// ugu_hip_visibility_test). Same names, members and meaning as UGU.
#pragma once

#include <array>
#include <functional>
#include <iostream>
#include <limits>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "ugu/camera.h"
#include "ugu/common.h"
#include "ugu/eigen_util.h"
#include "ugu/image.h"
#include "ugu/mesh.h"

namespace ugu {

// ViewSelectionCriteria (visibility_tester.h:22-35)
enum class ViewSelectionCriteria {
  kMinViewingAngle = 0,
  kMinDistance = 1,
  kMeanViewingAngle = 2,
  kMedianViewingAngle = 3,
  kMeanDistance = 4,
  kMedianDistance = 5,
  kMaxArea = 6,
  kMinIntensity = 7,
  kMedianIntensity = 8,
  kMode = 9,  // meaningful only for discrete keyframe colours (class ids)
  kModeViewingAngle = 10,
  kCustom = 11
};

// VisibilityTesterOption (visibility_tester.h:38-57)
struct VisibilityTesterOption {
  float backface_culling_angle_rad_th{radians(90.0f)};
  bool use_mask{true};
  bool use_depth{true};
  ColorInterpolation interp{ColorInterpolation::kBilinear};
  bool calc_stat_vertex_info{true};
  bool collect_face_info{true};
  bool calc_stat_face_info{true};
  float max_viewing_angle{radians(90.0f)};
  float max_distance_from_camera_to_vertex{std::numeric_limits<float>::max()};
  float max_depth_difference{std::numeric_limits<float>::max()};
  float eps = 0.1f;
  Eigen::Vector3f invalid_color{0.0f, 0.0f, 0.0f};
  void CopyTo(VisibilityTesterOption* dst) const { *dst = *this; }
};

struct VertexInfoPerKeyframe {
  int kf_id{-1};
  Eigen::Vector3f color{0.0f, 0.0f, 0.0f};
  float intensity = -1.f;
  Eigen::Vector2f projected_pos{0.0f, 0.0f};
  float viewing_angle{999.9f};
  float distance{-999.9f};
};

struct VertexInfo {
  std::vector<VertexInfoPerKeyframe> visible_keyframes;
  Eigen::Vector3f mean_color, median_color;
  int min_viewing_angle_index{-1}, min_viewing_angle_id{-1};
  Eigen::Vector3f min_viewing_angle_color;
  Eigen::Vector3f mean_viewing_angle_color, median_viewing_angle_color;
  int min_distance_index{-1}, min_distance_id{-1};
  Eigen::Vector3f min_distance_color;
  Eigen::Vector3f mean_distance_color, median_distance_color;
  float min_intensity = std::numeric_limits<float>::max();
  float median_intensity = std::numeric_limits<float>::max();
  Eigen::Vector3f median_intensity_color, min_intensity_color;
  float mean_viewing_angle{-1.0f}, median_viewing_angle{-1.0f};
  float mean_distance{-1.0f}, median_distance{-1.0f};
  // the most frequent colour (plain and weighted by inverse viewing angle)
  Eigen::Vector3f mode;
  int mode_frequency{-1};
  Eigen::Vector3f mode_viewing_angle_color;
  float mode_frequency_viewing_angle_color{-1.f};
  std::unordered_map<Eigen::Vector3f, int> occurrence;
  std::unordered_map<Eigen::Vector3f, float> occurrence_viewing_angle;
  // free for the custom statistics callback
  std::vector<std::vector<Eigen::Vector3f>> customs;
  Eigen::Vector3f custom_best;

  void Update(const VertexInfoPerKeyframe& info) { visible_keyframes.push_back(info); }
  void CalcStat(std::function<void(VertexInfo&)> vert_custom_func = nullptr);
  int VisibleFrom(int kf_id) const;
};

std::ostream& operator<<(std::ostream& os, const VertexInfoPerKeyframe& vi);
std::ostream& operator<<(std::ostream& os, const VertexInfo& vi);

struct FaceInfoPerKeyframe {
  int kf_id{-1};
  int face_id{-1};
  float area{-1.f};
  float viewing_angle{-1.f};
  float distance{std::numeric_limits<float>::max()};
  std::array<Eigen::Vector2f, 3> projected_tri;
};

struct FaceInfo {
  std::vector<FaceInfoPerKeyframe> visible_keyframes;
  int max_area_index{-1}, max_area_id{-1};
  float max_area{-1.0f};
  int min_viewing_angle_index{-1}, min_viewing_angle_id{-1};
  float min_viewing_angle = std::numeric_limits<float>::max();
  int min_distance_index{-1}, min_distance_id{-1};
  float min_distance = std::numeric_limits<float>::max();

  void Update(const FaceInfoPerKeyframe& info) { visible_keyframes.push_back(info); }
  void CalcStat();
};

struct VisibilityInfo {
  bool has_vertex_stat{false};
  bool has_face_stat{false};
  std::vector<VertexInfo> vertex_info_list;
  std::vector<FaceInfo> face_info_list;

  VisibilityInfo() {}
  explicit VisibilityInfo(const Mesh& mesh)
      : vertex_info_list(mesh.vertices().size()), face_info_list(mesh.vertex_indices().size()) {}
  void CalcStatVertexInfo(std::function<void(VertexInfo&)> vert_custom_func = nullptr);
  void CalcStatFaceInfo();
  std::string SerializeAsJson() const;
  void Update(int vertex_id, const VertexInfoPerKeyframe& info) {
    vertex_info_list[size_t(vertex_id)].Update(info);
  }
  void Update(int face_id, const FaceInfoPerKeyframe& info) {
    face_info_list[size_t(face_id)].Update(info);
  }
};

struct Keyframe;
using KeyframePtr = std::shared_ptr<Keyframe>;

struct Keyframe {
  int id = -1;
  std::string color_path;
  Image3b color;
  std::string depth_path;
  Image1f depth;
  std::string mask_path;
  Image1b mask;
  std::shared_ptr<const Camera> camera;
  static KeyframePtr Create() { return std::make_shared<Keyframe>(); }
};

class VisibilityTester {
  class Impl;
  std::unique_ptr<Impl> pimpl_;

 public:
  VisibilityTester();
  ~VisibilityTester();
  explicit VisibilityTester(const VisibilityTesterOption& option);
  void set_option(const VisibilityTesterOption& option);
  void set_data(const std::vector<Eigen::Vector3f>& vertices,
                const std::vector<Eigen::Vector3f>& normals,
                const std::vector<Eigen::Vector3i>& indices);
  bool PrepareData();
  void set_keyframe(std::shared_ptr<Keyframe> keyframe);
  // facewise=true: each corner tested 1e-3 towards its face centre; a
  // vertex's record comes from the first face (lowest id) that passes it
  bool Test(VisibilityInfo* info, bool facewise = true,
            std::function<void(VertexInfo&)> vert_custom_func = nullptr) const;
  bool Test(std::vector<std::shared_ptr<Keyframe>> keyframes, VisibilityInfo* info,
            bool facewise = true, std::function<void(VertexInfo&)> vert_custom_func = nullptr);
};

}  // namespace ugu
