// Host implementations of the ugu type subset in include/ugu/: logging
// (src/log.cc:21-62), radians/degrees, ConvertTo for RenderW
// (include/ugu/image_proc.h:453-520), Mesh::CalcFaceNormal / CalcNormal
// (src/mesh.cc:297-349) and PinholeCamera (include/ugu/camera.h:257-387).

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "ugu/camera.h"
#include "ugu/image.h"
#include "ugu/log.h"
#include "ugu/mesh.h"

namespace ugu {

namespace {
LogLevel g_level = LogLevel::kVerbose;

void vlog(LogLevel level, const char* format, va_list args) {
  if (level < g_level) return;
  std::vprintf(format, args);
}
}  // namespace

void set_log_level(LogLevel level) { g_level = level; }
LogLevel get_log_level() { return g_level; }

#define UGU_LOG_IMPL(NAME, LEVEL)         \
  void NAME(const char* format, ...) {    \
    va_list args;                         \
    va_start(args, format);               \
    vlog(LEVEL, format, args);            \
    va_end(args);                         \
  }
UGU_LOG_IMPL(LOGD, LogLevel::kDebug)
UGU_LOG_IMPL(LOGI, LogLevel::kInfo)
UGU_LOG_IMPL(LOGW, LogLevel::kWarning)
UGU_LOG_IMPL(LOGE, LogLevel::kError)
#undef UGU_LOG_IMPL

float radians(const float& degrees) { return degrees * static_cast<float>(pi) / 180.0f; }
double radians(const double& degrees) { return degrees * pi / 180.0; }
float degrees(const float& radians) { return radians * 180.0f / static_cast<float>(pi); }
double degrees(const double& radians) { return radians * 180.0 / pi; }

bool ConvertTo(const Image1f& src, Image1w* dst, float scale) {
  if (dst->cols != src.cols || dst->rows != src.rows) *dst = Image1w::zeros(src.rows, src.cols);
  for (int y = 0; y < src.rows; y++)
    for (int x = 0; x < src.cols; x++) {
      const double v = static_cast<double>(scale) * static_cast<double>(src.at<float>(y, x));
      // static_cast<uint16_t>(double) as x86-64 emits it: 32-bit truncation,
      // INT32_MIN for NaN / out of range, then the low 16 bits
      const int32_t iv = (v >= -2147483648.0 && v < 2147483648.0) ? static_cast<int32_t>(v)
                                                                   : INT32_MIN;
      dst->at<uint16_t>(y, x) = static_cast<uint16_t>(static_cast<uint32_t>(iv) & 0xffffu);
    }
  return true;
}

void Mesh::CalcFaceNormal() {
  face_normals_.clear();
  face_normals_.resize(vertex_indices_.size());
  for (size_t i = 0; i < vertex_indices_.size(); i++) {
    const auto& f = vertex_indices_[i];
    Eigen::Vector3f v1 = (vertices_[f[1]] - vertices_[f[0]]).normalized();
    Eigen::Vector3f v2 = (vertices_[f[2]] - vertices_[f[0]]).normalized();
    face_normals_[i] = v1.cross(v2).normalized();
  }
}

void Mesh::CalcNormal() {
  if (vertex_indices_.empty()) return;
  CalcFaceNormal();
  normal_indices_ = vertex_indices_;
  normals_.assign(vertices_.size(), Eigen::Vector3f(0.0f, 0.0f, 0.0f));
  std::vector<int> add_count(vertices_.size(), 0);
  for (size_t i = 0; i < vertex_indices_.size(); i++)
    for (int j = 0; j < 3; j++) {
      const int idx = vertex_indices_[i][j];
      normals_[idx] += face_normals_[i];
      add_count[idx]++;
    }
  for (size_t i = 0; i < vertices_.size(); i++) {
    if (add_count[i] > 0) {
      normals_[i] /= static_cast<float>(add_count[i]);
      normals_[i].normalize();
    } else {
      normals_[i].setZero();
      LOGW("vertex %d has invalid normal\n", static_cast<int>(i));
    }
  }
}

PinholeCamera::PinholeCamera() { set_c2w(Eigen::Affine3d::Identity()); }

PinholeCamera::PinholeCamera(int width, int height, float fov_y_deg) {
  set_size(width, height);
  principal_point_[0] = width_ * 0.5f - 0.5f;
  principal_point_[1] = height_ * 0.5f - 0.5f;
  set_fov_y(fov_y_deg);
  set_c2w(Eigen::Affine3d::Identity());
}

PinholeCamera::PinholeCamera(int width, int height, const Eigen::Affine3d& c2w,
                             float fov_y_deg) {
  set_size(width, height);
  set_c2w(c2w);
  principal_point_[0] = width_ * 0.5f - 0.5f;
  principal_point_[1] = height_ * 0.5f - 0.5f;
  set_fov_y(fov_y_deg);
}

PinholeCamera::PinholeCamera(int width, int height, const Eigen::Affine3d& c2w,
                             const Eigen::Vector2f& principal_point,
                             const Eigen::Vector2f& focal_length)
    : principal_point_(principal_point), focal_length_(focal_length) {
  set_size(width, height);
  set_c2w(c2w);
}

void PinholeCamera::set_size(int width, int height) {
  width_ = width;
  height_ = height;
}

void PinholeCamera::set_c2w(const Eigen::Affine3d& c2w) {
  c2w_ = c2w;
  w2c_ = c2w_.inverse();
}

float PinholeCamera::fov_x() const {
  return degrees(2 * std::atan(width_ * 0.5f / focal_length_[0]));
}
float PinholeCamera::fov_y() const {
  return degrees(2 * std::atan(height_ * 0.5f / focal_length_[1]));
}
void PinholeCamera::set_fov_x(float fov_x_deg) {
  focal_length_[0] = width_ * 0.5f / static_cast<float>(std::tan(radians(fov_x_deg) * 0.5));
  focal_length_[1] = focal_length_[0];
}
void PinholeCamera::set_fov_y(float fov_y_deg) {
  focal_length_[1] = height_ * 0.5f / static_cast<float>(std::tan(radians(fov_y_deg) * 0.5));
  focal_length_[0] = focal_length_[1];
}

void PinholeCamera::Project(const Eigen::Vector3f& camera_p, Eigen::Vector3f* image_p) const {
  (*image_p)[0] = focal_length_[0] / camera_p[2] * camera_p[0] + principal_point_[0];
  (*image_p)[1] = focal_length_[1] / camera_p[2] * camera_p[1] + principal_point_[1];
  (*image_p)[2] = camera_p[2];
}

}  // namespace ugu
