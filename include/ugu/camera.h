// Subset of ugu/camera.h (include/ugu/camera.h:26-154,257-387 in UGU):
// the Camera interface members the renderer uses and PinholeCamera.
// OpenCV axes: z forward, y down, x right.
#pragma once

#include <memory>

#include "ugu/common.h"

namespace ugu {

class Camera {
 public:
  virtual ~Camera() {}
  virtual int width() const = 0;
  virtual int height() const = 0;
  virtual const Eigen::Affine3d& c2w() const = 0;
  virtual const Eigen::Affine3d& w2c() const = 0;
  virtual void set_size(int width, int height) = 0;
  virtual void set_c2w(const Eigen::Affine3d& c2w) = 0;
  virtual void Project(const Eigen::Vector3f& camera_p, Eigen::Vector3f* image_p) const = 0;
};

using CameraPtr = std::shared_ptr<Camera>;

class PinholeCamera : public Camera {
 public:
  PinholeCamera();
  PinholeCamera(int width, int height, float fov_y_deg);
  PinholeCamera(int width, int height, const Eigen::Affine3d& c2w, float fov_y_deg);
  PinholeCamera(int width, int height, const Eigen::Affine3d& c2w,
                const Eigen::Vector2f& principal_point, const Eigen::Vector2f& focal_length);
  ~PinholeCamera() override {}

  int width() const override { return width_; }
  int height() const override { return height_; }
  const Eigen::Affine3d& c2w() const override { return c2w_; }
  const Eigen::Affine3d& w2c() const override { return w2c_; }
  void set_size(int width, int height) override;
  void set_c2w(const Eigen::Affine3d& c2w) override;

  float fov_x() const;
  float fov_y() const;
  void set_fov_x(float fov_x_deg);
  void set_fov_y(float fov_y_deg);
  const Eigen::Vector2f& principal_point() const { return principal_point_; }
  const Eigen::Vector2f& focal_length() const { return focal_length_; }
  void set_principal_point(const Eigen::Vector2f& p) { principal_point_ = p; }
  void set_focal_length(const Eigen::Vector2f& f) { focal_length_ = f; }

  void Project(const Eigen::Vector3f& camera_p, Eigen::Vector3f* image_p) const override;

 protected:
  int width_{-1};
  int height_{-1};
  Eigen::Affine3d c2w_;
  Eigen::Affine3d w2c_;
  Eigen::Vector2f principal_point_{-1.0f, -1.0f};
  Eigen::Vector2f focal_length_{-1.0f, -1.0f};
};

using PinholeCameraPtr = std::shared_ptr<PinholeCamera>;

}  // namespace ugu
