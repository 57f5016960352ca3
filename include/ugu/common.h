// Subset of ugu/common.h (include/ugu/common.h:45-60 in UGU) used by the
// renderer API.
#pragma once

#include "ugu/eigen_min.h"

namespace ugu {

static inline const double pi = 3.14159265358979323846;

// Interpolation method in texture uv space
enum class ColorInterpolation {
  kNn = 0,       // Nearest Neigbor
  kBilinear = 1  // Bilinear interpolation
};

float radians(const float& degrees);
double radians(const double& degrees);
float degrees(const float& radians);
double degrees(const double& radians);

}  // namespace ugu
