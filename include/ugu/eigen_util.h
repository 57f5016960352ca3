// ugu/eigen_util.h (UGU): std::hash for fixed-size Eigen matrices, so a colour
// can key the occurrence maps of VertexInfo (an unordered_map's contents do
// not depend on the hash; only its iteration order does).
#pragma once

#include <cstddef>
#include <functional>

#include "ugu/common.h"

namespace std {

template <typename Scalar, int Rows, int Cols>
struct hash<Eigen::Matrix<Scalar, Rows, Cols>> {
  size_t operator()(const Eigen::Matrix<Scalar, Rows, Cols>& m) const {
    size_t h = 1469598103934665603ull;
    for (int i = 0; i < Rows * Cols; i++) {
      h ^= std::hash<Scalar>()(m.data()[i]);
      h *= 1099511628211ull;
    }
    return h;
  }
};

}  // namespace std
