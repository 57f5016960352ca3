// Host-side construction of the raytracer's bounding volume hierarchy.
// Plain C++ (no HIP types) so the builder can also be compiled for host-only
// checks.
#pragma once

#include <stdint.h>

#include <vector>

namespace ugu_hip {

struct BvhBuild {
  // 16 floats per inner node: child A (lo.xyz, reference bits, hi.xyz, 0),
  // child B likewise. Node 0 is the root; children follow their parent.
  std::vector<float> nodes;
  std::vector<int32_t> leaves;  // (first triangle record, count) per leaf
  // 12 floats per triangle record: (v0.xyz, face id bits) (v1.xyz, 0) (v2.xyz, 0)
  std::vector<float> tris;
  int max_inner_depth = 0;  // the root has depth 0
  int dropped_faces = 0;    // faces with a non-finite coordinate (never hit)
};

// Builds the hierarchy over F faces given as records of 12 floats
// (x0 y0 z0 x1)(y1 z1 x2 y2)(z2 * * *). Inner nodes are at most
// max_inner_depth deep (leaves take whatever remains at that depth).
// Returns 0, -1 when no face has finite coordinates, -2 when the result fails
// its structure checks (forward inner references, leaf and triangle ranges,
// depth bound).
int build_bvh(const float* faces12, int32_t F, int max_inner_depth, BvhBuild* out);

}  // namespace ugu_hip
