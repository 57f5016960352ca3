// Binned-SAH bounding volume hierarchy for the raytracer kernel.
//
// Top-down: 16 centroid bins along the widest centroid axis, surface area
// heuristic with a traversal cost of one triangle test; ranges of at most 8
// faces become leaves when splitting does not pay. Below depth 32 the builder
// switches to balanced median splits, so the depth stays bounded for any
// input; at max_inner_depth whatever remains is one leaf. The reference's
// naive BVH (include/ugu/accel/bvh_naive.h) only prunes with box tests too, so
// the hierarchy's shape never changes which face is found.
//
// Child boxes are widened outward (see widen_lo/widen_hi) so that the float
// slab test of the kernel cannot reject a box holding a float64 hit.

#include "bvh_build.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace ugu_hip {
namespace {

constexpr int kBins = 16;
constexpr int kSahDepth = 32;
constexpr int kLeafMax = 8;

struct Box {
  float lo[3], hi[3];
  void reset() {
    for (int k = 0; k < 3; k++) {
      lo[k] = std::numeric_limits<float>::infinity();
      hi[k] = -std::numeric_limits<float>::infinity();
    }
  }
  void grow(const Box& b) {
    for (int k = 0; k < 3; k++) {
      lo[k] = std::min(lo[k], b.lo[k]);
      hi[k] = std::max(hi[k], b.hi[k]);
    }
  }
  void grow(const float p[3]) {
    for (int k = 0; k < 3; k++) {
      lo[k] = std::min(lo[k], p[k]);
      hi[k] = std::max(hi[k], p[k]);
    }
  }
  double area() const {
    double e[3];
    for (int k = 0; k < 3; k++) e[k] = std::max(0.0, double(hi[k]) - double(lo[k]));
    return e[0] * e[1] + e[1] * e[2] + e[2] * e[0];
  }
};

// Widening: |x| 2^-20 + extent 2^-16 + 1e-30 outward, rounded outward to float.
// Covers the float64 rounding of hits at the box surface and keeps a ray that
// runs exactly along a widened face (d = 0 on that axis) away from geometry.
float widen_lo(float x, double ext) {
  const double d = double(x) - (std::fabs(double(x)) * 0x1p-20 + ext * 0x1p-16 + 1e-30);
  float f = static_cast<float>(d);
  if (double(f) > d) f = std::nextafter(f, -std::numeric_limits<float>::infinity());
  return f;
}
float widen_hi(float x, double ext) {
  const double d = double(x) + (std::fabs(double(x)) * 0x1p-20 + ext * 0x1p-16 + 1e-30);
  float f = static_cast<float>(d);
  if (double(f) < d) f = std::nextafter(f, std::numeric_limits<float>::infinity());
  return f;
}

float bits_float(int32_t i) {
  float f;
  std::memcpy(&f, &i, 4);
  return f;
}

struct Builder {
  const float* rec;
  std::vector<Box> fbox;
  std::vector<float> cen;  // 3 per face
  std::vector<int32_t> idx;
  int max_depth;
  BvhBuild* out;

  struct Res {
    int32_t ref;
    Box box;
  };

  int32_t make_leaf(int b, int e) {
    const int32_t leaf = static_cast<int32_t>(out->leaves.size() / 2);
    out->leaves.push_back(static_cast<int32_t>(out->tris.size() / 12));
    out->leaves.push_back(e - b);
    for (int i = b; i < e; i++) {
      const int32_t f = idx[size_t(i)];
      const float* r = rec + size_t(f) * 12;
      const float t[12] = {r[0], r[1], r[2], bits_float(f), r[3], r[4], r[5], 0.0f,
                           r[6], r[7], r[8], 0.0f};
      out->tris.insert(out->tris.end(), t, t + 12);
    }
    return ~leaf;
  }

  // Partition point of [b, e) along the widest centroid axis.
  int split(int b, int e, const Box& cb, int depth, bool* leaf) {
    const int n = e - b;
    int axis = 0;
    double ext = -1.0;
    for (int k = 0; k < 3; k++) {
      const double x = double(cb.hi[k]) - double(cb.lo[k]);
      if (x > ext) {
        ext = x;
        axis = k;
      }
    }
    *leaf = false;
    if (!(ext > 0.0)) {
      // all centroids coincide: no plane separates them
      if (n <= kLeafMax) {
        *leaf = true;
        return b;
      }
      return b + n / 2;
    }
    if (depth < kSahDepth) {
      int cnt[kBins] = {};
      Box bb[kBins];
      for (auto& x : bb) x.reset();
      const double lo = cb.lo[axis], scale = kBins / ext;
      auto bin_of = [&](int32_t f) {
        int k = static_cast<int>((double(cen[size_t(f) * 3 + axis]) - lo) * scale);
        return std::min(kBins - 1, std::max(0, k));
      };
      for (int i = b; i < e; i++) {
        const int32_t f = idx[size_t(i)];
        const int k = bin_of(f);
        cnt[k]++;
        bb[k].grow(fbox[size_t(f)]);
      }
      double right_area[kBins];
      int right_cnt[kBins];
      Box acc;
      acc.reset();
      int c = 0;
      for (int k = kBins - 1; k > 0; k--) {
        acc.grow(bb[k]);
        c += cnt[k];
        right_area[k] = acc.area();
        right_cnt[k] = c;
      }
      acc.reset();
      c = 0;
      double best = std::numeric_limits<double>::infinity();
      int best_k = -1;
      for (int k = 1; k < kBins; k++) {
        acc.grow(bb[k - 1]);
        c += cnt[k - 1];
        if (c == 0 || right_cnt[k] == 0) continue;
        const double cost = acc.area() * c + right_area[k] * right_cnt[k];
        if (cost < best) {
          best = cost;
          best_k = k;
        }
      }
      Box all = acc;
      all.grow(bb[kBins - 1]);
      const double pa = all.area();
      if (best_k > 0) {
        // cost in triangle tests: 1 for the node + the expected child tests
        const double split_cost = 1.0 + (pa > 0.0 ? best / pa : double(n));
        if (n <= kLeafMax && split_cost >= double(n)) {
          *leaf = true;
          return b;
        }
        int32_t* first = idx.data() + b;
        int32_t* mid = std::partition(first, idx.data() + e,
                                      [&](int32_t f) { return bin_of(f) < best_k; });
        const int m = static_cast<int>(mid - idx.data());
        if (m > b && m < e) return m;
      }
    }
    if (n <= kLeafMax && depth >= kSahDepth) {
      *leaf = true;
      return b;
    }
    const int m = b + n / 2;
    std::nth_element(idx.data() + b, idx.data() + m, idx.data() + e, [&](int32_t p, int32_t q) {
      const float a = cen[size_t(p) * 3 + axis], c2 = cen[size_t(q) * 3 + axis];
      return a < c2 || (a == c2 && p < q);
    });
    return m;
  }

  Res build(int b, int e, int depth) {
    Res r;
    r.box.reset();
    Box cb;
    cb.reset();
    for (int i = b; i < e; i++) {
      const int32_t f = idx[size_t(i)];
      r.box.grow(fbox[size_t(f)]);
      cb.grow(&cen[size_t(f) * 3]);
    }
    const int n = e - b;
    if (n <= 2 || depth > max_depth) {
      r.ref = make_leaf(b, e);
      return r;
    }
    bool leaf = false;
    const int m = split(b, e, cb, depth, &leaf);
    if (leaf) {
      r.ref = make_leaf(b, e);
      return r;
    }
    const int32_t node = static_cast<int32_t>(out->nodes.size() / 16);
    out->nodes.resize(out->nodes.size() + 16);
    out->max_inner_depth = std::max(out->max_inner_depth, depth);
    const Res a = build(b, m, depth + 1);
    const Res c = build(m, e, depth + 1);
    write_child(node, 0, a);
    write_child(node, 1, c);
    r.ref = node;
    return r;
  }

  void write_child(int32_t node, int slot, const Res& c) {
    float* p = out->nodes.data() + size_t(node) * 16 + size_t(slot) * 8;
    double ext = 0.0;
    for (int k = 0; k < 3; k++) ext = std::max(ext, double(c.box.hi[k]) - double(c.box.lo[k]));
    for (int k = 0; k < 3; k++) {
      p[k] = widen_lo(c.box.lo[k], ext);
      p[4 + k] = widen_hi(c.box.hi[k], ext);
    }
    p[3] = bits_float(c.ref);
    p[7] = 0.0f;
  }
};

int32_t float_bits(float f) {
  int32_t i;
  std::memcpy(&i, &f, 4);
  return i;
}

// Structure checks the traversal relies on: inner references point forward
// (so a walk ends), leaf references and triangle runs are in range, every
// depth fits the stack bound the caller asked for.
bool valid(const BvhBuild& b) {
  const int64_t nn = static_cast<int64_t>(b.nodes.size() / 16);
  const int64_t nl = static_cast<int64_t>(b.leaves.size() / 2);
  const int64_t nt = static_cast<int64_t>(b.tris.size() / 12);
  if (nn < 1) return false;
  for (int64_t i = 0; i < nn; i++) {
    for (int slot = 0; slot < 2; slot++) {
      const int32_t r = float_bits(b.nodes[size_t(i) * 16 + size_t(slot) * 8 + 3]);
      if (r == INT32_MIN) {
        if (slot == 0) return false;
        continue;
      }
      if (r >= 0) {
        if (r <= i || r >= nn) return false;
      } else if (int64_t(~r) >= nl) {
        return false;
      }
    }
  }
  for (int64_t l = 0; l < nl; l++) {
    const int64_t first = b.leaves[size_t(l) * 2], cnt = b.leaves[size_t(l) * 2 + 1];
    if (first < 0 || cnt < 1 || first + cnt > nt) return false;
  }
  return true;
}

}  // namespace

int build_bvh(const float* faces12, int32_t F, int max_inner_depth, BvhBuild* out) {
  *out = BvhBuild{};
  Builder bl;
  bl.rec = faces12;
  bl.max_depth = max_inner_depth;
  bl.out = out;
  bl.fbox.resize(size_t(F));
  bl.cen.resize(size_t(F) * 3);
  bl.idx.reserve(size_t(F));
  for (int32_t f = 0; f < F; f++) {
    const float* r = faces12 + size_t(f) * 12;
    bool finite = true;
    for (int k = 0; k < 9; k++) finite = finite && std::isfinite(r[k]);
    if (!finite) {
      out->dropped_faces++;
      continue;
    }
    Box& b = bl.fbox[size_t(f)];
    b.reset();
    b.grow(r);
    b.grow(r + 3);
    b.grow(r + 6);
    for (int k = 0; k < 3; k++)
      bl.cen[size_t(f) * 3 + k] = 0.5f * b.lo[k] + 0.5f * b.hi[k];
    bl.idx.push_back(f);
  }
  const int n = static_cast<int>(bl.idx.size());
  if (n == 0) return -1;
  const Builder::Res root = bl.build(0, n, 0);
  if (root.ref < 0) {
    // a single leaf: wrap it in a root whose second child is empty
    out->nodes.assign(16, 0.0f);
    bl.write_child(0, 0, root);
    float* p = out->nodes.data() + 8;
    p[3] = bits_float(INT32_MIN);
    for (int k = 0; k < 3; k++) {
      p[k] = 1.0f;
      p[4 + k] = -1.0f;
    }
  }
  if (out->max_inner_depth > max_inner_depth || !valid(*out)) return -2;
  return 0;
}

}  // namespace ugu_hip
