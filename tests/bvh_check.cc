// Host-side check of the raytracer's hierarchy (built by tests/test_bvh_cpu.py
// with g++): for seeded triangle soups and rays, the traversal the kernels run
// (widened child boxes, float slab test with 2^-20 slack, nearer child first,
// pruning at the current closest t) must find the same closest t as testing
// every face with the same float64 Moller-Trumbore routine. Adversarial rays:
// axis-parallel, through triangle vertices and edges, grazing, origins on box
// faces, far-away and tiny scenes.
//
// usage: bvh_check <seed> <n_scenes>; prints the number of rays and
// mismatches, exits 1 on any mismatch or structure failure.

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../ugu_amd/csrc/bvh_build.h"
#include "../ugu_amd/csrc/rt_math.h"

namespace {

struct F4 {
  float x, y, z, w;
};

int32_t bits(float f) {
  int32_t i;
  std::memcpy(&i, &f, 4);
  return i;
}

float traverse(const ugu_hip::BvhBuild& b, float ox, float oy, float oz, float r0, float r1,
               float r2, int* max_sp) {
  const float ix = 1.0f / r0, iy = 1.0f / r1, iz = 1.0f / r2;
  float bt = __builtin_huge_valf();
  int stack[64];
  int sp = 0, ref = 0;
  for (;;) {
    if (ref >= 0) {
      const float* nd = &b.nodes[size_t(ref) * 16];
      const F4 alo{nd[0], nd[1], nd[2], 0}, ahi{nd[4], nd[5], nd[6], 0};
      const F4 blo{nd[8], nd[9], nd[10], 0}, bhi{nd[12], nd[13], nd[14], 0};
      const int ra = bits(nd[3]), rb = bits(nd[11]);
      const float tmax = ugu_hip::box_tmax(bt);
      const float ta = ugu_hip::box_entry(alo, ahi, ox, oy, oz, ix, iy, iz, tmax);
      const float tb = rb == INT32_MIN ? __builtin_huge_valf()
                                       : ugu_hip::box_entry(blo, bhi, ox, oy, oz, ix, iy, iz, tmax);
      const bool ha = ta != __builtin_huge_valf(), hb = tb != __builtin_huge_valf();
      if (ha && hb) {
        if (sp >= 64) return -1.0f;  // would overflow the kernels' stack
        stack[sp++] = ta <= tb ? rb : ra;
        if (sp > *max_sp) *max_sp = sp;
        ref = ta <= tb ? ra : rb;
        continue;
      }
      if (ha || hb) {
        ref = ha ? ra : rb;
        continue;
      }
    } else {
      const int32_t first = b.leaves[size_t(~ref) * 2], cnt = b.leaves[size_t(~ref) * 2 + 1];
      for (int k = 0; k < cnt; k++) {
        const float* t = &b.tris[size_t(first + k) * 12];
        const F4 p0{t[0], t[1], t[2], 0}, p1{t[4], t[5], t[6], 0}, p2{t[8], t[9], t[10], 0};
        float tt, u, v;
        if (ugu_hip::mt_hit<F4>(ox, oy, oz, r0, r1, r2, p0, p1, p2, tt, u, v) && tt < bt) bt = tt;
      }
    }
    if (sp == 0) break;
    ref = stack[--sp];
  }
  return bt;
}

float brute(const std::vector<float>& rec, int F, float ox, float oy, float oz, float r0,
            float r1, float r2) {
  float bt = __builtin_huge_valf();
  for (int f = 0; f < F; f++) {
    const float* t = &rec[size_t(f) * 12];
    bool finite = true;
    for (int k = 0; k < 9; k++) finite = finite && std::isfinite(t[k]);
    if (!finite) continue;
    const F4 p0{t[0], t[1], t[2], 0}, p1{t[3], t[4], t[5], 0}, p2{t[6], t[7], t[8], 0};
    float tt, u, v;
    if (ugu_hip::mt_hit<F4>(ox, oy, oz, r0, r1, r2, p0, p1, p2, tt, u, v) && tt < bt) bt = tt;
  }
  return bt;
}

}  // namespace

int main(int argc, char** argv) {
  const unsigned seed = argc > 1 ? unsigned(std::atoi(argv[1])) : 1;
  const int scenes = argc > 2 ? std::atoi(argv[2]) : 20;
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> U(-1.0f, 1.0f);
  long rays = 0, bad = 0;
  int max_sp = 0;
  for (int s = 0; s < scenes; s++) {
    // scene scale and offset: unit, tiny, huge, far from the origin
    const float scale = (s % 4 == 0) ? 1.0f : (s % 4 == 1 ? 1e-3f : (s % 4 == 2 ? 1e3f : 1.0f));
    const float off = (s % 4 == 3) ? 1e4f : 0.0f;
    const int F = 50 + int(rng() % 400);
    std::vector<float> rec(size_t(F) * 12, 0.0f);
    for (int f = 0; f < F; f++) {
      float c[3] = {U(rng), U(rng), U(rng)};
      const float sz = (f % 5 == 0) ? 0.5f : 0.08f;
      for (int k = 0; k < 3; k++)
        for (int a = 0; a < 3; a++) {
          float v = c[a] + sz * U(rng);
          if (f % 7 == 0 && a == 2) v = c[2];  // axis-aligned (constant z) faces
          rec[size_t(f) * 12 + k * 3 + a] = off + scale * v;
        }
    }
    ugu_hip::BvhBuild b;
    if (ugu_hip::build_bvh(rec.data(), F, 63, &b) != 0) {
      std::printf("build failed in scene %d\n", s);
      return 1;
    }
    for (int r = 0; r < 3000; r++) {
      float o[3], d[3];
      const int kind = r % 6;
      for (int a = 0; a < 3; a++) o[a] = off + scale * 2.5f * U(rng);
      const float* t = &rec[size_t(rng() % F) * 12];
      float target[3];
      if (kind == 1) {  // through a vertex
        for (int a = 0; a < 3; a++) target[a] = t[a];
      } else if (kind == 2) {  // through an edge midpoint
        for (int a = 0; a < 3; a++) target[a] = 0.5f * t[a] + 0.5f * t[3 + a];
      } else {
        for (int a = 0; a < 3; a++) target[a] = (t[a] + t[3 + a] + t[6 + a]) / 3.0f;
      }
      for (int a = 0; a < 3; a++) d[a] = target[a] - o[a];
      if (kind == 3) {  // axis-parallel
        const int ax = int(rng() % 3);
        for (int a = 0; a < 3; a++) d[a] = a == ax ? (U(rng) > 0 ? 1.0f : -1.0f) : 0.0f;
        for (int a = 0; a < 3; a++)
          if (a != ax) o[a] = target[a];
      }
      if (kind == 4) {  // grazing: in the plane of a constant-z face, origin on its height
        o[2] = t[2];
        d[2] = 0.0f;
      }
      if (kind == 5) {  // origin on a widened child box face of the root
        const int ax = int(rng() % 3);
        o[ax] = b.nodes[size_t(ax)];
      }
      const float n = std::sqrt(d[0] * d[0] + (d[1] * d[1] + d[2] * d[2]));
      if (!(n > 0.0f)) continue;
      for (int a = 0; a < 3; a++) d[a] /= n;
      const float tb = brute(rec, F, o[0], o[1], o[2], d[0], d[1], d[2]);
      const float tt = traverse(b, o[0], o[1], o[2], d[0], d[1], d[2], &max_sp);
      rays++;
      if (bits(tb) != bits(tt)) {
        if (bad < 5)
          std::printf("mismatch scene %d ray %d kind %d: brute %.9g bvh %.9g\n", s, r, kind, tb,
                      tt);
        bad++;
      }
    }
  }
  std::printf("rays %ld mismatches %ld max_stack %d\n", rays, bad, max_sp);
  return bad == 0 ? 0 : 1;
}
