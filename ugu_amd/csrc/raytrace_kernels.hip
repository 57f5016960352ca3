// HIP kernel of the gfx950 raytracer: Raytracer::Impl::Render
// (src/renderer/cpu/raytracer.cc:117-225) for a batch of views.
//
// One wave per 8x8 pixel tile, one lane per pixel. The wave walks the BVH
// (built on the host, nodes hold both children's boxes) as one packet of 64
// rays, nearer child first, and every lane tests the triangles of every leaf
// reached with the reference's float64 Moller-Trumbore routine
// (src/util/geom_util.cc:60-115, epsilon double(1e-8f) of bvh_naive.h:121,
// standard barycentric conversion).
// The closest hit is the smallest float t; equal t resolve to the lowest face
// id. The box test is float32 with widened boxes and a relative slack on the
// ray interval, so it only prunes boxes that cannot hold a hit the float64 test
// would accept. Then the pixel is culled or shaded as in raytracer.cc:160-218.

#include "device_common.h"
#include "rt_math.h"

#pragma clang fp contract(off)

namespace ugu_hip {

// Four waves per workgroup, each on its own tile (no cross-wave sharing).
constexpr int kRtWaves = 4;

// One wave traces the 64 rays of an 8x8 pixel tile through the hierarchy
// together: the node and leaf indices are wave-uniform (fetched with scalar
// loads), every lane tests its own ray against the node's two child boxes,
// and a child is entered when any lane's ray may reach it (nearer child first
// by a vote of the lanes). The stack of deferred children is one LDS word per
// level for the whole wave. Each lane keeps its own closest hit, and the
// closest hit does not depend on the visiting order, so the result is the
// per-ray one; the traversal only decides which tests are made.

__global__ __launch_bounds__(64 * kRtWaves) void k_raytrace(const DevView* __restrict__ views,
                                                            const BvhNode* __restrict__ nodes,
                                                            const int2* __restrict__ leaves,
                                                            const float4* __restrict__ tris,
                                                            int n_views, int total_tiles,
                                                            int chunk, DevMesh m,
                                                            DevOptions o) {
  __shared__ int32_t stacks[kRtWaves][kBvhStack];
  const int wave = threadIdx.x >> 6;
  int32_t* stack = stacks[wave];
  // Tiles of all views are numbered consecutively (views[i].tile_base).
  // Workgroups go round-robin over the 8 XCDs; each XCD here takes runs of
  // `chunk` consecutive workgroups (neighbouring tiles share its L2), and the
  // runs alternate over the XCDs. The grid is a multiple of 8 chunk
  // workgroups, which makes the mapping a permutation.
  const int i = blockIdx.x;
  const int k = i >> 3;
  const int lb = (((k / chunk) << 3) + (i & 7)) * chunk + (k % chunk);
  const int g = lb * kRtWaves + wave;
  if (g >= total_tiles) return;  // uniform over the wave
  int lo = 0, hi = n_views - 1;  // last view with tile_base <= g
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (views[mid].tile_base <= g) lo = mid; else hi = mid - 1;
  }
  const DevView& v = views[lo];
  const int tiles_x = (v.W + 7) >> 3;
  const int tile = g - v.tile_base;
  const int lane = threadIdx.x & 63;
  const int x = (tile % tiles_x) * 8 + (lane & 7);
  const int y = (tile / tiles_x) * 8 + (lane >> 3);
  const bool act = x < v.W && y < v.H;

  // ray_w / org_ray_w tables (camera.h:482-513)
  float r0, r1, r2;
  ray_w_at(v, static_cast<float>(act ? x : 0), static_cast<float>(act ? y : 0), r0, r1, r2);
  const float ox = v.c2w_t[0], oy = v.c2w_t[1], oz = v.c2w_t[2];
  const float ix = 1.0f / r0, iy = 1.0f / r1, iz = 1.0f / r2;
  const double dox = ox, doy = oy, doz = oz, ddx = r0, ddy = r1, ddz = r2;
  const float kInf = __builtin_huge_valf();

  float bt = kInf, bu = 0.0f, bv = 0.0f;
  float tmax = kInf;  // box_tmax(bt), updated with bt
  int bf = -1;
  int sp = 0;
  int ref = 0;  // the root is inner node 0
  // Inner references point to larger node indices and every node is entered
  // at most once, so the walk ends; the builder bounds the depth, and with it
  // the stack, by kBvhStack.
  for (;;) {
    if (ref >= 0) {
      const BvhNode nd = nodes[ref];
      const int ra = __float_as_int(nd.a_lo.w), rb = __float_as_int(nd.b_lo.w);
      // both tests on every lane (no branch between the node's loads), then
      // inactive lanes and the empty child are masked out
      const float ea = box_entry(nd.a_lo, nd.a_hi, ox, oy, oz, ix, iy, iz, tmax);
      const float eb = box_entry(nd.b_lo, nd.b_hi, ox, oy, oz, ix, iy, iz, tmax);
      const float ta = act ? ea : kInf;
      const float tb = (act && rb != kBvhEmpty) ? eb : kInf;
      const bool ha = __any(ta != kInf), hb = __any(tb != kInf);
      if (ha && hb) {
        const int va = __popcll(__ballot(ta < tb)), vb = __popcll(__ballot(tb < ta));
        const bool a_first = va >= vb;
        stack[sp] = a_first ? rb : ra;
        sp++;
        ref = a_first ? ra : rb;
        continue;
      }
      if (ha || hb) {
        ref = ha ? ra : rb;
        continue;
      }
    } else {
      const int2 lf = leaves[~ref];
      for (int k = 0; k < lf.y; k++) {
        const float4* tr = tris + int64_t(lf.x + k) * 3;
        const float4 p0 = tr[0], p1 = tr[1], p2 = tr[2];
        float t, u, w;
        if (act && mt_hit(dox, doy, doz, ddx, ddy, ddz, p0, p1, p2, t, u, w)) {
          const int f = __float_as_int(p0.w);
          if (t < bt || (t == bt && f < bf)) {
            bt = t;
            bu = u;
            bv = w;
            bf = f;
            tmax = box_tmax(bt);
          }
        }
      }
    }
    if (sp == 0) break;
    sp--;
    ref = __builtin_amdgcn_readfirstlane(stack[sp]);
  }
  if (!act) return;
  const int64_t pi = int64_t(y) * v.W + x;

  bool hit = bf >= 0;
  // back-face culling: face normal . ray_w > 0 -> background (raytracer.cc:166-172)
  if (hit && o.culling && backface_of(m.fnormals, bf, r0, r1, r2)) hit = false;

  if (!hit) {
    if (v.color) {
      v.color[pi * 3 + 0] = 0;
      v.color[pi * 3 + 1] = 0;
      v.color[pi * 3 + 2] = 0;
    }
    if (v.depth) v.depth[pi] = 0.0f;
    if (v.normal) {
      v.normal[pi * 3 + 0] = 0.0f;
      v.normal[pi * 3 + 1] = 0.0f;
      v.normal[pi * 3 + 2] = 0.0f;
    }
    if (v.mask) v.mask[pi] = 0;
    if (v.face_id) v.face_id[pi] = -1;
    return;
  }
  if (v.face_id) v.face_id[pi] = bf;
  if (v.mask) v.mask[pi] = 255;
  if (v.depth) {
    // hit_pos_c = w2c_R * (org + ray * t) + w2c_t; depth = z * depth_scale
    const float h0 = ox + r0 * bt, h1 = oy + r1 * bt, h2 = oz + r2 * bt;
    const float zc = sum3(v.w2c_R[6] * h0, v.w2c_R[7] * h1, v.w2c_R[8] * h2) + v.w2c_t[2];
    v.depth[pi] = zc * o.depth_scale;
  }
  float n0, n1, n2;
  FaceShade fs;
  face_shade(m, bf, o.shading_normal != 0, v.color && o.diffuse_color == 1, fs);
  if (o.shading_normal == 0) {
    const float* fn = m.fnormals + int64_t(bf) * 3;
    n0 = fn[0];
    n1 = fn[1];
    n2 = fn[2];
  } else {
    // (1 - u - v) n[i0] + u n[i1] + v n[i2] with the raytracer's (u, v)
    const float* na = fs.na;
    const float* nb = fs.nb;
    const float* nc = fs.nc;
    const float w = 1.0f - bu - bv;
    n0 = w * na[0] + bu * nb[0] + bv * nc[0];
    n1 = w * na[1] + bu * nb[1] + bv * nc[1];
    n2 = w * na[2] + bu * nb[2] + bv * nc[2];
  }
  if (v.normal) {
    v.normal[pi * 3 + 0] = sum3(v.w2c_R[0] * n0, v.w2c_R[1] * n1, v.w2c_R[2] * n2);
    v.normal[pi * 3 + 1] = sum3(v.w2c_R[3] * n0, v.w2c_R[4] * n1, v.w2c_R[5] * n2);
    v.normal[pi * 3 + 2] = sum3(v.w2c_R[6] * n0, v.w2c_R[7] * n1, v.w2c_R[8] * n2);
  }
  if (v.color) {
    uint8_t c[3];
    shade_u8(m, o, bf, fs, bu, bv, r0, r1, r2, n0, n1, n2, c);
    v.color[pi * 3 + 0] = c[0];
    v.color[pi * 3 + 1] = c[1];
    v.color[pi * 3 + 2] = c[2];
  }
}

// ---------------------------------------------------------------------------
// Visibility test: VisibilityTester::TestVertex (visibility_tester.cc:390-530)
// for one item per lane. Items are scattered over the image, so every lane
// walks the hierarchy on its own (stack of kBvhStack entries per lane in LDS).

struct DevVisItem {
  int32_t visible;
  float px, py, angle, dist, c0, c1, c2, intensity;
};

// closest float t over all faces (+inf: no hit), per-lane traversal
UGU_DEV float closest_t(const BvhNode* __restrict__ nodes, const int2* __restrict__ leaves,
                        const float4* __restrict__ tris, int32_t* stack, int lane, float ox,
                        float oy, float oz, float r0, float r1, float r2) {
  const float ix = 1.0f / r0, iy = 1.0f / r1, iz = 1.0f / r2;
  const double dox = ox, doy = oy, doz = oz;
  float bt = __builtin_huge_valf();
  float tmax = bt;  // box_tmax(bt), updated with bt
  int sp = 0, ref = 0;
  for (;;) {
    if (ref >= 0) {
      const BvhNode nd = nodes[ref];
      const int ra = __float_as_int(nd.a_lo.w), rb = __float_as_int(nd.b_lo.w);
      const float ta = box_entry(nd.a_lo, nd.a_hi, ox, oy, oz, ix, iy, iz, tmax);
      const float tb = rb == kBvhEmpty ? __builtin_huge_valf()
                                       : box_entry(nd.b_lo, nd.b_hi, ox, oy, oz, ix, iy, iz, tmax);
      const bool ha = ta != __builtin_huge_valf(), hb = tb != __builtin_huge_valf();
      if (ha && hb) {
        const bool a_first = ta <= tb;
        stack[sp * 64 + lane] = a_first ? rb : ra;
        sp++;
        ref = a_first ? ra : rb;
        continue;
      }
      if (ha || hb) {
        ref = ha ? ra : rb;
        continue;
      }
    } else {
      const int2 lf = leaves[~ref];
      for (int k = 0; k < lf.y; k++) {
        const float4* tr = tris + int64_t(lf.x + k) * 3;
        float t, u, w;
        if (mt_hit(dox, doy, doz, r0, r1, r2, tr[0], tr[1], tr[2], t, u, w) && t < bt) {
          bt = t;
          tmax = box_tmax(bt);
        }
      }
    }
    if (sp == 0) break;
    sp--;
    ref = stack[sp * 64 + lane];
  }
  return bt;
}

__global__ __launch_bounds__(64) void k_visibility(DevView v, const uint8_t* __restrict__ color,
                                                   const float* __restrict__ depth,
                                                   const uint8_t* __restrict__ mask,
                                                   DevVisOpt o, int n,
                                                   const float* __restrict__ pos,
                                                   const float* __restrict__ vert,
                                                   const float* __restrict__ nrm,
                                                   const BvhNode* __restrict__ nodes,
                                                   const int2* __restrict__ leaves,
                                                   const float4* __restrict__ tris,
                                                   DevVisItem* __restrict__ out) {
  __shared__ int32_t stack[kBvhStack * 64];
  const int lane = threadIdx.x;
  const int i = blockIdx.x * 64 + lane;
  if (i >= n) return;
  DevVisItem r;
  r.visible = 0;
  r.px = r.py = r.c0 = r.c1 = r.c2 = 0.0f;
  r.angle = 999.9f;
  r.dist = -999.9f;
  r.intensity = -1.0f;
  const float px = pos[i * 3], py = pos[i * 3 + 1], pz = pos[i * 3 + 2];
  // camera_p = w2c * world_p, then Project(camera_p, &image_p, &d)
  const float qx = sum3(v.w2c_R[0] * px, v.w2c_R[1] * py, v.w2c_R[2] * pz) + v.w2c_t[0];
  const float qy = sum3(v.w2c_R[3] * px, v.w2c_R[4] * py, v.w2c_R[5] * pz) + v.w2c_t[1];
  const float qz = sum3(v.w2c_R[6] * px, v.w2c_R[7] * py, v.w2c_R[8] * pz) + v.w2c_t[2];
  const float ix = v.fx / qz * qx + v.cx;
  const float iy = v.fy / qz * qy + v.cy;
  for (;;) {  // one pass; `break` = TestVertex returns false
    if (qz < 0.0f || ix < 0.0f || static_cast<float>(v.W - 1) < ix || iy < 0.0f ||
        static_cast<float>(v.H - 1) < iy)
      break;
    if (!(ix >= 0.0f && iy >= 0.0f)) break;  // NaN projections
    const int nx = static_cast<int>(roundf(ix)), ny = static_cast<int>(roundf(iy));
    const int64_t pi = int64_t(ny) * v.W + nx;
    if (o.use_mask && mask[pi] != 255) break;
    if (o.use_depth && depth[pi] < FLT_EPSILON) break;
    float r0, r1, r2;
    ray_w_at(v, ix, iy, r0, r1, r2);
    const float dot = -sum3(nrm[i * 3] * r0, nrm[i * 3 + 1] * r1, nrm[i * 3 + 2] * r2);
    const float angle = acosf(dot);
    if (o.th < angle) break;
    const float ox = v.c2w_t[0], oy = v.c2w_t[1], oz = v.c2w_t[2];
    const float t = closest_t(nodes, leaves, tris, stack, lane, ox, oy, oz, r0, r1, r2);
    const float vx = vert[i * 3], vy = vert[i * 3 + 1], vz = vert[i * 3 + 2];
    float hx = vx, hy = vy, hz = vz;
    const float distance = t < __builtin_huge_valf() ? t : FLT_MAX;
    if (t < __builtin_huge_valf()) {
      hx = ox + r0 * t;
      hy = oy + r1 * t;
      hz = oz + r2 * t;
      const float dx = hx - vx, dy = hy - vy, dz = hz - vz;
      if (sqrtf(sum3(dx * dx, dy * dy, dz * dz)) > o.eps) break;
    }
    if (o.use_depth) {
      const float hcz = sum3(v.w2c_R[6] * hx, v.w2c_R[7] * hy, v.w2c_R[8] * hz) + v.w2c_t[2];
      if (o.max_ddiff < fabsf(hcz - depth[pi])) break;
    }
    r.px = ix;
    r.py = iy;
    r.angle = angle;
    r.dist = distance;
    if (distance > o.max_dist || angle > o.max_angle) break;
    float c[3];
    if (o.interp == 0) {
      for (int k = 0; k < 3; k++) c[k] = static_cast<float>(color[pi * 3 + k]);
    } else {
      // BilinearInterpolation (include/ugu/util/image_util.h:155-193)
      int x0 = static_cast<int>(floorf(ix)), y0 = static_cast<int>(floorf(iy));
      int x1 = x0 + 1, y1 = y0 + 1;
      if (x0 < 0) x0 = 0;
      if (y0 < 0) y0 = 0;
      if (v.W <= x1) x1 = v.W - 1;
      if (v.H <= y1) y1 = v.H - 1;
      const float lu = ix - static_cast<float>(x0), lv = iy - static_cast<float>(y0);
      const uint8_t* c00 = color + (int64_t(y0) * v.W + x0) * 3;
      const uint8_t* c01 = color + (int64_t(y0) * v.W + x1) * 3;
      const uint8_t* c10 = color + (int64_t(y1) * v.W + x0) * 3;
      const uint8_t* c11 = color + (int64_t(y1) * v.W + x1) * 3;
      for (int k = 0; k < 3; k++) {
        const float f = (1.0f - lu) * (1.0f - lv) * static_cast<float>(c00[k]) +
                        lu * (1.0f - lv) * static_cast<float>(c01[k]) +
                        (1.0f - lu) * lv * static_cast<float>(c10[k]) +
                        lu * lv * static_cast<float>(c11[k]);
        c[k] = static_cast<float>(f2u8_x86(f));
      }
    }
    r.c0 = c[0];
    r.c1 = c[1];
    r.c2 = c[2];
    // Color2Gray (image_util.h:227-244): double arithmetic, cast to float
    r.intensity = static_cast<float>(0.2989 * double(c[0]) + 0.5870 * double(c[1]) +
                                     0.1140 * double(c[2]));
    r.visible = 1;
    break;
  }
  out[i] = r;
}

int launch_visibility(hipStream_t st, const DevView& v, const uint8_t* color, const float* depth,
                      const uint8_t* mask, const DevVisOpt& o, int n, const float* pos,
                      const float* vert, const float* nrm, const DevBvh& bvh, void* out) {
  if (n < 1) return 0;
  hipLaunchKernelGGL(k_visibility, dim3((n + 63) / 64), dim3(64), 0, st, v, color, depth, mask, o,
                     n, pos, vert, nrm, bvh.nodes, bvh.leaves, bvh.tris,
                     static_cast<DevVisItem*>(out));
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// views[i].tile_base = first 8x8 tile of view i in the concatenated tile list
int launch_raytrace(hipStream_t st, const DevView* views, int n_views, int total_tiles,
                    const DevBvh& bvh, const DevMesh& m, const DevOptions& o) {
  if (n_views < 1 || total_tiles < 1) return 0;
  // One view: each XCD takes one contiguous band of tiles (measured 0.61 vs
  // 0.75 ms at 1080p). Many views: runs of 16 workgroups alternate over the
  // XCDs (views differ in cost per tile; one band per XCD would leave some
  // XCDs with the expensive views: 13.8 vs 10.5 ms on the 64 ring views).
  const int64_t need = (int64_t(total_tiles) + kRtWaves - 1) / kRtWaves;
  const int64_t chunk = n_views == 1 ? (need + 7) / 8 : 16;
  const int64_t per = 8 * chunk;
  const int64_t blocks64 = (need + per - 1) / per * per;
  if (blocks64 > INT32_MAX) return -1;
  const int blocks = static_cast<int>(blocks64);
  hipLaunchKernelGGL(k_raytrace, dim3(blocks), dim3(64 * kRtWaves), 0, st, views, bvh.nodes,
                     bvh.leaves, bvh.tris, n_views, total_tiles, static_cast<int>(chunk), m, o);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace ugu_hip
