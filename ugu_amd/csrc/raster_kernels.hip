// HIP kernels of the gfx950 rasterizer.
//
// Pipeline for a batch of views (a single view is a batch of one):
//   k_face_soup   48-byte face records (positions, face normal) per mesh
//   k_clear       zero per-tile bin counters and the status words
//   k_setup_bin   per-face setup of the face loop (rasterizer.cc:162-193):
//                 z<0 reject, bbox, off-screen cull, pixel bbox, area; then
//                 appends (face id, skip depth) to the front or back segment
//                 of every screen tile its pixel bbox touches (a face of at
//                 most 4 bbox pixels only where it may cover a pixel centre)
//   k_raster      one workgroup per screen tile: stages the tile's faces in
//                 LDS (skipping faces proven hidden), expands (face, pixel)
//                 pairs so every lane tests one pair, resolves depth with a
//                 packed 64-bit (depth bits << 32 | face id) LDS atomicMin,
//                 then shades and writes every output pixel of the tile once
//                 (resolve loop rasterizer.cc:243-301, pixel_shader.h).
//   k_bwd_*       differentiable backward pass (no reference counterpart).
//
// Bit-exactness: all coverage arithmetic follows the reference's expression
// order; this file is compiled with -ffp-contract=off and correctly rounded
// fp32 division/sqrt (see the Makefile), and the pragma below forbids
// contraction locally as well.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>
#include <limits.h>

#include "device_common.h"

#pragma clang fp contract(off)

namespace ugu_hip {

// w2c_R * p + w2c_t, then PinholeCamera::Project (camera.h:380-387):
// x = fx / z * x + cx (divide first), y likewise, z kept.
UGU_DEV float4 project_point(const DevView& v, float px, float py, float pz) {
  float qx = sum3(v.w2c_R[0] * px, v.w2c_R[1] * py, v.w2c_R[2] * pz) + v.w2c_t[0];
  float qy = sum3(v.w2c_R[3] * px, v.w2c_R[4] * py, v.w2c_R[5] * pz) + v.w2c_t[1];
  float qz = sum3(v.w2c_R[6] * px, v.w2c_R[7] * py, v.w2c_R[8] * pz) + v.w2c_t[2];
  // fx / qz and fy / qz are the same correctly rounded quotient when fx == fy
  // (a wave-uniform branch: the second divide is not executed then)
  const float sx = v.fx / qz;
  float sy = sx;
  if (__builtin_expect(v.fy != v.fx, 0)) sy = v.fy / qz;
  float4 r;
  r.x = sx * qx + v.cx;
  r.y = sy * qy + v.cy;
  r.z = qz;
  r.w = 0.0f;
  return r;
}

// A face record is three float4: (x0 y0 z0 x1) (y1 z1 x2 y2) (z2 nx ny nz),
// the face's vertex positions and its face normal.
UGU_DEV void load_face_rec(const float4* soup, int f, float P[9], float N[3]) {
  const float4* r = soup + int64_t(f) * 3;
  const float4 a = r[0], b = r[1], c = r[2];
  P[0] = a.x; P[1] = a.y; P[2] = a.z;
  P[3] = a.w; P[4] = b.x; P[5] = b.y;
  P[6] = b.z; P[7] = b.w; P[8] = c.x;
  N[0] = c.y; N[1] = c.z; N[2] = c.w;
}

// True when the screen x, y of the three corners are all below 2^30 in
// magnitude (no NaN or infinity): the sum of their magnitudes propagates NaN
// and exceeds 2^30 whenever one of them does.
UGU_DEV bool small_screen(const float4& a, const float4& b, const float4& c) {
  return ((fabsf(a.x) + fabsf(b.x)) + (fabsf(c.x) + fabsf(a.y))) + (fabsf(b.y) + fabsf(c.y)) <
         1073741824.0f;
}

// Per-face setup of rasterizer.cc:164-193. Returns false when the
// reference skips the face or its pixel bbox is empty.
UGU_DEV bool face_setup(const float4& a, const float4& b, const float4& c,
                        int W, int H, int& x0, int& x1, int& y0, int& y1,
                        float& area) {
  if (a.z < 0.0f || b.z < 0.0f || c.z < 0.0f) return false;
  if (small_screen(a, b, c)) {
    // no NaN: std::min / std::max are the hardware min / max (their choice
    // between +0 and -0 aside, which ceil, floor and the comparisons below
    // do not see), and below 2^30 the x86 conversions are plain truncations
    const float xmin = fminf(fminf(a.x, b.x), c.x), xmax = fmaxf(fmaxf(a.x, b.x), c.x);
    const float ymin = fminf(fminf(a.y, b.y), c.y), ymax = fmaxf(fmaxf(a.y, b.y), c.y);
    if (xmin > static_cast<float>(W - 1) || xmax < 0.0f ||
        ymin > static_cast<float>(H - 1) || ymax < 0.0f)
      return false;
    x0 = max(0, static_cast<int>(ceilf(xmin)));
    x1 = min(W - 1, static_cast<int>(floorf(xmax)));
    y0 = max(0, static_cast<int>(ceilf(ymin)));
    y1 = min(H - 1, static_cast<int>(floorf(ymax)));
    area = edge_fn(a.x, a.y, b.x, b.y, c.x, c.y);
    if (fabsf(area) < FLT_MIN) return false;
    return x0 <= x1 && y0 <= y1;
  }
  float xmin = min3_std(a.x, b.x, c.x);
  float ymin = min3_std(a.y, b.y, c.y);
  float xmax = max3_std(a.x, b.x, c.x);
  float ymax = max3_std(a.y, b.y, c.y);
  if (xmin > static_cast<float>(W - 1) || xmax < 0.0f ||
      ymin > static_cast<float>(H - 1) || ymax < 0.0f)
    return false;
  x0 = max(0, f2i_x86(ceilf(xmin)));
  x1 = min(W - 1, f2i_sat(floorf(xmax)));
  y0 = max(0, f2i_x86(ceilf(ymin)));
  y1 = min(H - 1, f2i_sat(floorf(ymax)));
  area = edge_fn(a.x, a.y, b.x, b.y, c.x, c.y);
  if (fabsf(area) < FLT_MIN) return false;
  return x0 <= x1 && y0 <= y1;
}

// Coverage, perspective-correct weights and depth of one (face, pixel) pair
// (rasterizer.cc:196-226). Returns true when the pixel is covered.
UGU_DEV bool cover_test(float ax, float ay, float az, float bx, float by,
                        float bz, float cx, float cy, float cz, float area,
                        bool bf, float x, float y, float& z, float& w0o,
                        float& w1o, float& w2o) {
  float w0 = edge_fn(bx, by, cx, cy, x, y);
  float w1 = edge_fn(cx, cy, ax, ay, x, y);
  float w2 = edge_fn(ax, ay, bx, by, x, y);
  bool cov = (!bf && (w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f)) ||
             (bf && (w0 <= 0.0f && w1 <= 0.0f && w2 <= 0.0f));
  if (!cov) return false;
  w0 /= area;
  w1 /= area;
  w2 /= area;
  w0 /= az;
  w1 /= bz;
  w2 /= cz;
  z = 1.0f / (w0 + w1 + w2);
  w0o = w0 * z;
  w1o = w1 * z;
  w2o = w2 * z;
  return true;
}



// ---------------------------------------------------------------------------
// k_clear

// Four words per tile: front and back segment counters and the minimum skip
// depth of the back segment. The fourth word holds the tile's view (largest v
// with tile_base[v] <= tile), so the rasterizer reads its counters and finds
// its view with one load.
__global__ void k_clear(int32_t* counts, int32_t n, int32_t* status,
                        const int32_t* __restrict__ tile_base, int32_t n_views) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    int lo = 0, hi = n_views - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tile_base[mid] <= i) lo = mid; else hi = mid - 1;
    }
    // front count, back count, back-segment skip-depth minimum (bits; +inf), view
    reinterpret_cast<int4*>(counts)[i] = make_int4(0, 0, 0x7f800000, lo);
  }
  if (i < 4) status[i] = 0;
  if (i < kQueues) status[kQueueStride * (i + 1)] = 0;
}

// ---------------------------------------------------------------------------
// Mesh::CalcFaceNormal (src/mesh.cc:339-349) pieces, for the differentiable
// path where positions change between calls.

UGU_DEV void normalize3(float& x, float& y, float& z) {
  float n = sum3(x * x, y * y, z * z);
  if (n > 0.0f) {
    float s = sqrtf(n);
    x = x / s;
    y = y / s;
    z = z / s;
  }
}


// ---------------------------------------------------------------------------
// k_shade_rec: one 64-byte shading record per face (DevMesh::shade): the
// corners' uvs (uv[uv_indices]), the material id and the corners' vertex
// normals (normals[normal_indices]), copied bit for bit (zero where the mesh
// has none), so the resolve fetches a winner's shading inputs with one load.
// Layout (u0 v0 u1 v1)(u2 v2 mat n0x)(n0y n0z n1x n1y)(n1z n2x n2y n2z): the
// texture inputs are the first 32 bytes, the normals the last 36.
__global__ void k_shade_rec(DevMesh m, float4* __restrict__ out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.F) return;
  float u[6] = {0, 0, 0, 0, 0, 0}, n[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (m.uv && m.uvidx)
    for (int k = 0; k < 3; k++) {
      const float* q = m.uv + int64_t(m.uvidx[int64_t(f) * 3 + k]) * 2;
      u[2 * k] = q[0];
      u[2 * k + 1] = q[1];
    }
  if (m.normals && m.nidx)
    for (int k = 0; k < 3; k++) {
      const float* q = m.normals + int64_t(m.nidx[int64_t(f) * 3 + k]) * 3;
      n[3 * k] = q[0];
      n[3 * k + 1] = q[1];
      n[3 * k + 2] = q[2];
    }
  const int mat = m.matid ? m.matid[f] : 0;
  float4* r = out + int64_t(f) * 4;
  r[0] = make_float4(u[0], u[1], u[2], u[3]);
  r[1] = make_float4(u[4], u[5], __int_as_float(mat), n[0]);
  r[2] = make_float4(n[1], n[2], n[3], n[4]);
  r[3] = make_float4(n[5], n[6], n[7], n[8]);
}

// ---------------------------------------------------------------------------
// k_face_soup: one 48-byte record per face (three vertex positions, then the
// face normal) so setup and raster fetch a face with one dependent load
// instead of an index load and three vertex gathers. compute_normals:
// Mesh::CalcFaceNormal on device (differentiable path, the positions change
// every call; also written to fn_out); otherwise the mesh's face normals are
// copied (zero when it has none).

__global__ void k_face_soup(const float* __restrict__ verts, const int32_t* __restrict__ faces,
                            const float* __restrict__ fnormals, int32_t F, int compute_normals,
                            float4* __restrict__ soup, float* __restrict__ fn_out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  const int32_t* fi = faces + int64_t(f) * 3;
  float P[9];
  for (int k = 0; k < 3; k++) {
    const float* q = verts + int64_t(fi[k]) * 3;
    P[k * 3 + 0] = q[0];
    P[k * 3 + 1] = q[1];
    P[k * 3 + 2] = q[2];
  }
  float nx = 0.0f, ny = 0.0f, nz = 0.0f;
  if (compute_normals) {
    float ax = P[3] - P[0], ay = P[4] - P[1], az = P[5] - P[2];
    float bx = P[6] - P[0], by = P[7] - P[1], bz = P[8] - P[2];
    normalize3(ax, ay, az);
    normalize3(bx, by, bz);
    nx = ay * bz - az * by;
    ny = az * bx - ax * bz;
    nz = ax * by - ay * bx;
    normalize3(nx, ny, nz);
    if (fn_out) {
      fn_out[int64_t(f) * 3 + 0] = nx;
      fn_out[int64_t(f) * 3 + 1] = ny;
      fn_out[int64_t(f) * 3 + 2] = nz;
    }
  } else if (fnormals) {
    nx = fnormals[int64_t(f) * 3 + 0];
    ny = fnormals[int64_t(f) * 3 + 1];
    nz = fnormals[int64_t(f) * 3 + 2];
  }
  float4* r = soup + int64_t(f) * 3;
  r[0] = make_float4(P[0], P[1], P[2], P[3]);
  r[1] = make_float4(P[4], P[5], P[6], P[7]);
  r[2] = make_float4(P[8], nx, ny, nz);
}

// ---------------------------------------------------------------------------
// k_setup_bin: one thread per (face, view)

// Appends `entry` to a bin segment for every lane with `has`, in two halves.
// `seg` = 4 * tile for the front segment (filled upwards from the bin's first
// slot) and 4 * tile + 1 for the back segment (filled downwards from its last
// slot). append_issue groups the lanes that target the same segment (ballot +
// popcount rank) and the group leaders issue their atomicAdds together;
// append_finish broadcasts each group's base and stores the entries. A caller
// issues several appends before finishing any, so the wavefront waits for
// one atomic round trip for all of them. Entries of one segment are in no
// particular order (the rasterizer's result does not depend on it). When the
// two segments together overflow the bin they overwrite each other's slots;
// every slot the rasterizer reads still holds a valid entry, and the host
// renders again with larger bins.
struct AppendTicket {
  int ret;   // the group's first slot, in the leader lane
  int lr;    // leader lane | rank in the group << 8
  int seg;   // segment, -1 for lanes without an entry
};

// Lower bound of the skip depths of a wavefront's back-face entries (bits of
// a non-negative float; 0 when one has no bound): the leaders of back groups
// fold it into their tile's back-segment minimum.
UGU_DEV uint32_t back_zmin(bool back, float zskip) {
  uint32_t zmin = 0xffffffffu;
  if (__ballot(back))
    zmin = wave_min_u32(back ? (zskip >= 0.0f ? __float_as_uint(zskip) : 0u) : 0xffffffffu);
  return zmin;
}

UGU_DEV AppendTicket append_issue(const DevBins& bins, int seg, uint32_t zmin, bool has) {
  AppendTicket tk;
  tk.ret = 0;
  tk.lr = 0;
  tk.seg = has ? seg : -1;
  const uint64_t hasm = __builtin_amdgcn_ballot_w64(has);
  if (!hasm) return tk;
  // Each pass takes the lowest pending lane; the lanes with its segment form
  // one group, whose lowest lane (that one) leads it. A lane keeps its
  // group's mask; rank, leader and size follow from it after the loop.
  uint64_t todo = hasm, grp_mine = 0;
  while (todo) {
    const int l = __ffsll(static_cast<unsigned long long>(todo)) - 1;
    const int ls = __builtin_amdgcn_readlane(seg, l);
    const uint64_t grp = __builtin_amdgcn_ballot_w64(seg == ls) & hasm;
    grp_mine = (has && seg == ls) ? grp : grp_mine;
    todo &= ~grp;
  }
  const int rank = static_cast<int>(__builtin_amdgcn_mbcnt_hi(
      static_cast<uint32_t>(grp_mine >> 32),
      __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(grp_mine), 0u)));
  if (has && rank == 0) {
    tk.ret = atomicAdd(&bins.counts[seg], __popcll(grp_mine));
    if (seg & 1) atomicMin(reinterpret_cast<unsigned int*>(bins.counts) + (seg + 1), zmin);
  }
  const int leader = has ? __ffsll(static_cast<unsigned long long>(grp_mine)) - 1 : 0;
  tk.lr = leader | (rank << 8);
  return tk;
}

// Finishes the first `used` (wavefront-uniform) of R tickets: every group
// base is broadcast before the first store, so the wavefront waits once for
// all its atomics (a store issued before a later wait would be waited for as
// well).
template <int R>
UGU_DEV void append_finish(const DevBins& bins, const AppendTicket* tk, const int2* entry,
                           int used = R) {
  int slot[R];
#pragma unroll
  for (int r = 0; r < R; r++)
    slot[r] = r < used ? __shfl(tk[r].ret, tk[r].lr & 0xff) + (tk[r].lr >> 8) : 0;
#pragma unroll
  for (int r = 0; r < R; r++) {
    if (r < used && tk[r].seg >= 0 && slot[r] < bins.cap) {
      const int64_t row = int64_t(tk[r].seg >> 2) * bins.cap;
      bins.bins[row + ((tk[r].seg & 1) ? bins.cap - 1 - slot[r] : slot[r])] = *entry;
    }
  }
}

// Entry i of a tile's bin: the front segment's nA entries, then the back
// segment read from the bin's last slot downwards.
// Each entry is read once, so the load is non-temporal: the stream of entries
// does not evict the face records, shading records and texels the rasterizer
// reads again (64-view batch: rasterizer -2%).
UGU_DEV int2 bin_entry(const int2* bin, int nA, int cap, int i) {
  const int2* p = i < nA ? bin + i : bin + (cap - 1 - (i - nA));
  const long long q = __builtin_nontemporal_load(reinterpret_cast<const long long*>(p));
  return make_int2(static_cast<int>(q & 0xffffffffll), static_cast<int>(q >> 32));
}

// Skip depth of a face in a view: a lower bound of every depth the coverage
// test (cover_test) can compute for the face at a pixel of its bbox, or -1
// when no useful bound exists. Once every in-view pixel of a tile holds a depth
// d <= dmax, a face whose skip depth exceeds dmax (1 + 2^-8) can neither win a
// pixel nor produce a depth below FLT_MIN, so the rasterizer drops it without
// fetching it.
// Bound: with L the larger bbox extent, rounding moves each edge function by
// at most 8u L^2 (u = 2^-24) and the area by as much, so the normalised
// weights sum to at most 1 + 32u L^2 / |area|; with 2^-9 L^2 <= |area| that is
// 1 + 2^-10, five more roundings cost < 2^-21, and
// z >= min(za, zb, zc) (1 - 2^-10 - 2^-20). NaN or huge values fail every
// comparison and give -1.
UGU_DEV float skip_depth(const float4& a, const float4& b, const float4& c, float area) {
  float ex, ey;
  if (small_screen(a, b, c)) {
    // no NaN: the hardware min / max (see face_setup)
    ex = fmaxf(fmaxf(a.x, b.x), c.x) - fminf(fminf(a.x, b.x), c.x);
    ey = fmaxf(fmaxf(a.y, b.y), c.y) - fminf(fminf(a.y, b.y), c.y);
  } else {
    ex = max3_std(a.x, b.x, c.x) - min3_std(a.x, b.x, c.x);
    ey = max3_std(a.y, b.y, c.y) - min3_std(a.y, b.y, c.y);
  }
  if (!(ex <= 1048576.0f && ey <= 1048576.0f)) return -1.0f;
  const float L = ex > ey ? ex : ey;
  if (!(L * L * 0.001953125f <= fabsf(area))) return -1.0f;
  if (!(a.z > 0.0f && b.z > 0.0f && c.z > 0.0f)) return -1.0f;
  // three positive depths: std::min is the hardware min
  return fminf(fminf(a.z, b.z), c.z);
}

// True when the tile's current depth bound proves the face hidden.
UGU_DEV bool skip_face(float zskip, float dmax) {
  return dmax >= 1e-18f && zskip > dmax * 1.00390625f;
}

// Views per thread: the face's indices are loaded once and reused for
// kSetupViews consecutive views (grid.y = ceil(n_views / kSetupViews)).
constexpr int kSetupViews = 8;

// Bin appends a setup thread issues before waiting for their atomics: a face
// whose pixel bbox spans up to this many tiles costs one atomic round trip
// per view.
constexpr int kAppendRounds = 4;

// Chunks whose faces all have at most this many (face, pixel) pairs expand
// them through an LDS pair list; otherwise by binary search over the prefix.
constexpr int kScatterMaxPairs = 32;

// Faces whose pixel bbox holds at most this many pixel centres are tested at
// those centres before they are binned.
constexpr int kCoverProbe = 4;

// Tiles of a small face's bbox (at most kCoverProbe pixel centres, so at most
// 2 x 2 tiles) holding a centre the face may cover: bit (tx - tx0) + 2 (ty -
// ty0). The edge functions are the coverage test's (cover_test, same operands
// and order, so the same bits); a centre where they are neither all >= 0 nor
// all <= 0 is covered for neither back-face flag, so the rasterizer's pair
// test would reject the face there whatever the ray. A face with no such
// centre (most sub-pixel faces) gets no bin entry at all.
template <int TW, int TH>
UGU_DEV uint32_t probe_cover(const float4& a, const float4& b, const float4& c, int x0, int x1,
                             int y0, int y1, int tx0, int ty0) {
  static_assert(TW >= kCoverProbe && TH >= kCoverProbe, "a probed bbox spans at most 2 x 2 tiles");
  uint32_t m = 0u;
  int px = x0, py = y0;
#pragma unroll
  for (int k = 0; k < kCoverProbe; k++) {
    if (py <= y1) {
      const float x = static_cast<float>(px), y = static_cast<float>(py);
      const float w0 = edge_fn(b.x, b.y, c.x, c.y, x, y);
      const float w1 = edge_fn(c.x, c.y, a.x, a.y, x, y);
      const float w2 = edge_fn(a.x, a.y, b.x, b.y, x, y);
      // all >= 0 or all <= 0 (a NaN fails both in cover_test; here fminf /
      // fmaxf drop it, which only keeps more centres)
      if (fminf(fminf(w0, w1), w2) >= 0.0f || fmaxf(fmaxf(w0, w1), w2) <= 0.0f)
        m |= 1u << ((static_cast<int>(static_cast<unsigned>(px) / TW) - tx0) +
                    2 * (static_cast<int>(static_cast<unsigned>(py) / TH) - ty0));
      if (++px > x1) {
        px = x0;
        py++;
      }
    }
  }
  return m;
}

template <int TW, int TH>
__global__ __launch_bounds__(256) void k_setup_bin(
    const DevView* __restrict__ views, int n_views, DevMesh mesh,
    DevBins bins) {
  // The block's 256 faces share their vertices (about 150 distinct ones on a
  // regular mesh): every distinct vertex is projected once per view into LDS
  // and the faces read their corners from there, about 0.6 projections per
  // face instead of three. Without the block lists each face projects its
  // own corners.
  __shared__ float4 sp[768];
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const bool shared = mesh.blk_off != nullptr;
  int vbase = 0, nu = 0;
  float P[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  int loc[3] = {0, 0, 0};
  if (shared) {
    vbase = mesh.blk_off[blockIdx.x];
    nu = mesh.blk_off[blockIdx.x + 1] - vbase;
    if (f < mesh.F) {
      const uint32_t lc = mesh.face_loc[f];
      loc[0] = lc & 1023u;
      loc[1] = (lc >> 10) & 1023u;
      loc[2] = lc >> 20;
    }
    for (int j = 0; j < 3; j++) {
      const int k = threadIdx.x + j * 256;
      if (k < nu) {
        const float* q = mesh.verts + int64_t(mesh.blk_vid[vbase + k]) * 3;
        P[j * 3 + 0] = q[0];
        P[j * 3 + 1] = q[1];
        P[j * 3 + 2] = q[2];
      }
    }
  } else if (f < mesh.F) {
    float N[3];
    load_face_rec(mesh.soup, f, P, N);
  }
  const int vb = blockIdx.y * kSetupViews;
  const int ve = min(vb + kSetupViews, n_views);
  for (int vi = vb; vi < ve; vi++) {
    const DevView& v = views[vi];
    bool has = false, back = false;
    float zskip = -1.0f;
    int tx0 = 0, tx1 = -1, ty0 = 0, ty1 = -1;
    // tiles of a probed face's bbox that receive it, bit (tx - tx0) + 2 (ty -
    // ty0); all ones for a face that is not probed (every bbox tile)
    uint32_t tmask = ~0u;
    float4 a, b, c;
    if (shared) {
      for (int j = 0; j < 3; j++) {
        const int k = threadIdx.x + j * 256;
        if (k < nu) sp[k] = project_point(v, P[j * 3 + 0], P[j * 3 + 1], P[j * 3 + 2]);
      }
      __syncthreads();
      a = sp[loc[0]];
      b = sp[loc[1]];
      c = sp[loc[2]];
      __syncthreads();  // the next view overwrites the window
    } else {
      a = project_point(v, P[0], P[1], P[2]);
      b = project_point(v, P[3], P[4], P[5]);
      c = project_point(v, P[6], P[7], P[8]);
    }
    if (f < mesh.F) {
      int x0, x1, y0, y1;
      float area;
      if (face_setup(a, b, c, v.W, v.H, x0, x1, y0, y1, area)) {
        has = true;
        // faces that are clockwise on screen are mostly hidden back faces:
        // they go to the back segment so the rasterizer sees them last
        back = !(area > 0.0f);
        zskip = skip_depth(a, b, c, area);
        tx0 = x0 / TW;
        tx1 = x1 / TW;
        ty0 = y0 / TH;
        ty1 = y1 / TH;
        if ((x1 - x0 + 1) * (y1 - y0 + 1) <= kCoverProbe) {
          tmask = probe_cover<TW, TH>(a, b, c, x0, x1, y0, y1, tx0, ty0);
          has = tmask != 0u;
        }
      }
    }
    // one append per tile of the face's pixel bbox (of a probed small face,
    // per tile where it may cover a centre): up to kAppendRounds
    // rounds are issued before any is finished, further rounds (faces over
    // more tiles) one at a time
    const int2 entry = make_int2(f, __float_as_int(zskip));
    // one bound for all rounds: the minimum over every back face of the
    // wavefront is also below the skip depth of each back face of a round
    const uint32_t zmin = back_zmin(has && back, zskip);
    const int nx = tx1 - tx0 + 1;
    const int rest = has ? max(nx * (ty1 - ty0 + 1) - kAppendRounds, 0) : 0;
    int tx = tx0, ty = ty0;
    // a probed face visits only its marked tiles, one per round
    const bool probed = tmask != ~0u;
    uint32_t pm = tmask;
    AppendTicket tk[kAppendRounds];
    int used = 0;  // rounds issued: until no lane has a tile left
#pragma unroll
    for (int r = 0; r < kAppendRounds; r++) {
      if (!__builtin_amdgcn_ballot_w64(has)) break;
      used = r + 1;
      int ax = tx, ay = ty;
      if (probed) {
        const int bit = __ffs(pm) - 1;
        ax = tx0 + (bit & 1);
        ay = ty0 + (bit >> 1);
      }
      const int seg = has ? 4 * (v.tile_base + ay * v.tiles_x + ax) + (back ? 1 : 0) : -1;
      tk[r] = append_issue(bins, seg, zmin, has);
      if (probed) {
        pm &= pm - 1u;
        has = has && pm != 0u;
      } else if (has && ++tx > tx1) {
        tx = tx0;
        if (++ty > ty1) has = false;
      }
    }
    append_finish<kAppendRounds>(bins, tk, &entry, used);
    // Faces over more than kAppendRounds tiles (large triangles): their
    // remaining tiles, all lanes' together, 64 per round -- lane j of a round
    // takes pair base + j of the wavefront's list, finds its face's lane (the
    // last lane whose exclusive prefix is <= the pair) and the tile ordinal.
    if (__any(rest > 0)) {
      const int lane = __lane_id();
      const int incl = wave_incl_sum(rest);
      const int excl = incl - rest;
      const int total = __builtin_amdgcn_readlane(incl, 63);
      for (int base = 0; base < total; base += 64) {
        const int p = base + lane;
        int src = 0;
        for (int step = 32; step >= 1; step >>= 1)
          if (__shfl(excl, src + step) <= p) src += step;
        const int k = p - __shfl(excl, src) + kAppendRounds;
        // lanes past the last pair (act false) may land on a face without
        // tiles (nx 0): clamp so their unused division stays defined
        const int snx = max(__shfl(nx, src), 1);
        const int sty = __shfl(ty0, src) + k / snx;
        const int stx = __shfl(tx0, src) + k % snx;
        const int sback = __shfl(back ? 1 : 0, src);
        const int2 e = make_int2(__shfl(f, src), __shfl(__float_as_int(zskip), src));
        const bool act = p < total;
        const AppendTicket t1 = append_issue(
            bins, act ? 4 * (v.tile_base + sty * v.tiles_x + stx) + sback : -1, zmin, act);
        append_finish<1>(bins, &t1, &e);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// k_raster

template <int NT>
struct RasterLds {
  unsigned long long key[NT];  // packed (depth bits << 32 | face id)
  float ray[3][NT];            // ray_w of each tile pixel
  float fv[13][NT];            // staged faces: 3 projected verts, area, face normal
  int32_t fid[NT];
  int32_t box[NT];             // lx0 | ly0 << 8 | nx << 16 (tile-local pixel bbox)
  // exclusive prefix of pair counts (binary-search expansion) or the
  // (face slot | pair index << 8) list of one round (pair-list expansion)
  int32_t pairs[2 * NT + 1];
  int32_t wsum[NT / 64 + 1];
  int32_t wbig[NT / 64];       // per-wave largest pair count of a face
  uint32_t flag[NT / 32];      // pixel saw a depth outside [FLT_MIN, inf] (bit per pixel)
  uint32_t wmax[NT / 64];       // per-wave max of the pixels' depth bits
  int32_t any_flag;
  int32_t item;
  int4 runseg[4];              // counters of the persistent block's current run
};

// Exact sequential z-buffer for one pixel (slow path): the result of the
// reference's update rule `if (d < FLT_MIN || z < d)` (rasterizer.cc:228-237),
// applied in increasing face id over the faces of the tile's bin that cover the
// pixel, starting from d = 0. Used only for pixels where some covering face
// produced a depth below FLT_MIN or NaN, where the packed atomicMin order is
// not the sequential order.
//
// Without NaN depths the rule has a closed form, computed in two passes over
// the bin: a covering face with depth below FLT_MIN is always taken, so the
// last such face E (largest id) resets the pixel, the next face is taken
// unconditionally, and from there the strict-less rule keeps the first
// smallest depth. The winner is therefore the smallest (depth, id) among the
// covering faces after E with depth >= FLT_MIN, or E itself when there is none.
// A NaN depth is taken only right after a reset and then blocks every later
// face; when a covering face has one, the rule is simulated in id order
// instead (one pass over the bin per covering face).
UGU_DEV void sequential_pixel(const DevView& v, const DevMesh& mesh,
                              const int2* bin, int nA, int cap, int n,
                              int gx, int gy, float r0, float r1, float r2,
                              int& out_fid) {
  const float x = static_cast<float>(gx), y = static_cast<float>(gy);
  // face of bin entry i and its depth at the pixel; false when it does not
  // cover the pixel
  auto covers = [&](int i, int& f, float& z) {
    f = bin_entry(bin, nA, cap, i).x;
    float P[9], N[3];
    load_face_rec(mesh.soup, f, P, N);
    const float4 a = project_point(v, P[0], P[1], P[2]);
    const float4 b = project_point(v, P[3], P[4], P[5]);
    const float4 c = project_point(v, P[6], P[7], P[8]);
    int x0, x1, y0, y1;
    float area;
    if (!face_setup(a, b, c, v.W, v.H, x0, x1, y0, y1, area)) return false;
    if (gx < x0 || gx > x1 || gy < y0 || gy > y1) return false;
    const bool bf = mesh.fnormals != nullptr && sum3(N[0] * r0, N[1] * r1, N[2] * r2) > 0.0f;
    float w0, w1, w2;
    return cover_test(a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, area, bf, x, y, z, w0, w1,
                      w2);
  };
  // pass 1: the last reset face and whether any depth is NaN
  int emax = -1;
  bool any_nan = false;
  for (int i = 0; i < n; i++) {
    int f;
    float z;
    if (!covers(i, f, z)) continue;
    if (z != z) any_nan = true;
    else if (z < FLT_MIN) emax = max(emax, f);
  }
  if (!any_nan) {
    // pass 2: first smallest depth after the reset
    unsigned long long best = ~0ull;
    for (int i = 0; i < n; i++) {
      int f;
      float z;
      if (!covers(i, f, z) || f <= emax || !(z >= FLT_MIN)) continue;
      const unsigned long long key =
          (static_cast<unsigned long long>(__float_as_uint(z)) << 32) | static_cast<unsigned int>(f);
      best = key < best ? key : best;
    }
    out_fid = best != ~0ull ? static_cast<int>(best & 0xffffffffull) : emax;
    return;
  }
  // some depth is NaN: the rule in id order
  int cur = -1;
  float d = 0.0f;
  int dfid = -1;
  for (;;) {
    int best = INT_MAX;
    float bestz = 0.0f;
    for (int i = 0; i < n; i++) {
      int f;
      float z;
      if (covers(i, f, z) && f > cur && f < best) {
        best = f;
        bestz = z;
      }
    }
    if (best == INT_MAX) break;
    if (d < FLT_MIN || bestz < d) {
      d = bestz;
      dfid = best;
    }
    cur = best;
  }
  out_fid = dfid;
}


// Bilinear sample of an f32 RGB texture with clamp-to-edge addressing
// (differentiable path). Returns the top-left texel and the local weights.
UGU_DEV void ftex_coords(int tw, int th, float iu, float iv, int& x0, int& y0,
                         int& x1, int& y1, float& lu, float& lv) {
  const float fx = iu * static_cast<float>(tw - 1);
  const float fy = (1.0f - iv) * static_cast<float>(th - 1);
  const float flx = floorf(fx), fly = floorf(fy);
  lu = fx - flx;
  lv = fy - fly;
  int ix = f2i_sat(flx), iy = f2i_sat(fly);
  x0 = min(max(ix, 0), tw - 1);
  y0 = min(max(iy, 0), th - 1);
  x1 = min(max(ix + 1, 0), tw - 1);
  y1 = min(max(iy + 1, 0), th - 1);
}

// ---------------------------------------------------------------------------
// Differentiable image model, evaluated in float64 for one pixel of a given
// face (visibility comes from the float32 rasterizer; the shading and its
// derivatives are computed in double so the gradients match the exact model,
// not the reference's float32 rounding of thin triangles).

struct DiffEval {
  double X[3], Y[3], iZ[3], Qx[3], Qy[3];  // iZ = 1 / camera depth
  double e[3], A, iA, q[3], s, is, b0, b1, b2;
  double c[3], t[3], dcdu[3], dcdv[3], dtdu[3], dtdv[3];
  double lu, lv;
  int tx0, ty0, tx1, ty1;
};

UGU_DEV int clamp_texel(double v, int hi) {
  if (!(v >= 0.0)) return 0;
  if (v > static_cast<double>(hi)) return hi;
  return static_cast<int>(v);
}

UGU_DEV void diff_eval(const DevView& v, const DevMesh& m, int fid, int px, int py,
                       const int32_t* vid, DiffEval& e) {
  const double x = px, y = py;
  for (int i = 0; i < 3; i++) {
    const float* P = m.verts + int64_t(vid[i]) * 3;
    double Q[3];
    for (int r = 0; r < 3; r++)
      Q[r] = double(v.w2c_R[r * 3 + 0]) * P[0] + double(v.w2c_R[r * 3 + 1]) * P[1] +
             double(v.w2c_R[r * 3 + 2]) * P[2] + double(v.w2c_t[r]);
    e.Qx[i] = Q[0];
    e.Qy[i] = Q[1];
    e.iZ[i] = 1.0 / Q[2];
    e.X[i] = double(v.fx) * Q[0] * e.iZ[i] + double(v.cx);
    e.Y[i] = double(v.fy) * Q[1] * e.iZ[i] + double(v.cy);
  }
  auto E = [](double ax, double ay, double bx, double by, double cx, double cy) {
    return (cx - ax) * (by - ay) - (cy - ay) * (bx - ax);
  };
  e.A = E(e.X[0], e.Y[0], e.X[1], e.Y[1], e.X[2], e.Y[2]);
  e.e[0] = E(e.X[1], e.Y[1], e.X[2], e.Y[2], x, y);
  e.e[1] = E(e.X[2], e.Y[2], e.X[0], e.Y[0], x, y);
  e.e[2] = E(e.X[0], e.Y[0], e.X[1], e.Y[1], x, y);
  // reciprocals once, then products: the model is evaluated in float64 and
  // compared with a float64 reference, so a few ulps of rounding are free
  e.iA = 1.0 / e.A;
  for (int i = 0; i < 3; i++) e.q[i] = e.e[i] * e.iA * e.iZ[i];
  e.s = e.q[0] + e.q[1] + e.q[2];
  e.is = 1.0 / e.s;
  e.b1 = e.q[1] * e.is;
  e.b2 = e.q[2] * e.is;
  e.b0 = 1.0 - e.b1 - e.b2;
  for (int k = 0; k < 3; k++) {
    e.c[k] = 1.0;
    e.t[k] = 1.0;
    e.dcdu[k] = e.dcdv[k] = e.dtdu[k] = e.dtdv[k] = 0.0;
  }
  e.lu = e.lv = 0.0;
  e.tx0 = e.ty0 = e.tx1 = e.ty1 = 0;
  if (m.vcolors) {
    const float* C0 = m.vcolors + int64_t(vid[0]) * 3;
    const float* C1 = m.vcolors + int64_t(vid[1]) * 3;
    const float* C2 = m.vcolors + int64_t(vid[2]) * 3;
    for (int k = 0; k < 3; k++) {
      e.c[k] = e.b0 * C0[k] + e.b1 * C1[k] + e.b2 * C2[k];
      e.dcdu[k] = double(C1[k]) - C0[k];
      e.dcdv[k] = double(C2[k]) - C0[k];
    }
  }
  if (m.ftex) {
    const int32_t* ui = m.uvidx + int64_t(fid) * 3;
    const float* a = m.uv + int64_t(ui[0]) * 2;
    const float* b = m.uv + int64_t(ui[1]) * 2;
    const float* c = m.uv + int64_t(ui[2]) * 2;
    const double iu = e.b0 * a[0] + e.b1 * b[0] + e.b2 * c[0];
    const double iv = e.b0 * a[1] + e.b1 * b[1] + e.b2 * c[1];
    const int tw = m.ftex_w, th = m.ftex_h;
    const double fxp = iu * (tw - 1), fyp = (1.0 - iv) * (th - 1);
    const double flx = floor(fxp), fly = floor(fyp);
    e.lu = fxp - flx;
    e.lv = fyp - fly;
    e.tx0 = clamp_texel(flx, tw - 1);
    e.ty0 = clamp_texel(fly, th - 1);
    e.tx1 = clamp_texel(flx + 1.0, tw - 1);
    e.ty1 = clamp_texel(fly + 1.0, th - 1);
    const float* p00 = m.ftex + (int64_t(e.ty0) * tw + e.tx0) * 3;
    const float* p01 = m.ftex + (int64_t(e.ty0) * tw + e.tx1) * 3;
    const float* p10 = m.ftex + (int64_t(e.ty1) * tw + e.tx0) * 3;
    const float* p11 = m.ftex + (int64_t(e.ty1) * tw + e.tx1) * 3;
    const double sx = tw - 1, sy = -(th - 1.0);
    const double dlu_du = sx * (double(b[0]) - a[0]), dlu_dv = sx * (double(c[0]) - a[0]);
    const double dlv_du = sy * (double(b[1]) - a[1]), dlv_dv = sy * (double(c[1]) - a[1]);
    const double lu = e.lu, lv = e.lv;
    for (int k = 0; k < 3; k++) {
      e.t[k] = (1 - lu) * (1 - lv) * p00[k] + lu * (1 - lv) * p01[k] +
               (1 - lu) * lv * p10[k] + lu * lv * p11[k];
      const double da = (1 - lv) * (double(p01[k]) - p00[k]) + lv * (double(p11[k]) - p10[k]);
      const double db = (1 - lu) * (double(p10[k]) - p00[k]) + lu * (double(p11[k]) - p01[k]);
      e.dtdu[k] = da * dlu_du + db * dlv_du;
      e.dtdv[k] = da * dlu_dv + db * dlv_dv;
    }
  }
}

// G-buffer planes of one pixel (see UguGBuffer in ugu_hip.h). u, v: the
// winner's perspective-correct weights of its second and third vertex; P: its
// world vertex positions (null when no position plane is written); uvm: the
// mesh when the uv plane is written and the mesh has uvs; n: the shading normal.
template <bool STREAM>
UGU_DEV void write_gbuffer(const DevView& v, int64_t pidx, bool covered, float u, float vv,
                           bool has_p, const float* P, const DevMesh* uvm, float n0, float n1,
                           float n2, int fid = 0) {
  float b[3] = {0.0f, 0.0f, 0.0f}, t[3] = {0.0f, 0.0f, 0.0f};
  float pw[3] = {0.0f, 0.0f, 0.0f}, pc[3] = {0.0f, 0.0f, 0.0f};
  if (covered) {
    const float w = 1.0f - u - vv;
    b[0] = u;
    b[1] = vv;
    b[2] = 1.0f - u - vv;
    if (uvm) {
      const int32_t* ui = uvm->uvidx + int64_t(fid) * 3;
      const float* a = uvm->uv + int64_t(ui[0]) * 2;
      const float* bb = uvm->uv + int64_t(ui[1]) * 2;
      const float* c = uvm->uv + int64_t(ui[2]) * 2;
      t[0] = w * a[0] + u * bb[0] + vv * c[0];
      t[1] = w * a[1] + u * bb[1] + vv * c[1];
      t[2] = 1.0f - t[0] - t[1];
    }
    if (has_p) {
      for (int k = 0; k < 3; k++) pw[k] = w * P[k] + u * P[3 + k] + vv * P[6 + k];
      for (int r = 0; r < 3; r++)
        pc[r] = v.w2c_t[r] + sum3(v.w2c_R[r * 3 + 0] * pw[0], v.w2c_R[r * 3 + 1] * pw[1],
                                  v.w2c_R[r * 3 + 2] * pw[2]);
    }
  } else {
    n0 = n1 = n2 = 0.0f;
  }
  if (v.bary) put3<STREAM>(v.bary, pidx, b[0], b[1], b[2]);
  if (v.uvo) put3<STREAM>(v.uvo, pidx, t[0], t[1], t[2]);
  if (v.pos_wld) put3<STREAM>(v.pos_wld, pidx, pw[0], pw[1], pw[2]);
  if (v.pos_cam) put3<STREAM>(v.pos_cam, pidx, pc[0], pc[1], pc[2]);
  if (v.nrm_wld) put3<STREAM>(v.nrm_wld, pidx, n0, n1, n2);
  if (v.geo_id) put<STREAM>(v.geo_id, pidx, covered ? 0 : -1);
}

// One screen tile; `seg` is the tile's counter record (front count, back
// count, back-segment minimum skip depth, view index).
template <int TW, int TH, int MODE>
UGU_DEV void raster_tile(RasterLds<TW * TH>& s, int gtile, const int4 seg, const DevView* views,
                         const DevMesh& mesh,
                         const DevBins& bins, const DevOptions& opt) {
  constexpr int NT = TW * TH;
  // output planes streamed past L2 when the tile's rows are long enough
  constexpr bool S = TW >= 16;
  const int t = threadIdx.x;

  const int count = seg.x + seg.y;
  const DevView& v = views[seg.w];
  const int lt = gtile - v.tile_base;
  const int tyi = lt / v.tiles_x;
  const int txi = lt - tyi * v.tiles_x;
  const int px0 = txi * TW, py0 = tyi * TH;

  // this thread's pixel
  const int lx = t % TW, ly = t / TW;
  const int gx = px0 + lx, gy = py0 + ly;
  if (count == 0) {
    // no face touches the tile: every pixel keeps the initial values of
    // ValidateAndInitBeforeRender (util_private.cc:68-87)
    if (gx < v.W && gy < v.H) {
      const int64_t pidx = int64_t(gy) * v.W + gx;
      if (MODE == 0) {
        if (v.color) put3<S>(v.color, pidx, uint8_t(0), uint8_t(0), uint8_t(0));
        if (v.depth) put<S>(v.depth, pidx, 0.0f);
        if (v.normal) put3<S>(v.normal, pidx, 0.0f, 0.0f, 0.0f);
        if (v.mask) put<S>(v.mask, pidx, uint8_t(0));
        write_gbuffer<S>(v, pidx, false, 0.0f, 0.0f, false, nullptr, nullptr, 0.0f, 0.0f, 0.0f);
      } else {
        put3<S>(v.image, pidx, 0.0f, 0.0f, 0.0f);
      }
      if (v.face_id) put<S>(v.face_id, pidx, -1);
    }
    return;
  }
  // front segment first, then the back segment (both clipped to the bin)
  const int nA = seg.x < bins.cap ? seg.x : bins.cap;
  const int nB = seg.y < bins.cap - nA ? seg.y : bins.cap - nA;
  const int n = nA + nB;
  const int2* bin = bins.bins + int64_t(gtile) * bins.cap;
  // the first chunk's entries are loaded while the rays are computed; every
  // chunk then loads the next chunk's entries before its own pairs
  const int chunks_a = (nA + NT - 1) / NT, chunks = chunks_a + (nB + NT - 1) / NT;
  int2 e_next = make_int2(0, 0);
  if (t < (chunks_a > 0 ? nA : n)) e_next = bin_entry(bin, nA, bins.cap, t);

  float r0, r1, r2;
  ray_w_at(v, static_cast<float>(gx), static_cast<float>(gy), r0, r1, r2);
  s.key[t] = ~0ull;
  if (t < NT / 32) s.flag[t] = 0u;
  s.ray[0][t] = r0;
  s.ray[1][t] = r1;
  s.ray[2][t] = r2;
  if (t == 0) s.any_flag = 0;
  // When face ids fit 24 bits the depth key also carries the LDS slot of the
  // face (below the id, so ties still go to the lowest id). After every chunk
  // each thread copies the staged record of its pixel's current winner into
  // registers when the winner changed, so the resolve needs no second gather
  // and projection of the winning face.
  static_assert(NT <= 256, "slot keys hold 8-bit LDS slots");
  const bool slot_key = mesh.F <= (1 << 24);
  unsigned long long seen = ~0ull;  // key whose coverage the registers hold
  float wz = 0.0f, ww0 = 0.0f, ww1 = 0.0f, ww2 = 0.0f;
  bool wbf = false;
  // overflowing tiles report their true occupancy; the host re-renders with
  // bins that large (one same-address atomic per overflowing tile only)
  if (t == 0 && count > bins.cap) atomicMax(&bins.status[0], count);
  __syncthreads();

  // Largest depth any in-view pixel of the tile holds after the previous
  // chunks (NaN while some pixel has none). Keys only decrease, so a face
  // whose skip depth exceeds it is dropped before it is fetched. The back
  // segment starts a new chunk so that all of its faces are tested against
  // the depths the front faces left.
  float dmax = __uint_as_float(0xffffffffu);
  for (int ch = 0; ch < chunks; ch++) {
    // the whole back segment is skipped when even its smallest skip depth
    // exceeds the tile's depth bound
    if (ch == chunks_a && skip_face(__uint_as_float(static_cast<uint32_t>(seg.z)), dmax)) break;
    const int base = ch < chunks_a ? ch * NT : nA + (ch - chunks_a) * NT;
    const int end = ch < chunks_a ? nA : n;
    const bool more = ch + 1 < chunks;
    const int i = base + t;
    const int2 e = e_next;
    if (more) {
      const int nbase = ch + 1 < chunks_a ? (ch + 1) * NT : nA + (ch + 1 - chunks_a) * NT;
      const int nend = ch + 1 < chunks_a ? nA : n;
      e_next = nbase + t < nend ? bin_entry(bin, nA, bins.cap, nbase + t) : make_int2(0, 0);
    }
    // stage one chunk of faces
    int c = 0;
    if (i < end && !skip_face(__int_as_float(e.y), dmax)) {
      const int f = e.x;
      float P[9], N[3];
      load_face_rec(mesh.soup, f, P, N);
      const float4 a = project_point(v, P[0], P[1], P[2]);
      const float4 b = project_point(v, P[3], P[4], P[5]);
      const float4 cc = project_point(v, P[6], P[7], P[8]);
      int x0, x1, y0, y1;
      float area;
      if (face_setup(a, b, cc, v.W, v.H, x0, x1, y0, y1, area)) {
        const int lx0 = max(x0, px0) - px0, lx1 = min(x1, px0 + TW - 1) - px0;
        const int ly0 = max(y0, py0) - py0, ly1 = min(y1, py0 + TH - 1) - py0;
        if (lx0 <= lx1 && ly0 <= ly1) {
          const int nx = lx1 - lx0 + 1;
          c = nx * (ly1 - ly0 + 1);
          s.box[t] = lx0 | (ly0 << 8) | (nx << 16);
        }
      }
      s.fv[0][t] = a.x; s.fv[1][t] = a.y; s.fv[2][t] = a.z;
      s.fv[3][t] = b.x; s.fv[4][t] = b.y; s.fv[5][t] = b.z;
      s.fv[6][t] = cc.x; s.fv[7][t] = cc.y; s.fv[8][t] = cc.z;
      s.fv[9][t] = area;
      s.fv[10][t] = N[0]; s.fv[11][t] = N[1]; s.fv[12][t] = N[2];
      s.fid[t] = f;
    }
    // block exclusive scan of the pair counts; a chunk without any pair
    // (every face skipped or off the tile) ends at the barrier
    const int lane = t & 63, wid = t >> 6;
    const int x = wave_incl_sum(c);
    const int cmax = wave_incl_max(c);
    if (lane == 63) {
      s.wsum[wid] = x;
      s.wbig[wid] = cmax;
    }
    if (!__syncthreads_or(c > 0)) continue;
    int wbase = 0, total = 0, big = 0;
    for (int w = 0; w < NT / 64; w++) {
      if (w < wid) wbase += s.wsum[w];
      total += s.wsum[w];
      big = max(big, s.wbig[w]);
    }
    const int pre = wbase + x - c;  // first pair of this thread's face
    const bool has_fn = mesh.fnormals != nullptr;
    // One (face slot, tile pixel) pair: coverage, depth and the packed LDS
    // atomicMin.
    auto test_pair = [&](int j, int pix) {
      const int pxl = static_cast<int>(static_cast<unsigned>(pix) % TW);
      const int pyl = static_cast<int>(static_cast<unsigned>(pix) / TW);
      const float rx = s.ray[0][pix], ry = s.ray[1][pix], rz = s.ray[2][pix];
      const bool bf = has_fn && sum3(s.fv[10][j] * rx, s.fv[11][j] * ry, s.fv[12][j] * rz) > 0.0f;
      float z, w0, w1, w2;
      if (cover_test(s.fv[0][j], s.fv[1][j], s.fv[2][j], s.fv[3][j], s.fv[4][j],
                     s.fv[5][j], s.fv[6][j], s.fv[7][j], s.fv[8][j], s.fv[9][j],
                     bf, static_cast<float>(px0 + pxl), static_cast<float>(py0 + pyl),
                     z, w0, w1, w2)) {
        if (z >= FLT_MIN) {
          const unsigned int lowbits =
              slot_key ? (static_cast<unsigned int>(s.fid[j]) << 8) | static_cast<unsigned int>(j)
                       : static_cast<unsigned int>(s.fid[j]);
          const unsigned long long key =
              (static_cast<unsigned long long>(__float_as_uint(z)) << 32) | lowbits;
          atomicMin(&s.key[pix], key);
        } else {
          atomicOr(&s.flag[pix >> 5], 1u << (pix & 31));
          s.any_flag = 1;
        }
      }
    };
    if (big <= kScatterMaxPairs) {
      // small faces: every face writes its pairs (slot | tile pixel << 8)
      // into an LDS list, then every lane takes one pair
      for (int r0 = 0; r0 < total; r0 += 2 * NT) {
        const int lo = max(pre, r0), hi = min(pre + c, r0 + 2 * NT);
        if (lo < hi) {
          const int bx = s.box[t];
          const int nx = (bx >> 16) & 0xff, x0l = bx & 0xff;
          int ry = (lo - pre) / nx;
          int rx = (lo - pre) - ry * nx;
          int pix = (((bx >> 8) & 0xff) + ry) * TW + x0l + rx;
          for (int q = lo; q < hi; q++) {
            s.pairs[q - r0] = t | (pix << 8);
            if (++rx == nx) {
              rx = 0;
              pix += TW - nx + 1;
            } else {
              pix++;
            }
          }
        }
        __syncthreads();
        const int np = min(total - r0, 2 * NT);
        for (int pp = t; pp < np; pp += NT) {
          const int pr = s.pairs[pp];
          test_pair(pr & 0xff, pr >> 8);
        }
        __syncthreads();
      }
    } else {
      // some face covers many pixels: binary search of the pair's face over
      // the prefix (balanced whatever the face sizes)
      s.pairs[t] = pre;
      if (t == NT - 1) s.pairs[NT] = total;
      __syncthreads();
      for (int p = t; p < total; p += NT) {
        int jl = 0, jh = NT - 1;
        while (jl < jh) {
          const int mid = (jl + jh + 1) >> 1;
          if (s.pairs[mid] <= p) jl = mid; else jh = mid - 1;
        }
        const int r = p - s.pairs[jl];
        const int bx = s.box[jl];
        const int nx = (bx >> 16) & 0xff;
        const int qy = r / nx;
        test_pair(jl, (((bx >> 8) & 0xff) + qy) * TW + (bx & 0xff) + (r - qy * nx));
      }
      __syncthreads();
    }
    const unsigned long long k = s.key[t];
    if (more) {
      // depth bits of this pixel: all ones when it has no key yet, 0 outside
      // the image (no face reaches those pixels)
      uint32_t db = (gx < v.W && gy < v.H) ? static_cast<uint32_t>(k >> 32) : 0u;
      db = wave_max_u32(db);
      if ((t & 63) == 0) s.wmax[t >> 6] = db;
    }
    if (slot_key && k != seen) {
      // the pixel's winner changed: evaluate its coverage from the staged
      // record with the pair test's arithmetic (same bits), so the resolve
      // needs neither the record nor a second test
      const int sl = static_cast<int>(k & 0xffull);
      wbf = has_fn && sum3(s.fv[10][sl] * r0, s.fv[11][sl] * r1, s.fv[12][sl] * r2) > 0.0f;
      cover_test(s.fv[0][sl], s.fv[1][sl], s.fv[2][sl], s.fv[3][sl], s.fv[4][sl], s.fv[5][sl],
                 s.fv[6][sl], s.fv[7][sl], s.fv[8][sl], s.fv[9][sl], wbf,
                 static_cast<float>(gx), static_cast<float>(gy), wz, ww0, ww1, ww2);
      seen = k;
    }
    if (more) {
      // the next chunk's staging overwrites the records read above
      __syncthreads();
      uint32_t m = s.wmax[0];
      for (int w = 1; w < NT / 64; w++) m = max(m, s.wmax[w]);
      dmax = __uint_as_float(m);
    }
  }

  // ---- resolve (rasterizer.cc:243-301) for this thread's pixel; the
  // winner's depth, weights and back-face flag are in registers (slot keys)
  // or computed again below
  int fid = -1;
  bool have = false;
  if (s.any_flag && ((s.flag[t >> 5] >> (t & 31)) & 1u)) {
    sequential_pixel(v, mesh, bin, nA, bins.cap, n, gx, gy, r0, r1, r2, fid);
  } else if (seen != ~0ull) {
    fid = static_cast<int>(static_cast<unsigned int>(seen & 0xffffffffull) >> 8);
    have = true;
  } else if (!slot_key) {
    const unsigned long long key = s.key[t];
    if (key != ~0ull) fid = static_cast<int>(key & 0xffffffffull);
  }
  __syncthreads();  // LDS is reused by the next tile of a persistent block
  if (gx >= v.W || gy >= v.H) return;

  float z = 0.0f, w0 = 0.0f, w1 = 0.0f, w2 = 0.0f;
  bool bf = false;
  if (fid >= 0) {
    if (have) {
      z = wz;
      w0 = ww0;
      w1 = ww1;
      w2 = ww2;
      bf = wbf;
    } else {
      float P[9], N[3];
      load_face_rec(mesh.soup, fid, P, N);
      const float4 a = project_point(v, P[0], P[1], P[2]);
      const float4 b = project_point(v, P[3], P[4], P[5]);
      const float4 c = project_point(v, P[6], P[7], P[8]);
      const float area = edge_fn(a.x, a.y, b.x, b.y, c.x, c.y);
      bf = mesh.fnormals != nullptr && sum3(N[0] * r0, N[1] * r1, N[2] * r2) > 0.0f;
      cover_test(a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, area, bf,
                 static_cast<float>(gx), static_cast<float>(gy), z, w0, w1, w2);
    }
    if (opt.culling && bf) {
      fid = -1;
      z = 0.0f;
    }
  }
  const int64_t pidx = int64_t(gy) * v.W + gx;

  if (MODE == 0) {
    uint8_t col[3] = {0, 0, 0};
    float nrm[3] = {0.0f, 0.0f, 0.0f};
    uint8_t msk = 0;
    if (fid > 0) {
      msk = 255;
      float n0, n1, n2;
      FaceShade fs;
      face_shade(mesh, fid, opt.shading_normal != 0, v.color && opt.diffuse_color == 1, fs);
      if (opt.shading_normal == 0) {
        const float* fn = mesh.fnormals + int64_t(fid) * 3;
        n0 = fn[0]; n1 = fn[1]; n2 = fn[2];
      } else {
        const float* na = fs.na;
        const float* nb = fs.nb;
        const float* nc = fs.nc;
        n0 = w0 * na[0] + w1 * nb[0] + w2 * nc[0];
        n1 = w0 * na[1] + w1 * nb[1] + w2 * nc[1];
        n2 = w0 * na[2] + w1 * nb[2] + w2 * nc[2];
      }
      if (v.normal) {
        nrm[0] = sum3(v.w2c_R[0] * n0, v.w2c_R[1] * n1, v.w2c_R[2] * n2);
        nrm[1] = sum3(v.w2c_R[3] * n0, v.w2c_R[4] * n1, v.w2c_R[5] * n2);
        nrm[2] = sum3(v.w2c_R[6] * n0, v.w2c_R[7] * n1, v.w2c_R[8] * n2);
      }
      if (v.color) shade_u8(mesh, opt, fid, fs, w1, w2, r0, r1, r2, n0, n1, n2, col);
      if (v.bary || v.uvo || v.pos_wld || v.pos_cam || v.nrm_wld || v.geo_id) {
        // P is passed with a flag, not as a pointer that may be null: a select
        // between pointers would keep the array out of registers
        float P[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, N[3];
        const bool pos = v.pos_wld || v.pos_cam;
        if (pos) load_face_rec(mesh.soup, fid, P, N);
        write_gbuffer<S>(v, pidx, true, w1, w2, pos, P,
                         v.uvo && mesh.uv && mesh.uvidx ? &mesh : nullptr, n0, n1, n2, fid);
      }
    } else {
      write_gbuffer<S>(v, pidx, false, 0.0f, 0.0f, false, nullptr, nullptr, 0.0f, 0.0f, 0.0f);
    }
    if (v.color) put3<S>(v.color, pidx, col[0], col[1], col[2]);
    if (v.depth) put<S>(v.depth, pidx, z);
    if (v.normal) put3<S>(v.normal, pidx, nrm[0], nrm[1], nrm[2]);
    if (v.mask) put<S>(v.mask, pidx, msk);
    if (v.face_id) put<S>(v.face_id, pidx, fid);
  } else {
    // differentiable forward: image = colour term * texture term (float64)
    float img[3] = {0.0f, 0.0f, 0.0f};
    if (fid > 0) {
      DiffEval e;
      diff_eval(v, mesh, fid, gx, gy, mesh.faces + int64_t(fid) * 3, e);
      for (int k = 0; k < 3; k++) img[k] = static_cast<float>(e.c[k] * e.t[k]);
    }
    put3<S>(v.image, pidx, img[0], img[1], img[2]);
    if (v.face_id) put<S>(v.face_id, pidx, fid);
  }
}

template <int TW, int TH, int MODE>
// 7 waves per SIMD for the rendering kernels (72 VGPRs, no scratch; 7 blocks
// of 22.9 KB fill the 160 KB of LDS), 6 for the differentiable forward, whose
// extra saved values would spill at 72: the kernel is latency-bound
// (dependent gathers, LDS round trips, barriers), so occupancy pays
// (64-view batch: 0.90 -> 0.87 ms at 7).
__global__ __launch_bounds__(TW * TH, MODE == 0 ? 7 : 6) void k_raster(
    const DevView* __restrict__ views, DevMesh mesh,
    DevBins bins, DevOptions opt,
    int persistent) {
  __shared__ RasterLds<TW * TH> s;
  if (!persistent) {
    const int gtile = xcd_remap(blockIdx.x, gridDim.x);
    raster_tile<TW, TH, MODE>(s, gtile, reinterpret_cast<const int4*>(bins.counts)[gtile], views,
                              mesh, bins, opt);
    return;
  }
  // persistent: pull runs of kQueueRun consecutive (view, tile) items until
  // every queue is drained. Run r belongs to queue r % kQueues; a block starts
  // at queue blockIdx.x % kQueues and, once that one is drained, takes the
  // next queue's runs. Each head only grows and every run below n_runs is
  // handed out once, so a block that finds all kQueues heads past their
  // runs (one atomic each, at most kQueues in a row) leaves with every run
  // claimed. The next run is claimed when the last tile of the current one
  // starts. The barrier before reading s.item publishes thread 0's write (an
  // empty tile returns from raster_tile without a barrier); the one after
  // keeps thread 0 from overwriting it before every thread has read it.
  // `item` is the same in every thread, so every thread leaves the loop
  // together. The counter records of a run's tiles are read together into
  // LDS, one round trip per run (most tiles of a batch are empty and need
  // nothing else).
  constexpr int kQueueRun = 2;
  static_assert(kQueueRun <= 4, "RasterLds::runseg holds one run");
  const int n_runs = (bins.total_tiles + kQueueRun - 1) / kQueueRun;
  int queue = static_cast<int>(blockIdx.x % kQueues);  // used by thread 0
  auto claim = [&]() {
    for (int k = 0; k < kQueues; k++) {
      const int r = atomicAdd(&bins.status[kQueueStride * (queue + 1)], 1);
      if (r < (n_runs - queue + kQueues - 1) / kQueues) return (queue + kQueues * r) * kQueueRun;
      queue = (queue + 1) % kQueues;
    }
    return bins.total_tiles;
  };
  if (threadIdx.x == 0) s.item = claim();
  __syncthreads();
  int item = s.item;
  __syncthreads();
  while (item < bins.total_tiles) {
    const int end = min(item + kQueueRun, bins.total_tiles);
    if (threadIdx.x < end - item)
      s.runseg[threadIdx.x] = reinterpret_cast<const int4*>(bins.counts)[item + threadIdx.x];
    __syncthreads();
    for (int it = item; it < end; it++) {
      if (it == end - 1 && threadIdx.x == 0) s.item = claim();
      raster_tile<TW, TH, MODE>(s, it, s.runseg[it - item], views, mesh, bins, opt);
    }
    __syncthreads();
    item = s.item;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Differentiable backward pass: dL/d(vertices, vertex colours, texture) at
// fixed visibility. The covered pixels of all views are sorted by face id
// into (face, view, pixel) items with a counting sort:
//   k_bwd_count    one thread per pixel: per-face pixel counts
//   k_scan_*       exclusive scan of the counts (face offsets, item total)
//   k_bwd_scatter  one thread per pixel: item at offset + (count - 1), the
//                  counts decrement back to zero
//   k_bwd_items    one thread per item: shading derivatives in float64;
//                  contributions of contiguous items of one face (all its
//                  pixels in every view) are summed across the wavefront,
//                  then one atomic per vertex component and run
// Grouping by face turns a face's gradient into a handful of atomics instead
// of one set per pixel.

__global__ void k_zero_i32(int32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// face 0 is never shaded (rasterizer.cc:250), so it has no gradient
// Runs of equal face ids among a wavefront's consecutive pixels: `head` for
// the first lane of a run, `first` = that lane, `len` = the run's length (in
// the head lane). Lanes without a face form runs of their own.
struct FaceRun {
  bool head;
  int first, len;
};

UGU_DEV FaceRun face_run(int key) {
  const int lane = __lane_id();
  const int prev = wave_prev(key, key);
  FaceRun r;
  r.head = lane == 0 || prev != key;
  const uint64_t heads = __ballot(r.head);
  const uint64_t upto = heads & ((2ull << lane) - 1ull);
  r.first = 63 - __clzll(static_cast<long long>(upto));
  const uint64_t after = heads & ~((2ull << lane) - 1ull);
  r.len = (after ? __ffsll(static_cast<unsigned long long>(after)) - 1 : 64) - lane;
  return r;
}

// One thread per pixel; one atomic per run of equal face ids.
__global__ __launch_bounds__(256) void k_bwd_count(const DevView* __restrict__ views,
                                                   int32_t F, int32_t* __restrict__ cnt) {
  const DevView& v = views[blockIdx.y];
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (int64_t(blockIdx.x) * blockDim.x >= int64_t(v.W) * v.H) return;  // whole block idle
  const int fid = p < v.W * v.H ? v.face_id[p] : 0;
  const bool valid = fid > 0 && fid < F;
  const FaceRun r = face_run(valid ? fid : -1 - __lane_id());
  if (valid && r.head) atomicAdd(&cnt[fid], r.len);
}

// One thread per pixel: the run's head takes len slots from the face's
// counter (which counts back down to zero); lane j of the run writes slot
// base - 1 - j of the face's range.
__global__ __launch_bounds__(256) void k_bwd_scatter(const DevView* __restrict__ views,
                                                     int32_t F, const int32_t* __restrict__ off,
                                                     int32_t* __restrict__ cnt,
                                                     int4* __restrict__ items) {
  const DevView& v = views[blockIdx.y];
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (int64_t(blockIdx.x) * blockDim.x >= int64_t(v.W) * v.H) return;  // whole block idle
  const int fid = p < v.W * v.H ? v.face_id[p] : 0;
  const bool valid = fid > 0 && fid < F;
  const FaceRun r = face_run(valid ? fid : -1 - __lane_id());
  int base = 0;
  if (valid && r.head) base = atomicAdd(&cnt[fid], -r.len);
  base = __shfl(base, r.first);
  if (valid) items[off[fid] + base - 1 - (__lane_id() - r.first)] = make_int4(fid, blockIdx.y, p, 0);
}

// Exclusive scan of n counters (n + 1 outputs, out[n] = total) in three
// kernels: per-block scans of 4096 values, a scan of the block sums, the add.
constexpr int kScanBlock = 4096;

__global__ __launch_bounds__(1024) void k_scan_blocks(const int32_t* __restrict__ in, int n,
                                                      int32_t* __restrict__ out,
                                                      int32_t* __restrict__ sums) {
  __shared__ int32_t wsum[16];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int64_t b0 = int64_t(blockIdx.x) * kScanBlock + t * 4;
  int v[4], tot = 0;
  for (int k = 0; k < 4; k++) {
    v[k] = b0 + k < n ? in[b0 + k] : 0;
    tot += v[k];
  }
  const int x = wave_incl_sum(tot);
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  int base = 0, all = 0;
  for (int w = 0; w < 16; w++) {
    if (w < wid) base += wsum[w];
    all += wsum[w];
  }
  base += x - tot;
  for (int k = 0; k < 4; k++) {
    if (b0 + k < n) out[b0 + k] = base;
    base += v[k];
  }
  if (t == 0) sums[blockIdx.x] = all;
}

// one block: exclusive scan of the block sums in place, total in sums[nb]
__global__ __launch_bounds__(1024) void k_scan_sums(int32_t* __restrict__ sums, int nb) {
  __shared__ int32_t wsum[16];
  __shared__ int32_t carry;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  if (t == 0) carry = 0;
  __syncthreads();
  for (int c0 = 0; c0 < nb; c0 += 1024) {
    const int val = c0 + t < nb ? sums[c0 + t] : 0;
    const int x = wave_incl_sum(val);
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    int base = carry, all = 0;
    for (int w = 0; w < 16; w++) {
      if (w < wid) base += wsum[w];
      all += wsum[w];
    }
    if (c0 + t < nb) sums[c0 + t] = base + x - val;
    __syncthreads();
    if (t == 0) carry += all;
    __syncthreads();
  }
  if (t == 0) sums[nb] = carry;
}

__global__ void k_scan_add(int32_t* __restrict__ out, int n, const int32_t* __restrict__ sums,
                           int nb) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] += sums[i / kScanBlock];
  if (i == 0) out[n] = sums[nb];
}


// Texel gradients of a block's items are first summed in an LDS hash table
// (open addressing on the texel index, linear probing); a block's items cover
// a few faces in every view that sees them, so many items hit the same texels.
// Texels that find no free slot go straight to global atomics.
constexpr int kTexSlots = 2048;
constexpr int kTexProbes = 8;

UGU_DEV int tex_slot(int key) {
  return static_cast<int>((static_cast<uint32_t>(key) * 2654435761u) >> 21);  // 11 bits
}

// Items per thread of k_bwd_items: a block handles kBwdRounds x 256
// consecutive items (round r: items r*256 .. r*256+255 of the block's range,
// so each wavefront sees consecutive items) and merges the texel gradients
// of all of them in its LDS table before the global atomics.
constexpr int kBwdRounds = 4;

// Geometry of the image model at one pixel of a face, in float64 (the values
// diff_eval computes): screen positions, inverse camera depths and camera x, y
// of the corners, the edge functions' quotients q_i = e_i / (A Z_i), 1 / A,
// 1 / (q0 + q1 + q2) and the weights b1, b2 of the second and third corner.
struct DiffGeo {
  double X[3], Y[3], iZ[3], Qx[3], Qy[3], q[3], iA, is, b1, b2;
};

UGU_DEV void diff_geo(const DevView& v, const float* __restrict__ verts, const int32_t* vid,
                      int px, int py, DiffGeo& e) {
  for (int i = 0; i < 3; i++) {
    const float* P = verts + int64_t(vid[i]) * 3;
    double Q[3];
    for (int r = 0; r < 3; r++)
      Q[r] = double(v.w2c_R[r * 3 + 0]) * P[0] + double(v.w2c_R[r * 3 + 1]) * P[1] +
             double(v.w2c_R[r * 3 + 2]) * P[2] + double(v.w2c_t[r]);
    e.Qx[i] = Q[0];
    e.Qy[i] = Q[1];
    e.iZ[i] = 1.0 / Q[2];
    e.X[i] = double(v.fx) * Q[0] * e.iZ[i] + double(v.cx);
    e.Y[i] = double(v.fy) * Q[1] * e.iZ[i] + double(v.cy);
  }
  auto E = [](double ax, double ay, double bx, double by, double cx, double cy) {
    return (cx - ax) * (by - ay) - (cy - ay) * (bx - ax);
  };
  const double x = px, y = py;
  const double A = E(e.X[0], e.Y[0], e.X[1], e.Y[1], e.X[2], e.Y[2]);
  e.iA = 1.0 / A;
  e.q[0] = E(e.X[1], e.Y[1], e.X[2], e.Y[2], x, y) * e.iA * e.iZ[0];
  e.q[1] = E(e.X[2], e.Y[2], e.X[0], e.Y[0], x, y) * e.iA * e.iZ[1];
  e.q[2] = E(e.X[0], e.Y[0], e.X[1], e.Y[1], x, y) * e.iA * e.iZ[2];
  e.is = 1.0 / (e.q[0] + e.q[1] + e.q[2]);
  e.b1 = e.q[1] * e.is;
  e.b2 = e.q[2] * e.is;
}

// One thread per item. The work is ordered so that few float64 values are
// live at a time (the kernel is bound by the latency of its gathers, so
// registers buy occupancy): the weights, then the colour and texture terms
// channel by channel into the colour and texel gradients and dL/du, dL/dv,
// whose atomics are issued next; the geometry is then evaluated again for the
// position gradients.
__global__ __launch_bounds__(256) void k_bwd_items(
    const DevView* __restrict__ views, DevMesh mesh, const int4* __restrict__ items,
    const int32_t* __restrict__ cursor, DevGrads g) {
  static_assert(kTexSlots == 2048, "tex_slot returns 11 bits");
  __shared__ int32_t hkey[kTexSlots];
  __shared__ float hval[3][kTexSlots];
  const int n_items = *cursor;
  const int64_t first = int64_t(blockIdx.x) * blockDim.x * kBwdRounds;
  if (first >= n_items) return;  // whole block idle
  const bool tex = g.texture && mesh.ftex;  // uniform over the block
  if (tex) {
    for (int j = threadIdx.x; j < kTexSlots; j += blockDim.x) {
      hkey[j] = -1;
      hval[0][j] = hval[1][j] = hval[2][j] = 0.0f;
    }
    __syncthreads();
  }
  for (int round = 0; round < kBwdRounds; round++) {
  const int64_t i = first + int64_t(round) * blockDim.x + threadIdx.x;
  const bool valid = i < n_items;
  int4 item = make_int4(-1, 0, 0, 0);
  if (valid) item = items[i];
  const int fid = item.x;
  int vid[3] = {0, 0, 0};
  double gu = 0.0, gv = 0.0, b1 = 0.0, b2 = 0.0;
  int px = 0, py = 0;
  const DevView* vp = views;

  // ---- colour and texture terms: dL/dC, dL/dT and dL/du, dL/dv
  // (dL/dc = G t, dL/dt = G c; u = b1, v = b2, b0 = 1 - u - v)
  {
    float gcol[9], gtex[12];
    for (int k = 0; k < 9; k++) gcol[k] = 0.0f;
    for (int k = 0; k < 12; k++) gtex[k] = 0.0f;
    int tx0 = 0, ty0 = 0, tx1 = 0, ty1 = 0;
    if (valid) {
      vp = views + item.y;
      const DevView& v = *vp;
      const int p = item.z;
      px = p % v.W;
      py = p / v.W;
      const int32_t* fi = mesh.faces + int64_t(fid) * 3;
      vid[0] = fi[0]; vid[1] = fi[1]; vid[2] = fi[2];
      const float* gp = v.grad + int64_t(p) * 3;
      const double G[3] = {gp[0], gp[1], gp[2]};
      {
        DiffGeo e;
        diff_geo(v, mesh.verts, vid, px, py, e);
        b1 = e.b1;
        b2 = e.b2;
      }
      const double b0 = 1.0 - b1 - b2;
      const float *C0 = nullptr, *C1 = nullptr, *C2 = nullptr;
      if (mesh.vcolors) {
        C0 = mesh.vcolors + int64_t(vid[0]) * 3;
        C1 = mesh.vcolors + int64_t(vid[1]) * 3;
        C2 = mesh.vcolors + int64_t(vid[2]) * 3;
      }
      double lu = 0.0, lv = 0.0, dlu_du = 0.0, dlu_dv = 0.0, dlv_du = 0.0, dlv_dv = 0.0;
      const float *p00 = nullptr, *p01 = nullptr, *p10 = nullptr, *p11 = nullptr;
      if (mesh.ftex) {
        const int32_t* ui = mesh.uvidx + int64_t(fid) * 3;
        const float* a = mesh.uv + int64_t(ui[0]) * 2;
        const float* b = mesh.uv + int64_t(ui[1]) * 2;
        const float* c = mesh.uv + int64_t(ui[2]) * 2;
        const double iu = b0 * a[0] + b1 * b[0] + b2 * c[0];
        const double iv = b0 * a[1] + b1 * b[1] + b2 * c[1];
        const int tw = mesh.ftex_w, th = mesh.ftex_h;
        const double fxp = iu * (tw - 1), fyp = (1.0 - iv) * (th - 1);
        const double flx = floor(fxp), fly = floor(fyp);
        lu = fxp - flx;
        lv = fyp - fly;
        tx0 = clamp_texel(flx, tw - 1);
        ty0 = clamp_texel(fly, th - 1);
        tx1 = clamp_texel(flx + 1.0, tw - 1);
        ty1 = clamp_texel(fly + 1.0, th - 1);
        p00 = mesh.ftex + (int64_t(ty0) * tw + tx0) * 3;
        p01 = mesh.ftex + (int64_t(ty0) * tw + tx1) * 3;
        p10 = mesh.ftex + (int64_t(ty1) * tw + tx0) * 3;
        p11 = mesh.ftex + (int64_t(ty1) * tw + tx1) * 3;
        const double sx = tw - 1, sy = -(th - 1.0);
        dlu_du = sx * (double(b[0]) - a[0]);
        dlu_dv = sx * (double(c[0]) - a[0]);
        dlv_du = sy * (double(b[1]) - a[1]);
        dlv_dv = sy * (double(c[1]) - a[1]);
      }
      const double wq[4] = {(1 - lu) * (1 - lv), lu * (1 - lv), (1 - lu) * lv, lu * lv};
      for (int k = 0; k < 3; k++) {
        double cc = 1.0, dcdu = 0.0, dcdv = 0.0;
        if (C0) {
          cc = b0 * C0[k] + b1 * C1[k] + b2 * C2[k];
          dcdu = double(C1[k]) - C0[k];
          dcdv = double(C2[k]) - C0[k];
        }
        double t = 1.0, dtdu = 0.0, dtdv = 0.0;
        if (p00) {
          t = wq[0] * p00[k] + wq[1] * p01[k] + wq[2] * p10[k] + wq[3] * p11[k];
          const double da = (1 - lv) * (double(p01[k]) - p00[k]) + lv * (double(p11[k]) - p10[k]);
          const double db = (1 - lu) * (double(p10[k]) - p00[k]) + lu * (double(p11[k]) - p01[k]);
          dtdu = da * dlu_du + db * dlv_du;
          dtdv = da * dlu_dv + db * dlv_dv;
        }
        const double gc = G[k] * t, gt = G[k] * cc;
        gcol[0 * 3 + k] = static_cast<float>(gc * b0);
        gcol[1 * 3 + k] = static_cast<float>(gc * b1);
        gcol[2 * 3 + k] = static_cast<float>(gc * b2);
        for (int q = 0; q < 4; q++) gtex[q * 3 + k] = static_cast<float>(gt * wq[q]);
        gu += gc * dcdu + gt * dtdu;
        gv += gc * dcdv + gt * dtdv;
      }
    }

    // wavefront pre-reduction over runs of equal face id, then atomics
    // from each run's last lane
    const int lane = __lane_id();
    const int prev_fid = wave_prev(fid, INT_MIN);
    const int next_fid = wave_next(fid, INT_MIN);
    const bool head = valid && (lane == 0 || prev_fid != fid);
    const bool tail = valid && (lane == 63 || next_fid != fid);
    const uint64_t heads = __ballot(head);
    const uint64_t vmask = __ballot(valid);
    const bool reduce = heads != vmask;  // some run longer than one lane
    const int run = valid ? __popcll(heads & ((2ull << lane) - 1ull)) : -1 - lane;
    if (g.colors && mesh.vcolors) {
      for (int k = 0; k < 9; k++) {
        const float val = reduce ? run_sum(gcol[k], run) : gcol[k];
        if (tail) unsafeAtomicAdd(&g.colors[int64_t(vid[k / 3]) * 3 + (k % 3)], val);
      }
    }
    if (tex && valid) {
      const int tw = mesh.ftex_w;
      const int tidx[4] = {ty0 * tw + tx0, ty0 * tw + tx1, ty1 * tw + tx0, ty1 * tw + tx1};
      for (int q = 0; q < 4; q++) {
        const int key = tidx[q];
        int h = tex_slot(key);
        bool done = false;
        for (int probe = 0; probe < kTexProbes && !done; probe++) {
          const int old = atomicCAS(&hkey[h], -1, key);
          if (old == -1 || old == key) {
            for (int k = 0; k < 3; k++) atomicAdd(&hval[k][h], gtex[q * 3 + k]);
            done = true;
          }
          h = (h + 1) & (kTexSlots - 1);
        }
        if (!done)
          for (int k = 0; k < 3; k++)
            unsafeAtomicAdd(&g.texture[int64_t(key) * 3 + k], gtex[q * 3 + k]);
      }
    }
  }

  // ---- positions (the geometry again): b_i = q_i / s
  if (g.verts) {
    float gpos[9];
    for (int k = 0; k < 9; k++) gpos[k] = 0.0f;
    if (valid) {
      const DevView& v = *vp;
      DiffGeo e;
      diff_geo(v, mesh.verts, vid, px, py, e);
      // dL/dq_j = (dL/db_j - (gu b1 + gv b2)) / s
      const double S = gu * b1 + gv * b2;
      const double gq[3] = {-S * e.is, (gu - S) * e.is, (gv - S) * e.is};
      // q_j = e_j / (A Z_j)
      double ge[3], gZ[3], gA = 0.0;
      for (int j = 0; j < 3; j++) {
        ge[j] = gq[j] * e.iA * e.iZ[j];
        gZ[j] = -gq[j] * e.q[j] * e.iZ[j];
        gA -= gq[j] * e.q[j] * e.iA;
      }
      // E(a,b,c): dE/da = (c.y-b.y, b.x-c.x), dE/db = (a.y-c.y, c.x-a.x),
      //           dE/dc = (b.y-a.y, a.x-b.x)
      // e0 = E(v1, v2, p), e1 = E(v2, v0, p), e2 = E(v0, v1, p), A = E(v0, v1, v2)
      double gX[3] = {0, 0, 0}, gY[3] = {0, 0, 0};
      const double x = px, y = py;
      const int ea[3] = {1, 2, 0}, eb[3] = {2, 0, 1};
      for (int j = 0; j < 3; j++) {
        const int ia = ea[j], ib = eb[j];
        gX[ia] += ge[j] * (y - e.Y[ib]);
        gY[ia] += ge[j] * (e.X[ib] - x);
        gX[ib] += ge[j] * (e.Y[ia] - y);
        gY[ib] += ge[j] * (x - e.X[ia]);
      }
      gX[0] += gA * (e.Y[2] - e.Y[1]);
      gY[0] += gA * (e.X[1] - e.X[2]);
      gX[1] += gA * (e.Y[0] - e.Y[2]);
      gY[1] += gA * (e.X[2] - e.X[0]);
      gX[2] += gA * (e.Y[1] - e.Y[0]);
      gY[2] += gA * (e.X[0] - e.X[1]);
      // X = fx Qx / Qz + cx, Y = fy Qy / Qz + cy, Z = Qz; Q = R P + t
      for (int j = 0; j < 3; j++) {
        const double iz = e.iZ[j];
        const double gQx = gX[j] * v.fx * iz;
        const double gQy = gY[j] * v.fy * iz;
        const double gQz = gZ[j] - gX[j] * v.fx * e.Qx[j] * iz * iz -
                           gY[j] * v.fy * e.Qy[j] * iz * iz;
        for (int cdim = 0; cdim < 3; cdim++)
          gpos[j * 3 + cdim] = static_cast<float>(double(v.w2c_R[0 * 3 + cdim]) * gQx +
                                                  double(v.w2c_R[1 * 3 + cdim]) * gQy +
                                                  double(v.w2c_R[2 * 3 + cdim]) * gQz);
      }
    }
    const int lane = __lane_id();
    const int prev_fid = wave_prev(fid, INT_MIN);
    const int next_fid = wave_next(fid, INT_MIN);
    const bool head = valid && (lane == 0 || prev_fid != fid);
    const bool tail = valid && (lane == 63 || next_fid != fid);
    const uint64_t heads = __ballot(head);
    const uint64_t vmask = __ballot(valid);
    const bool reduce = heads != vmask;
    const int run = valid ? __popcll(heads & ((2ull << lane) - 1ull)) : -1 - lane;
    for (int k = 0; k < 9; k++) {
      const float val = reduce ? run_sum(gpos[k], run) : gpos[k];
      if (tail) unsafeAtomicAdd(&g.verts[int64_t(vid[k / 3]) * 3 + (k % 3)], val);
    }
  }
  }  // rounds
  if (tex) {
    __syncthreads();
    for (int j = threadIdx.x; j < kTexSlots; j += blockDim.x) {
      const int key = hkey[j];
      if (key >= 0)
        for (int k = 0; k < 3; k++) unsafeAtomicAdd(&g.texture[int64_t(key) * 3 + k], hval[k][j]);
    }
  }
}

// ---------------------------------------------------------------------------
// RenderW depth conversion: uint16(1.0 * (double)d), truncating as x86 does.
__global__ void k_depth_to_u16(const float* __restrict__ d, int64_t n,
                               uint16_t* __restrict__ out) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 1.0 * static_cast<double>(d[i]);
  int32_t iv = (s >= -2147483648.0 && s < 2147483648.0) ? static_cast<int32_t>(s) : INT32_MIN;
  out[i] = static_cast<uint16_t>(static_cast<uint32_t>(iv) & 0xffffu);
}

// ---------------------------------------------------------------------------
// host-side launchers (called from ugu_hip.cpp)

int launch_clear(hipStream_t st, int32_t* counts, int32_t n, int32_t* status,
                 const int32_t* tile_base, int32_t n_views) {
  int m = n > 4 ? n : 4;
  hipLaunchKernelGGL(k_clear, dim3((m + 255) / 256), dim3(256), 0, st, counts, n, status,
                     tile_base, n_views);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_shade_rec(hipStream_t st, const DevMesh& m, float4* out) {
  if (m.F <= 0) return 0;
  hipLaunchKernelGGL(k_shade_rec, dim3((m.F + 255) / 256), dim3(256), 0, st, m, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_face_soup(hipStream_t st, const float* verts, const int32_t* faces,
                     const float* fnormals, int32_t F, int compute_normals, float4* soup,
                     float* fn_out) {
  if (F <= 0) return 0;
  hipLaunchKernelGGL(k_face_soup, dim3((F + 255) / 256), dim3(256), 0, st, verts, faces,
                     fnormals, F, compute_normals, soup, fn_out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_setup_bin(hipStream_t st, int tw, int th, const DevView* views,
                     int n_views, const DevMesh& mesh,
                     const DevBins& bins) {
  if (mesh.F <= 0) return 0;
  dim3 grid((mesh.F + 255) / 256, (n_views + kSetupViews - 1) / kSetupViews), block(256);
  if (tw == 8 && th == 8)
    hipLaunchKernelGGL((k_setup_bin<8, 8>), grid, block, 0, st, views, n_views, mesh, bins);
  else if (tw == 16 && th == 16)
    hipLaunchKernelGGL((k_setup_bin<16, 16>), grid, block, 0, st, views, n_views, mesh, bins);
  else if (tw == 32 && th == 8)
    hipLaunchKernelGGL((k_setup_bin<32, 8>), grid, block, 0, st, views, n_views, mesh, bins);
  else
    return -2;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int TW, int TH, int MODE>
static int launch_raster_t(hipStream_t st, const DevView* views,
                           const DevMesh& mesh,
                           const DevBins& bins, const DevOptions& opt,
                           int persistent_blocks) {
  if (bins.total_tiles <= 0) return 0;
  const int grid = persistent_blocks > 0 ? persistent_blocks : bins.total_tiles;
  hipLaunchKernelGGL((k_raster<TW, TH, MODE>), dim3(grid), dim3(TW * TH), 0, st,
                     views, mesh, bins, opt, persistent_blocks > 0 ? 1 : 0);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_raster(hipStream_t st, int tw, int th, int mode, const DevView* views,
                  const DevMesh& mesh, const DevBins& bins,
                  const DevOptions& opt, int persistent_blocks) {
#define UGU_RASTER_CASE(TWV, THV)                                                  \
  if (tw == TWV && th == THV)                                                      \
    return mode == 0 ? launch_raster_t<TWV, THV, 0>(st, views, mesh, bins,   \
                                                    opt, persistent_blocks)        \
                     : launch_raster_t<TWV, THV, 1>(st, views, mesh, bins,   \
                                                    opt, persistent_blocks);
  UGU_RASTER_CASE(8, 8)
  UGU_RASTER_CASE(16, 16)
  UGU_RASTER_CASE(32, 8)
#undef UGU_RASTER_CASE
  return -2;
}

int raster_blocks_per_cu(int tw, int th, int mode) {
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
#define UGU_OCC_CASE(TWV, THV)                                                       \
  if (tw == TWV && th == THV) {                                                      \
    if (mode == 0)                                                                   \
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_raster<TWV, THV, 0>,   \
                                                       TWV * THV, 0);                \
    else                                                                             \
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_raster<TWV, THV, 1>,   \
                                                       TWV * THV, 0);                \
  }
  UGU_OCC_CASE(8, 8)
  UGU_OCC_CASE(16, 16)
  UGU_OCC_CASE(32, 8)
#undef UGU_OCC_CASE
  return e == hipSuccess ? nb : 0;
}

// Scratch of the backward pass: cnt (F + 1), off (F + 1) and the block sums.
int64_t backward_scratch_ints(int32_t F) {
  return 2 * (int64_t(F) + 1) + (int64_t(F) + kScanBlock - 1) / kScanBlock + 1;
}

int launch_backward(hipStream_t st, const DevView* views, int n_views,
                    int max_pixels, int64_t total_pixels, const DevMesh& mesh,
                    const DevGrads& g, int4* items, int32_t* scratch, hipEvent_t before_items) {
  if (max_pixels <= 0 || total_pixels <= 0 || mesh.F <= 0) return 0;
  const int F = mesh.F;
  int32_t* cnt = scratch;
  int32_t* off = scratch + (F + 1);
  int32_t* sums = off + (F + 1);
  const int nb = (F + kScanBlock - 1) / kScanBlock;
  const dim3 pgrid((max_pixels + 255) / 256, n_views);
  hipLaunchKernelGGL(k_zero_i32, dim3((F + 256) / 256), dim3(256), 0, st, cnt, F + 1);
  hipLaunchKernelGGL(k_bwd_count, pgrid, dim3(256), 0, st, views, F, cnt);
  hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(1024), 0, st, cnt, F, off, sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, st, sums, nb);
  hipLaunchKernelGGL(k_scan_add, dim3((F + 255) / 256), dim3(256), 0, st, off, F, sums, nb);
  hipLaunchKernelGGL(k_bwd_scatter, pgrid, dim3(256), 0, st, views, F, off, cnt, items);
  if (before_items) (void)hipEventRecord(before_items, st);
  // at most one item per pixel; blocks past the item count (off[F]) exit at once
  hipLaunchKernelGGL(k_bwd_items,
                     dim3(static_cast<unsigned>((total_pixels + 256 * kBwdRounds - 1) /
                                                (256 * kBwdRounds))),
                     dim3(256), 0, st, views, mesh, items, off + F, g);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_depth_to_u16(hipStream_t st, const float* d, int64_t n, uint16_t* out) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_depth_to_u16, dim3((n + 255) / 256), dim3(256), 0, st, d, n, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace ugu_hip
