// Device helpers shared by the rasterizer and raytracer kernels: the
// reference's scalar conversions and reductions, the camera ray, texture
// reads and the pixel shaders. Every file that includes this one is compiled
// with -ffp-contract=off and correctly rounded fp32 division/sqrt.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>
#include <limits.h>

#include "raster_types.h"

#pragma clang fp contract(off)

namespace ugu_hip {

#define UGU_DEV __device__ __forceinline__

// Output pointers read from descriptors in memory are generic to the
// compiler, which then emits flat stores (counted against both the vector
// memory and the LDS wait counters). Casting to the global address space
// gives plain global stores.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
UGU_DEV gptr<T> to_global(T* p) {
  return (gptr<T>)p;
}

// Output planes are written once and never read back by the kernel that
// writes them. With STREAM they are stored non-temporally: streamed to memory
// instead of allocated in L2, where they would evict the records and texels
// the kernel reads again (64-view batch: rasterizer -6.5%). That pays when a
// wavefront's row segments are long; short ones (8-pixel tiles: 24 B of
// colour) turn into partial-line writes that L2 would otherwise have merged
// with the neighbouring tile's.
template <bool STREAM, typename T>
UGU_DEV void put(T* plane, int64_t i, T value) {
  if (STREAM)
    __builtin_nontemporal_store(value, to_global(plane) + i);
  else
    to_global(plane)[i] = value;
}

// Pixel i of a three-channel plane, its channels stored back to back (one
// 12-byte store for floats; stores to another plane in between would keep
// the compiler from merging them, as the planes may alias).
template <bool STREAM, typename T>
UGU_DEV void put3(T* plane, int64_t i, T c0, T c1, T c2) {
  put<STREAM>(plane, i * 3, c0);
  put<STREAM>(plane, i * 3 + 1, c1);
  put<STREAM>(plane, i * 3 + 2, c2);
}

// ---------------------------------------------------------------------------
// Wavefront scans and reductions through the DPP cross-lane paths: four
// shifts within each row of 16 lanes, then the row broadcasts of lanes 15
// and 31 -- register moves, where __shfl_* is an LDS round trip per step.
// Lanes whose source lies outside their row (or in a row the mask leaves
// out) take `id`, the operation's identity.

template <typename Op>
UGU_DEV int wave_scan_dpp(int x, int id, Op op) {
  x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x111, 0xf, 0xf, false));  // row_shr:1
  x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x112, 0xf, 0xf, false));  // row_shr:2
  x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x114, 0xf, 0xf, false));  // row_shr:4
  x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x118, 0xf, 0xf, false));  // row_shr:8
  x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x142, 0xa, 0xf, false));  // row_bcast:15
  x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return x;
}

// Inclusive prefix sum in lane order.
UGU_DEV int wave_incl_sum(int x) {
  return wave_scan_dpp(x, 0, [](int a, int b) { return a + b; });
}

// Inclusive running maximum of non-negative values.
UGU_DEV int wave_incl_max(int x) {
  return wave_scan_dpp(x, 0, [](int a, int b) { return max(a, b); });
}

// Sums `val` over contiguous runs of equal `run` ids (an inclusive scan
// segmented by run, through the DPP paths of wave_scan_dpp); the run's last
// lane ends up with the run total. A source lane outside the row or the
// row mask brings run INT_MIN, which no run has.
UGU_DEV float run_sum(float val, int run) {
  constexpr int kNone = INT_MIN;
#define UGU_RUN_STEP(ctrl, rows)                                                          \
  {                                                                                     \
    const int o = __builtin_amdgcn_update_dpp(0, __float_as_int(val), ctrl, rows, 0xf, false); \
    const int orun = __builtin_amdgcn_update_dpp(kNone, run, ctrl, rows, 0xf, false);   \
    if (orun == run) val += __int_as_float(o);                                          \
  }
  UGU_RUN_STEP(0x111, 0xf)  // row_shr:1
  UGU_RUN_STEP(0x112, 0xf)  // row_shr:2
  UGU_RUN_STEP(0x114, 0xf)  // row_shr:4
  UGU_RUN_STEP(0x118, 0xf)  // row_shr:8
  UGU_RUN_STEP(0x142, 0xa)  // row_bcast:15
  UGU_RUN_STEP(0x143, 0xc)  // row_bcast:31
#undef UGU_RUN_STEP
  return val;
}

// The previous / next lane's value (lane 0 / lane 63 get `edge`).
UGU_DEV int wave_prev(int v, int edge) {
  return __builtin_amdgcn_update_dpp(edge, v, 0x138, 0xf, 0xf, false);  // wave_shr:1
}
UGU_DEV int wave_next(int v, int edge) {
  return __builtin_amdgcn_update_dpp(edge, v, 0x130, 0xf, 0xf, false);  // wave_shl:1
}

// Maximum / minimum over the wavefront, the same in every lane.
UGU_DEV uint32_t wave_max_u32(uint32_t v) {
  const int r = wave_scan_dpp(static_cast<int>(v), 0, [](int a, int b) {
    return static_cast<int>(max(static_cast<uint32_t>(a), static_cast<uint32_t>(b)));
  });
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(r, 63));
}
UGU_DEV uint32_t wave_min_u32(uint32_t v) {
  const int r = wave_scan_dpp(static_cast<int>(v), -1, [](int a, int b) {
    return static_cast<int>(min(static_cast<uint32_t>(a), static_cast<uint32_t>(b)));
  });
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(r, 63));
}

// ---------------------------------------------------------------------------
// scalar helpers

// Eigen's unrolled reduction of a fixed 3-vector: x0 + (x1 + x2)
UGU_DEV float sum3(float a, float b, float c) { return a + (b + c); }

// EdgeFunction(a, b, c) = (c0-a0)(b1-a1) - (c1-a1)(b0-a0)  (rasterizer.cc:25-28)
UGU_DEV float edge_fn(float ax, float ay, float bx, float by, float cx, float cy) {
  return (cx - ax) * (by - ay) - (cy - ay) * (bx - ax);
}

// float -> int32 as x86-64 cvttss2si (what the reference's static_cast
// compiles to): truncation, INT32_MIN for NaN and out-of-range values.
UGU_DEV int32_t f2i_x86(float f) {
  return (f >= -2147483648.0f && f < 2147483648.0f) ? static_cast<int32_t>(f)
                                                    : INT32_MIN;
}
// saturating float -> int32 for the pixel bbox upper bound
UGU_DEV int32_t f2i_sat(float f) {
  if (f >= 2147483648.0f) return INT32_MAX;
  if (f >= -2147483648.0f) return static_cast<int32_t>(f);
  return INT32_MIN;
}
UGU_DEV uint8_t f2u8_x86(float f) {
  return static_cast<uint8_t>(static_cast<uint32_t>(f2i_x86(f)) & 0xffu);
}

// std::min({a,b,c}) / std::max({a,b,c}) including their NaN behaviour
UGU_DEV float min3_std(float a, float b, float c) {
  float m = a;
  if (b < m) m = b;
  if (c < m) m = c;
  return m;
}
UGU_DEV float max3_std(float a, float b, float c) {
  float m = a;
  if (m < b) m = b;
  if (m < c) m = c;
  return m;
}

// PinholeCamera::ray_w(float x, float y) (camera.h:438-448):
// d = ((x-cx)/fx, (y-cy)/fy, 1); d.normalize(); ray = c2w_R_f * d
UGU_DEV void ray_w_at(const DevView& v, float x, float y, float& r0, float& r1,
                      float& r2) {
  float d0 = (x - v.cx) / v.fx;
  float d1 = (y - v.cy) / v.fy;
  float d2 = 1.0f;
  float z = sum3(d0 * d0, d1 * d1, d2 * d2);
  if (z > 0.0f) {
    float s = sqrtf(z);
    d0 = d0 / s;
    d1 = d1 / s;
    d2 = d2 / s;
  }
  r0 = sum3(v.c2w_R[0] * d0, v.c2w_R[1] * d1, v.c2w_R[2] * d2);
  r1 = sum3(v.c2w_R[3] * d0, v.c2w_R[4] * d1, v.c2w_R[5] * d2);
  r2 = sum3(v.c2w_R[6] * d0, v.c2w_R[7] * d1, v.c2w_R[8] * d2);
}

// backface = face_normal . ray_w > 0 (rasterizer.cc:200)
UGU_DEV bool backface_of(const float* fn, int64_t fid, float r0, float r1, float r2) {
  if (fn == nullptr) return false;
  const float* n = fn + fid * 3;
  return sum3(n[0] * r0, n[1] * r1, n[2] * r2) > 0.0f;
}

// ImageBase::at on an RGB8 texture (image.h:416-427) with the linear index
// clamped to the image. Device textures are stored RGBA8 (alpha unused) so a
// texel is one aligned 32-bit load.
UGU_DEV uchar4 tex_at(const uint8_t* base, int w, int h, int row, int col) {
  int64_t lin = static_cast<int64_t>(row) * w + col;
  int64_t n = static_cast<int64_t>(w) * h;
  lin = lin < 0 ? 0 : (lin >= n ? n - 1 : lin);
  return reinterpret_cast<const uchar4*>(base)[lin];
}

// Shading inputs of one face: its corners' vertex normals (need_n) and its
// corners' uvs and material (need_uv); fields not asked for are left unset.
// With the mesh's shading records (one 64-byte record per face, built at
// PrepareMesh) they are the record's first 32 bytes (uvs, material) and last
// 48 bytes (normals); otherwise gathered through the index arrays
// (normal_indices -> normals, uv_indices -> uv, material_ids).
struct FaceShade {
  float na[3], nb[3], nc[3];
  float ua[2], ub[2], uc[2];
  int mat;
};

UGU_DEV void face_shade(const DevMesh& m, int fid, bool need_n, bool need_uv, FaceShade& s) {
  if (m.shade) {
    const float4* r = m.shade + int64_t(fid) * 4;
    auto uvs = [&](const float4& a, const float4& b) {
      s.ua[0] = a.x; s.ua[1] = a.y; s.ub[0] = a.z; s.ub[1] = a.w;
      s.uc[0] = b.x; s.uc[1] = b.y;
      s.mat = __float_as_int(b.z);
    };
    auto nrms = [&](const float4& b, const float4& c, const float4& d) {
      s.na[0] = b.w; s.na[1] = c.x; s.na[2] = c.y;
      s.nb[0] = c.z; s.nb[1] = c.w; s.nb[2] = d.x;
      s.nc[0] = d.y; s.nc[1] = d.z; s.nc[2] = d.w;
    };
    if (need_n && need_uv) {
      const float4 a = r[0], b = r[1], c = r[2], d = r[3];
      uvs(a, b);
      nrms(b, c, d);
    } else if (need_uv) {
      const float4 a = r[0], b = r[1];
      uvs(a, b);
    } else if (need_n) {
      const float4 b = r[1], c = r[2], d = r[3];
      nrms(b, c, d);
    }
    return;
  }
  if (need_n) {
    const int32_t* ni = m.nidx + int64_t(fid) * 3;
    const float* na = m.normals + int64_t(ni[0]) * 3;
    const float* nb = m.normals + int64_t(ni[1]) * 3;
    const float* nc = m.normals + int64_t(ni[2]) * 3;
    for (int k = 0; k < 3; k++) {
      s.na[k] = na[k];
      s.nb[k] = nb[k];
      s.nc[k] = nc[k];
    }
  }
  if (need_uv) {
    const int32_t* ui = m.uvidx + int64_t(fid) * 3;
    const float* a = m.uv + int64_t(ui[0]) * 2;
    const float* b = m.uv + int64_t(ui[1]) * 2;
    const float* c = m.uv + int64_t(ui[2]) * 2;
    s.ua[0] = a[0]; s.ua[1] = a[1];
    s.ub[0] = b[0]; s.ub[1] = b[1];
    s.uc[0] = c[0]; s.uc[1] = c[1];
    s.mat = m.matid[fid];
  }
}

// Colour of one shaded pixel of the reference path: colourizer then shader
// (pixel_shader.h:215-218 with the classes at :222-422). `fs` holds the face's
// uvs and material when the diffuse colour is the texture.
UGU_DEV void shade_u8(const DevMesh& m, const DevOptions& o, int fid, const FaceShade& fs,
                      float u, float vv, float r0, float r1, float r2, float n0, float n1,
                      float n2, uint8_t c[3]) {
  if (o.diffuse_color == 2) {
    // DiffuseVertexColorColorizer (pixel_shader.h:234-256)
    const int32_t* fi = m.faces + int64_t(fid) * 3;
    const float* c0 = m.vcolors + int64_t(fi[0]) * 3;
    const float* c1 = m.vcolors + int64_t(fi[1]) * 3;
    const float* c2 = m.vcolors + int64_t(fi[2]) * 3;
    const float w = 1.0f - u - vv;
    for (int k = 0; k < 3; k++) c[k] = f2u8_x86(w * c0[k] + u * c1[k] + vv * c2[k]);
  } else if (o.diffuse_color == 1) {
    const TexInfo t = m.n_tex == 1 ? m.tex0 : m.tex[fs.mat];
    const uint8_t* base = m.texdata + t.offset;
    const float* a = fs.ua;
    const float* b = fs.ub;
    const float* cc = fs.uc;
    const float w = 1.0f - u - vv;
    const float iu = w * a[0] + u * b[0] + vv * cc[0];
    const float iv = w * a[1] + u * b[1] + vv * cc[1];
    const float fx = iu * static_cast<float>(t.w - 1);
    const float fy = (1.0f - iv) * static_cast<float>(t.h - 1);
    if (o.interp == 0) {
      // DiffuseTextureNnColorizer (pixel_shader.h:260-298)
      const uchar4 dt = tex_at(base, t.w, t.h, f2i_x86(roundf(fy)), f2i_x86(roundf(fx)));
      c[0] = dt.x;
      c[1] = dt.y;
      c[2] = dt.z;
    } else {
      // DiffuseTextureBilinearColorizer (pixel_shader.h:302-359)
      const int x0 = f2i_x86(floorf(fx));
      const int y0 = f2i_x86(floorf(fy));
      const int x1 = x0 + 1, y1 = y0 + 1;
      const float lu = fx - static_cast<float>(x0);
      const float lv = fy - static_cast<float>(y0);
      const uchar4 q00 = tex_at(base, t.w, t.h, y0, x0);
      const uchar4 q01 = tex_at(base, t.w, t.h, y0, x1);
      const uchar4 q10 = tex_at(base, t.w, t.h, y1, x0);
      const uchar4 q11 = tex_at(base, t.w, t.h, y1, x1);
      const uint8_t p00[3] = {q00.x, q00.y, q00.z}, p01[3] = {q01.x, q01.y, q01.z};
      const uint8_t p10[3] = {q10.x, q10.y, q10.z}, p11[3] = {q11.x, q11.y, q11.z};
      for (int k = 0; k < 3; k++) {
        float ic = (1.0f - lu) * (1.0f - lv) * static_cast<float>(p00[k]) +
                   lu * (1.0f - lv) * static_cast<float>(p01[k]) +
                   (1.0f - lu) * lv * static_cast<float>(p10[k]) +
                   lu * lv * static_cast<float>(p11[k]);
        c[k] = f2u8_x86(ic);
      }
    }
  } else {
    // DiffuseDefaultColorizer (pixel_shader.h:222-230)
    c[0] = c[1] = c[2] = 255;
  }
  if (o.diffuse_shading == 1) {
    // DiffuseLambertianShader (pixel_shader.h:370-388); light_dir = ray_w
    float coeff = -sum3(r0 * n0, r1 * n1, r2 * n2);
    if (coeff < 0.0f) coeff = 0.0f;
    for (int k = 0; k < 3; k++) c[k] = f2u8_x86(coeff * static_cast<float>(c[k]));
  } else if (o.diffuse_shading == 2) {
    // DiffuseOrenNayarShader (pixel_shader.h:392-422); light_dir = ray_w
    const float dot_light = -sum3(r0 * n0, r1 * n1, r2 * n2);
    const float theta_i = acosf(dot_light);
    const float dot_ray = -sum3(r0 * n0, r1 * n1, r2 * n2);
    const float theta_r = acosf(dot_ray);
    const float bl0 = -n0 * dot_light - r0, bl1 = -n1 * dot_light - r1,
                bl2 = -n2 * dot_light - r2;
    const float phi_diff_cos = fmaxf(0.0f, sum3(bl0 * bl0, bl1 * bl1, bl2 * bl2));
    const float alpha = fmaxf(theta_i, theta_r);
    const float beta = fminf(theta_i, theta_r);
    const float coeff = fmaxf(0.0f, dot_light) *
                        (o.on_A + (o.on_B * phi_diff_cos * sinf(alpha) * tanf(beta)));
    for (int k = 0; k < 3; k++) c[k] = f2u8_x86(coeff * static_cast<float>(c[k]));
  }
}

UGU_DEV int xcd_remap(int b, int nb) {
  // Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
  // contiguous run of tiles so neighbouring tiles share its L2.
  const int q = nb >> 3, r = nb & 7;
  const int xcd = b & 7, idx = b >> 3;
  const int start = xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q;
  return start + idx;
}

}  // namespace ugu_hip
