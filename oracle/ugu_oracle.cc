// CPU restatement of UGU's software rasterizer, used ONLY as the test oracle
// and as the CPU baseline of the benchmark. Nothing in the product library
// links or calls this file: tests, the smoke check and the benchmark's
// cpu_baseline leg load it as a checker.
//
// It restates, expression by expression, with fp-contract off:
//   src/renderer/cpu/rasterizer.cc:105-307          Rasterizer::Impl::Render
//   src/renderer/cpu/rasterizer.cc:25-28            EdgeFunction
//   src/renderer/cpu/util_private.cc:12-89          ValidateAndInitBeforeRender
//   include/ugu/renderer/cpu/pixel_shader.h:143-422 colourizers and shaders
//   include/ugu/camera.h:380-387                    PinholeCamera::Project
//   include/ugu/camera.h:438-448,482-513            ray_c/ray_w and InitRayTable
//   include/ugu/image_proc.h:453-520                ConvertTo (RenderW)
//
// Eigen evaluation order. Fixed-size 3-term reductions (dot, squaredNorm,
// each row of Matrix3f*Vector3f) are evaluated by Eigen's unrolled redux as
// x0 + (x1 + x2); this is isolated in sum3() below. Element-wise expressions
// (a*x + b*y + c*z per coefficient) are plain left-to-right.
//
// Float->int conversions follow what GCC emits on x86-64 for the reference's
// static_casts (cvttss2si: truncation, INT32_MIN for NaN/out of range; the
// narrow casts keep the low bits). Two documented departures, both where the
// reference has undefined behaviour:
//   * the pixel bbox upper bound uses a saturating conversion (the reference's
//     int32 overflow turns into a 2^31-long loop writing wrapped pixels);
//   * texture reads whose linear index falls outside the image are clamped to
//     the image (the reference reads outside the buffer there).
//
// Parity pins for this file (closed forms, brute force, invariants) live in
// tests/test_oracle_*.py.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <limits>
#include <vector>
#include <algorithm>

#ifdef _OPENMP
#include <omp.h>
#endif

// Planted-mistake switch for the oracle's own pin tests
// (tests/test_oracle_mutations.py builds the oracle with -DORACLE_MUTATION=k
// and checks that the pins fail); 0 = the faithful restatement.
#ifndef ORACLE_MUTATION
#define ORACLE_MUTATION 0
#endif
#define MUT(k) (ORACLE_MUTATION == (k))

#if defined(__FP_FAST_FMAF) || defined(__FAST_MATH__)
#error "oracle must be compiled without FMA contraction or fast-math"
#endif

extern "C" {

struct OracleMesh {
  int32_t num_vertices;
  const float* vertices;        // V x 3
  int32_t num_faces;
  const int32_t* faces;         // F x 3
  const float* face_normals;    // F x 3 or null
  int32_t num_normals;
  const float* normals;         // N x 3 or null
  const int32_t* normal_indices;// F x 3 or null
  int32_t num_vertex_colors;
  const float* vertex_colors;   // V x 3 or null (0..255 scale)
  int32_t num_uv;
  const float* uv;              // U x 2 or null
  const int32_t* uv_indices;    // F x 3 or null
  const int32_t* material_ids;  // F or null
  int32_t num_materials;
  const int32_t* tex_width;     // per material
  const int32_t* tex_height;    // per material
  const uint8_t* const* tex_data;  // per material, H x W x 3 RGB
};

struct OracleCamera {
  int32_t width, height;
  float fx, fy, cx, cy;
  float w2c_R[9];  // row-major, = camera->w2c().rotation().cast<float>()
  float w2c_t[3];  // = camera->w2c().translation().cast<float>()
  float c2w_R[9];  // row-major, = c2w.matrix().block<3,3>(0,0).cast<float>()
};

// G-buffer planes (null = not requested): the fields of ugu::GBuffer
// (include/ugu/renderer/base.h:34-77) the rasterizer's resolve can fill, as
// defined by the library's UguGBuffer: covered = mask 255; u, v = the
// winner's weights w1, w2; w = (1 - u) - v.
struct OracleGBuf {
  float* bary;        // (u, v, (1 - u) - v)
  float* uv;          // (s, t, (1 - s) - t), (s, t) interpolated as the colourizers do
  float* pos_wld;     // w P0 + u P1 + v P2
  float* pos_cam;     // w2c_t + w2c_R pos_wld
  float* normal_wld;  // shading normal
  int32_t* geo_id;    // 0, -1 where not covered
};

struct OracleOptions {
  int32_t diffuse_color;    // 0 none, 1 texture, 2 vertex
  int32_t interp;           // 0 nn, 1 bilinear
  int32_t shading_normal;   // 0 face, 1 vertex
  int32_t diffuse_shading;  // 0 none, 1 lambertian, 2 oren-nayar
  float depth_scale;
  int32_t backface_culling;
  float oren_nayar_sigma;
};

}  // extern "C"

namespace {

// error codes mirror the reference's LOGE branches (util_private.cc:19-66)
enum {
  kOk = 0,
  kErrNoCamera = -1,
  kErrMeshNotInitialized = -2,
  kErrCullingNoFaceNormal = -3,
  kErrTextureEmpty = -4,
  kErrTextureNthEmpty = -5,
  kErrVertexColorEmpty = -6,
  kErrFaceNormalEmpty = -7,
  kErrVertexNormalEmpty = -8,
  kErrAllNull = -9,
  kErrDepthNull = -10,
};

struct V3 {
  float v[3];
  float& operator[](int i) { return v[i]; }
  const float& operator[](int i) const { return v[i]; }
};

// Eigen's unrolled 3-term reduction: x0 + (x1 + x2)
inline float sum3(float a, float b, float c) { return a + (b + c); }

inline float dot3(const V3& a, const V3& b) {
  return sum3(a[0] * b[0], a[1] * b[1], a[2] * b[2]);
}

inline V3 matvec(const float* R, const V3& x) {
  V3 r;
  for (int i = 0; i < 3; i++)
    r[i] = sum3(R[i * 3 + 0] * x[0], R[i * 3 + 1] * x[1], R[i * 3 + 2] * x[2]);
  return r;
}

// x86-64 cvttss2si: truncate, INT32_MIN for NaN / out of range.
inline int32_t f2i_x86(float f) {
  if (f >= -2147483648.0f && f < 2147483648.0f) return static_cast<int32_t>(f);
  return std::numeric_limits<int32_t>::min();
}
inline int32_t d2i_x86(double f) {
  if (f >= -2147483648.0 && f < 2147483648.0) return static_cast<int32_t>(f);
  return std::numeric_limits<int32_t>::min();
}
// saturating conversion used for the pixel bbox (see header comment)
inline int32_t f2i_sat(float f) {
  if (f >= 2147483648.0f) return std::numeric_limits<int32_t>::max();
  if (f >= -2147483648.0f) return static_cast<int32_t>(f);
  return std::numeric_limits<int32_t>::min();  // also NaN
}
inline uint8_t f2u8_x86(float f) {
  return static_cast<uint8_t>(static_cast<uint32_t>(f2i_x86(f)) & 0xffu);
}

// rasterizer.cc:25-28
inline float EdgeFunction(const V3& a, const V3& b, const V3& c) {
  return (c[0] - a[0]) * (b[1] - a[1]) - (c[1] - a[1]) * (b[0] - a[0]);
}

// std::min({a,b,c}) / std::max({a,b,c}) (min_element/max_element order)
inline float min3_std(float a, float b, float c) {
  float m = a;
  if (b < m) m = b;
  if (c < m) m = c;
  return m;
}
inline float max3_std(float a, float b, float c) {
  float m = a;
  if (m < b) m = b;
  if (m < c) m = c;
  return m;
}

V3 load3(const float* p, int64_t i) {
  V3 r;
  r[0] = p[i * 3 + 0];
  r[1] = p[i * 3 + 1];
  r[2] = p[i * 3 + 2];
  return r;
}

struct Tex {
  int32_t w, h;
  const uint8_t* d;
  // ImageBase::at (image.h:416-427) with the linear index clamped to the image
  const uint8_t* at(int32_t row, int32_t col) const {
    int64_t lin = static_cast<int64_t>(row) * w + col;
    int64_t n = static_cast<int64_t>(w) * h;
    if (lin < 0) lin = 0;
    if (lin >= n) lin = n - 1;
    return d + lin * 3;
  }
};

// camera.h:438-448 (ray_c then c2w_R_f * dir), evaluated for one pixel
inline V3 ray_w_at(const OracleCamera& cam, float x, float y) {
  V3 d;
  d[0] = (x - cam.cx) / cam.fx;
  d[1] = (y - cam.cy) / cam.fy;
  d[2] = 1.0f;
  // Eigen normalize(): z = squaredNorm(); if (z > 0) *this /= sqrt(z)
  float z = sum3(d[0] * d[0], d[1] * d[1], d[2] * d[2]);
  if (z > 0.0f) {
    float s = std::sqrt(z);
    d[0] = d[0] / s;
    d[1] = d[1] / s;
    d[2] = d[2] / s;
  }
  return matvec(cam.c2w_R, d);
}

// Per-pixel tables exactly as PinholeCamera::InitRayTable (camera.h:482-513):
// all four tables are filled (OpenMP over rows) so the CPU timing matches
// the reference's work.
struct RayTables {
  std::vector<V3> org_c, org_w, ray_c, ray_w;
};

void InitRayTable(const OracleCamera& cam, const V3& c2w_t, RayTables* t) {
  const size_t n = static_cast<size_t>(cam.width) * cam.height;
  t->org_c.resize(n);
  t->org_w.resize(n);
  t->ray_c.resize(n);
  t->ray_w.resize(n);
#pragma omp parallel for schedule(dynamic, 1)
  for (int y = 0; y < cam.height; y++) {
    for (int x = 0; x < cam.width; x++) {
      size_t i = static_cast<size_t>(y) * cam.width + x;
      t->org_c[i] = V3{{0.0f, 0.0f, 0.0f}};
      t->org_w[i] = c2w_t;
      V3 d;
      d[0] = (static_cast<float>(x) - cam.cx) / cam.fx;
      d[1] = (static_cast<float>(y) - cam.cy) / cam.fy;
      d[2] = 1.0f;
      float z = sum3(d[0] * d[0], d[1] * d[1], d[2] * d[2]);
      if (z > 0.0f) {
        float s = std::sqrt(z);
        d[0] = d[0] / s;
        d[1] = d[1] / s;
        d[2] = d[2] / s;
      }
      t->ray_c[i] = d;
      t->ray_w[i] = ray_w_at(cam, static_cast<float>(x), static_cast<float>(y));
    }
  }
}

// pixel_shader.h:143-148
struct OrenNayarParam {
  float sigma, A, B;
  explicit OrenNayarParam(float s) : sigma(s) {
    float sigma_2 = sigma * sigma;
    A = 1.0f - (0.5f * sigma_2 / (sigma_2 + 0.33f));
    B = 0.45f * sigma_2 / (sigma_2 + 0.09f);
  }
};

struct ShadeCtx {
  const OracleMesh* m;
  const OracleOptions* o;
  std::vector<Tex> tex;
  OrenNayarParam on;
  explicit ShadeCtx(float sigma) : on(sigma) {}
};

// interp of a 2-vector: (1-u-v)*a + u*b + v*c per coefficient (pixel_shader.h:277-279)
inline void interp_uv(const OracleMesh& m, int fid, float u, float v, float out[2]) {
  const int32_t* ui = m.uv_indices + static_cast<int64_t>(fid) * 3;
  const float* a = m.uv + static_cast<int64_t>(ui[0]) * 2;
  const float* b = m.uv + static_cast<int64_t>(ui[1]) * 2;
  const float* c = m.uv + static_cast<int64_t>(ui[2]) * 2;
  float w = 1.0f - u - v;
  for (int k = 0; k < 2; k++) out[k] = w * a[k] + u * b[k] + v * c[k];
}

// Colourizers (pixel_shader.h:222-359) and shaders (:363-422) for one pixel.
void ShadePixel(const ShadeCtx& s, uint8_t* c, float u, float v, int fid,
                const V3& ray_w, const V3& light_dir, const V3& n) {
  const OracleMesh& m = *s.m;
  // colourizer
  if (s.o->diffuse_color == 2) {
    // DiffuseVertexColorColorizer (:234-256)
    const int32_t* f = m.faces + static_cast<int64_t>(fid) * 3;
    V3 c0 = load3(m.vertex_colors, f[0]);
    V3 c1 = load3(m.vertex_colors, f[1]);
    V3 c2 = load3(m.vertex_colors, f[2]);
    float w = 1.0f - u - v;
    for (int k = 0; k < 3; k++) {
      float ic = w * c0[k] + u * c1[k] + v * c2[k];
      c[k] = f2u8_x86(ic);
    }
  } else if (s.o->diffuse_color == 1) {
    const Tex& t = s.tex[m.material_ids[fid]];
    float iuv[2];
    interp_uv(m, fid, u, v, iuv);
    float fx = iuv[0] * static_cast<float>(t.w - 1);
    float fy = (1.0f - iuv[1]) * static_cast<float>(t.h - 1);
    if (s.o->interp == 0) {
      // DiffuseTextureNnColorizer (:260-298)
      int tx = f2i_x86(MUT(7) ? std::floor(fx) : std::round(fx));
      int ty = f2i_x86(MUT(7) ? std::floor(fy) : std::round(fy));
      const uint8_t* dt = t.at(ty, tx);
      for (int k = 0; k < 3; k++) {
        float ic = static_cast<float>(dt[k]);
        c[k] = f2u8_x86(ic);
      }
    } else {
      // DiffuseTextureBilinearColorizer (:302-359)
      int x0 = f2i_x86(std::floor(fx));
      int y0 = f2i_x86(std::floor(fy));
      int x1 = x0 + 1;
      int y1 = y0 + 1;
      float lu = fx - static_cast<float>(x0);
      float lv = fy - static_cast<float>(y0);
      if (MUT(8)) std::swap(lu, lv);
      const uint8_t* a = t.at(y0, x0);
      const uint8_t* b = t.at(y0, x1);
      const uint8_t* cc = t.at(y1, x0);
      const uint8_t* d = t.at(y1, x1);
      for (int k = 0; k < 3; k++) {
        float ic = (1.0f - lu) * (1.0f - lv) * static_cast<float>(a[k]) +
                   lu * (1.0f - lv) * static_cast<float>(b[k]) +
                   (1.0f - lu) * lv * static_cast<float>(cc[k]) +
                   lu * lv * static_cast<float>(d[k]);
        c[k] = f2u8_x86(ic);
      }
    }
  } else {
    // DiffuseDefaultColorizer (:222-230)
    c[0] = c[1] = c[2] = 255;
  }
  // shader
  if (s.o->diffuse_shading == 1) {
    // DiffuseLambertianShader (:370-388)
    float coeff = MUT(9) ? dot3(light_dir, n) : -dot3(light_dir, n);
    if (coeff < 0.0f) coeff = 0.0f;
    for (int k = 0; k < 3; k++) c[k] = f2u8_x86(coeff * static_cast<float>(c[k]));
  } else if (s.o->diffuse_shading == 2) {
    // DiffuseOrenNayarShader (:392-422)
    float dot_light = -dot3(light_dir, n);
    float theta_i = std::acos(dot_light);
    float dot_ray = -dot3(ray_w, n);
    float theta_r = std::acos(dot_ray);
    V3 bl, br;
    for (int k = 0; k < 3; k++) {
      bl[k] = -n[k] * dot_light - light_dir[k];
      br[k] = -n[k] * dot_light - ray_w[k];
    }
    float phi_diff_cos = std::max(0.0f, dot3(bl, br));
    float alpha = std::max(theta_i, theta_r);
    float beta = std::min(theta_i, theta_r);
    float A = s.on.A, B = s.on.B;
    float coeff = std::max(0.0f, dot_light) *
                  (A + (B * phi_diff_cos * std::sin(alpha) * std::tan(beta)));
    for (int k = 0; k < 3; k++) c[k] = f2u8_x86(coeff * static_cast<float>(c[k]));
  }
}

int Validate(const OracleMesh& m, const OracleOptions& o, bool any_output) {
  // util_private.cc:12-66 (camera and PrepareMesh are always present here)
  if (o.backface_culling && m.face_normals == nullptr) return kErrCullingNoFaceNormal;
  if (o.diffuse_color == 1 && m.num_materials <= 0) return kErrTextureEmpty;
  if (o.diffuse_color == 1) {
    for (int i = 0; i < m.num_materials; i++)
      if (m.tex_width[i] < 1 || m.tex_height[i] < 1 || m.tex_data[i] == nullptr)
        return kErrTextureNthEmpty;
  }
  if (o.diffuse_color == 2 && m.vertex_colors == nullptr) return kErrVertexColorEmpty;
  if (o.shading_normal == 0 && m.face_normals == nullptr) return kErrFaceNormalEmpty;
  if (o.shading_normal == 1 && m.normals == nullptr) return kErrVertexNormalEmpty;
  if (!any_output) return kErrAllNull;
  return kOk;
}

}  // namespace

extern "C" {

// Restatement of Rasterizer::Impl::Render (rasterizer.cc:105-307).
// Output arrays are caller-owned, sized W*H*{3,1,3,1,1}; null = not requested.
// c2w_t (camera position) only feeds the org_ray_w table, as in InitRayTable.
// Returns 0 or a negative error code; *elapsed_ms receives the wall time of
// the main loop (rasterizer.cc:121-122,303-304), ray table included.
int oracle_render_gbuffer(const OracleMesh* mesh, const OracleCamera* cam,
                          const float* c2w_t, const OracleOptions* opt, uint8_t* color,
                          float* depth, float* normal, uint8_t* mask, int32_t* face_id,
                          const OracleGBuf* gbuf, double* elapsed_ms);

int oracle_render(const OracleMesh* mesh, const OracleCamera* cam,
                  const float* c2w_t, const OracleOptions* opt, uint8_t* color,
                  float* depth, float* normal, uint8_t* mask, int32_t* face_id,
                  double* elapsed_ms) {
  return oracle_render_gbuffer(mesh, cam, c2w_t, opt, color, depth, normal, mask, face_id,
                               nullptr, elapsed_ms);
}

// oracle_render plus the G-buffer planes of *gbuf (may be null).
int oracle_render_gbuffer(const OracleMesh* mesh, const OracleCamera* cam,
                          const float* c2w_t, const OracleOptions* opt, uint8_t* color,
                          float* depth, float* normal, uint8_t* mask, int32_t* face_id,
                          const OracleGBuf* gbuf, double* elapsed_ms) {
  if (cam == nullptr) return kErrNoCamera;
  if (mesh == nullptr) return kErrMeshNotInitialized;
  const OracleMesh& m = *mesh;
  const OracleOptions& o = *opt;
  const OracleGBuf g = gbuf ? *gbuf : OracleGBuf{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  const bool any_g = g.bary || g.uv || g.pos_wld || g.pos_cam || g.normal_wld || g.geo_id;
  int err = Validate(m, o, color || depth || normal || mask || face_id || any_g);
  if (err) return err;

  const int W = cam->width, H = cam->height;
  const size_t npx = static_cast<size_t>(W) * H;
  float* const gplanes[5] = {g.bary, g.uv, g.pos_wld, g.pos_cam, g.normal_wld};
  for (float* gp : gplanes)
    if (gp) std::fill(gp, gp + npx * 3, 0.0f);
  if (g.geo_id) std::fill(g.geo_id, g.geo_id + npx, -1);
  // ValidateAndInitBeforeRender: colour 0, depth 0, normal 0, mask 0, fid -1
  if (color) std::memset(color, 0, npx * 3);
  if (depth) std::fill(depth, depth + npx, 0.0f);
  if (normal) std::fill(normal, normal + npx * 3, 0.0f);
  if (mask) std::memset(mask, 0, npx);
  if (face_id) std::fill(face_id, face_id + npx, -1);

  ShadeCtx sc(0.3f);  // CopyTo never copies oren_nayar_sigma (renderer.h:30-37)
  sc.m = mesh;
  sc.o = opt;
  for (int i = 0; i < m.num_materials; i++)
    sc.tex.push_back(Tex{m.tex_width[i], m.tex_height[i], m.tex_data[i]});

  auto t0 = std::chrono::steady_clock::now();

  RayTables rt;
  V3 ct{{c2w_t ? c2w_t[0] : 0.0f, c2w_t ? c2w_t[1] : 0.0f, c2w_t ? c2w_t[2] : 0.0f}};
  InitRayTable(*cam, ct, &rt);

  // vertex loop (rasterizer.cc:125-136)
  const int V = m.num_vertices;
  std::vector<V3> camera_vertices(V), camera_normals(V), image_vertices(V);
  std::vector<float> camera_depth_list(V);
  for (int i = 0; i < V; i++) {
    V3 rv = matvec(cam->w2c_R, load3(m.vertices, i));
    for (int k = 0; k < 3; k++) camera_vertices[i][k] = rv[k] + cam->w2c_t[k];
    // the reference also rotates normals[i] (unused); only valid indices here
    if (m.normals && i < m.num_normals)
      camera_normals[i] = matvec(cam->w2c_R, load3(m.normals, i));
    camera_depth_list[i] = camera_vertices[i][2];
    // PinholeCamera::Project (camera.h:380-387): divide first, then multiply
    const V3& p = camera_vertices[i];
    image_vertices[i][0] = cam->fx / p[2] * p[0] + cam->cx;
    image_vertices[i][1] = cam->fy / p[2] * p[1] + (MUT(10) ? cam->cx : cam->cy);
    image_vertices[i][2] = p[2];
  }

  std::vector<float> depth_internal;
  float* depth_ = depth;
  if (depth_ == nullptr) {
    depth_internal.assign(npx, 0.0f);
    depth_ = depth_internal.data();
  }
  std::vector<int32_t> fid_internal;
  int32_t* face_id_ = face_id;
  if (face_id_ == nullptr) {
    fid_internal.assign(npx, -1);
    face_id_ = fid_internal.data();
  }
  std::vector<uint8_t> backface_image(npx, 0);
  std::vector<float> weight_image(npx * 3, 0.0f);

  // face loop (rasterizer.cc:161-241)
  const float flt_min = std::numeric_limits<float>::min();
  for (int i = 0; i < m.num_faces; i++) {
    const int32_t* face = m.faces + static_cast<int64_t>(i) * 3;
    const V3& v0_i = image_vertices[face[0]];
    const V3& v1_i = image_vertices[face[1]];
    const V3& v2_i = image_vertices[face[2]];
    if (v0_i[2] < 0.0f || v1_i[2] < 0.0f || v2_i[2] < 0.0f) continue;
    float xmin = min3_std(v0_i[0], v1_i[0], v2_i[0]);
    float ymin = min3_std(v0_i[1], v1_i[1], v2_i[1]);
    float xmax = max3_std(v0_i[0], v1_i[0], v2_i[0]);
    float ymax = max3_std(v0_i[1], v1_i[1], v2_i[1]);
    if (xmin > static_cast<float>(W - 1) || xmax < 0 ||
        ymin > static_cast<float>(H - 1) || ymax < 0)
      continue;
    int32_t x0 = std::max(int32_t(0), f2i_x86(std::ceil(xmin)));
    int32_t x1 = std::min(W - 1, f2i_sat(std::floor(xmax)));
    int32_t y0 = std::max(int32_t(0), f2i_x86(std::ceil(ymin)));
    int32_t y1 = std::min(H - 1, f2i_sat(std::floor(ymax)));
    float area = EdgeFunction(v0_i, v1_i, v2_i);
    if (std::abs(area) < flt_min) continue;
    const V3 fn = load3(m.face_normals ? m.face_normals : m.vertices, i);
    for (int32_t y = y0; y <= y1; ++y) {
      for (int32_t x = x0; x <= x1; ++x) {
        const V3& ray_w = rt.ray_w[static_cast<size_t>(y) * W + x];
        bool backface = m.face_normals ? (MUT(2) ? dot3(fn, ray_w) < 0 : dot3(fn, ray_w) > 0) : false;
        V3 pixel_sample{{static_cast<float>(x), static_cast<float>(y), 0.0f}};
        float w0 = MUT(1) ? EdgeFunction(v2_i, v1_i, pixel_sample)
                          : EdgeFunction(v1_i, v2_i, pixel_sample);
        float w1 = EdgeFunction(v2_i, v0_i, pixel_sample);
        float w2 = EdgeFunction(v0_i, v1_i, pixel_sample);
        if ((!backface && (w0 >= 0 && w1 >= 0 && w2 >= 0)) ||
            (backface && (w0 <= 0 && w1 <= 0 && w2 <= 0))) {
          w0 /= area;
          w1 /= area;
          w2 /= area;
          if (MUT(4)) {
            // affine (not perspective-correct) interpolation
            pixel_sample[2] = w0 * v0_i[2] + w1 * v1_i[2] + w2 * v2_i[2];
          } else {
            w0 /= v0_i[2];
            w1 /= v1_i[2];
            w2 /= v2_i[2];
            pixel_sample[2] = 1.0f / (w0 + w1 + w2);
            w0 = w0 * pixel_sample[2];
            w1 = w1 * pixel_sample[2];
            w2 = w2 * pixel_sample[2];
          }
          size_t pi = static_cast<size_t>(y) * W + x;
          float& d = depth_[pi];
          if (d < flt_min || (MUT(3) ? pixel_sample[2] <= d : pixel_sample[2] < d)) {
            d = pixel_sample[2];
            face_id_[pi] = i;
            weight_image[pi * 3 + 0] = w0;
            weight_image[pi * 3 + 1] = w1;
            weight_image[pi * 3 + 2] = w2;
            backface_image[pi] = backface ? 255 : 0;
          }
        }
      }
    }
  }

  // resolve loop (rasterizer.cc:243-301)
  for (int y = 0; y < H; y++) {
    for (int x = 0; x < W; x++) {
      size_t pi = static_cast<size_t>(y) * W + x;
      const uint8_t bf = backface_image[pi];
      int32_t& fid = face_id_[pi];
      if (o.backface_culling && bf == 255 && !MUT(13)) {
        depth_[pi] = 0.0f;
        fid = -1;
        continue;
      }
      if (MUT(5) ? fid >= 0 : fid > 0) {
        const V3& ray_w = rt.ray_w[pi];
        float w0 = weight_image[pi * 3 + 0];
        float w1 = weight_image[pi * 3 + 1];
        float w2 = weight_image[pi * 3 + 2];
        if (mask) mask[pi] = 255;
        V3 sn{{0.0f, 0.0f, 0.0f}};
        if (o.shading_normal == 0) {
          sn = load3(m.face_normals, fid);
        } else {
          const int32_t* ni = m.normal_indices + static_cast<int64_t>(fid) * 3;
          V3 na = load3(m.normals, ni[0]);
          V3 nb = load3(m.normals, ni[1]);
          V3 nc = load3(m.normals, ni[2]);
          for (int k = 0; k < 3; k++) sn[k] = w0 * na[k] + w1 * nb[k] + w2 * nc[k];
        }
        if (normal) {
          V3 snc = matvec(cam->w2c_R, sn);
          if (MUT(12)) {
            float Rt[9];
            for (int a = 0; a < 3; a++)
              for (int b = 0; b < 3; b++) Rt[a * 3 + b] = cam->w2c_R[b * 3 + a];
            snc = matvec(Rt, sn);
          }
          for (int k = 0; k < 3; k++) normal[pi * 3 + k] = snc[k];
        }
        if (color) {
          V3 light_dir = ray_w;
          if (MUT(6))
            ShadePixel(sc, color + pi * 3, w2, w1, fid, ray_w, light_dir, sn);
          else
            ShadePixel(sc, color + pi * 3, w1, w2, fid, ray_w, light_dir, sn);
        }
        if (any_g) {
          const float u = w1, v = w2, w = 1.0f - u - v;
          if (g.bary) {
            g.bary[pi * 3 + 0] = MUT(14) ? v : u;
            g.bary[pi * 3 + 1] = MUT(14) ? u : v;
            g.bary[pi * 3 + 2] = 1.0f - u - v;
          }
          if (g.uv && m.uv && m.uv_indices) {
            float st[2];
            interp_uv(m, fid, u, v, st);
            g.uv[pi * 3 + 0] = st[0];
            g.uv[pi * 3 + 1] = st[1];
            g.uv[pi * 3 + 2] = 1.0f - st[0] - st[1];
          }
          if (g.pos_wld || g.pos_cam) {
            const int32_t* f = m.faces + static_cast<int64_t>(fid) * 3;
            const V3 p0 = load3(m.vertices, f[0]), p1 = load3(m.vertices, f[1]),
                     p2 = load3(m.vertices, f[2]);
            V3 pw;
            for (int k = 0; k < 3; k++) pw[k] = w * p0[k] + u * p1[k] + v * p2[k];
            float Rt[9];
            for (int a = 0; a < 3; a++)
              for (int b = 0; b < 3; b++) Rt[a * 3 + b] = cam->w2c_R[b * 3 + a];
            const V3 rp = matvec(MUT(15) ? Rt : cam->w2c_R, pw);
            for (int k = 0; k < 3; k++) {
              if (g.pos_wld) g.pos_wld[pi * 3 + k] = pw[k];
              if (g.pos_cam) g.pos_cam[pi * 3 + k] = cam->w2c_t[k] + rp[k];
            }
          }
          if (g.normal_wld)
            for (int k = 0; k < 3; k++) g.normal_wld[pi * 3 + k] = sn[k];
          if (g.geo_id) g.geo_id[pi] = 0;
        }
      }
    }
  }

  auto t1 = std::chrono::steady_clock::now();
  if (elapsed_ms)
    *elapsed_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  return kOk;
}

// ---------------------------------------------------------------------------
// Raytracer::Impl::Render (raytracer.cc:117-225) with an exact closest hit:
// every face is tested (no BVH) with the reference's float64 Moller-Trumbore
// routine IntersectImpl (geom_util.cc:60-115) at the naive BVH's epsilon
// (bvh_naive.h:121, 1e-8f) and its standard-barycentric conversion; the
// closest hit is the smallest float t, equal t -> lowest face id.

// Eigen's dot of two float64 3-vectors: one Packet2d for coefficients 0..1
// and a scalar tail (redux_impl, LinearVectorizedTraversal): (x0 + x1) + x2.
static inline double sum3d(double a, double b, double c) { return (a + b) + c; }

// Eigen's cross for 3-vectors: (a1 b2 - a2 b1, a2 b0 - a0 b2, a0 b1 - a1 b0)
static inline void cross_d(const double a[3], const double b[3], double r[3]) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double dot_d(const double a[3], const double b[3]) {
  return sum3d(a[0] * b[0], a[1] * b[1], a[2] * b[2]);
}

// returns true and (t, u, v) as floats on a hit
static bool moller_trumbore(const double o[3], const double d[3], const double v0[3],
                     const double v1[3], const double v2[3], double eps, float* t,
                     float* u, float* v) {
  double e1[3], e2[3], alpha[3], r[3], beta[3];
  for (int k = 0; k < 3; k++) {
    e1[k] = v1[k] - v0[k];
    e2[k] = v2[k] - v0[k];
  }
  cross_d(d, e2, alpha);
  const double det = dot_d(e1, alpha);
  if (-eps < det && det < eps) return false;
  const double inv_det = 1.0 / det;
  for (int k = 0; k < 3; k++) r[k] = o[k] - v0[k];
  const double uu = dot_d(alpha, r) * inv_det;
  if (uu < 0.0f || uu > 1.0f) return false;
  cross_d(r, e1, beta);
  const double vv = dot_d(d, beta) * inv_det;
  if (vv < 0.0f || uu + vv > 1.0f) return false;
  const double tt = dot_d(e2, beta) * inv_det;
  if (tt < 0.0f) return false;
  // standard barycentric: u <- 1 - u - v, v <- u
  const double w = 1 - uu - vv;
  *t = static_cast<float>(tt);
  *u = static_cast<float>(MUT(16) ? uu : w);
  *v = static_cast<float>(MUT(16) ? w : uu);
  return true;
}

int oracle_raytrace(const OracleMesh* mesh, const OracleCamera* cam, const float* c2w_t,
                    const OracleOptions* opt, uint8_t* color, float* depth, float* normal,
                    uint8_t* mask, int32_t* face_id, double* elapsed_ms) {
  if (cam == nullptr) return kErrNoCamera;
  if (mesh == nullptr) return kErrMeshNotInitialized;
  const OracleMesh& m = *mesh;
  const OracleOptions& o = *opt;
  int err = Validate(m, o, color || depth || normal || mask || face_id);
  if (err) return err;
  const int W = cam->width, H = cam->height;
  const size_t npx = static_cast<size_t>(W) * H;
  if (color) std::memset(color, 0, npx * 3);
  if (depth) std::fill(depth, depth + npx, 0.0f);
  if (normal) std::fill(normal, normal + npx * 3, 0.0f);
  if (mask) std::memset(mask, 0, npx);
  if (face_id) std::fill(face_id, face_id + npx, -1);
  auto t0 = std::chrono::steady_clock::now();
  V3 ct{{c2w_t ? c2w_t[0] : 0.0f, c2w_t ? c2w_t[1] : 0.0f, c2w_t ? c2w_t[2] : 0.0f}};
  RayTables rt;
  InitRayTable(*cam, ct, &rt);
  ShadeCtx sc(0.3f);  // Raytracer::set_option also goes through CopyTo (raytracer.cc:72-74)
  sc.m = &m;
  sc.o = &o;
  for (int i = 0; i < m.num_materials; i++) sc.tex.push_back(Tex{m.tex_width[i], m.tex_height[i], m.tex_data[i]});
  const double eps = static_cast<double>(1e-8f);
  const int F = m.num_faces;
#pragma omp parallel for schedule(dynamic, 1)
  for (int y = 0; y < H; y++) {
    for (int x = 0; x < W; x++) {
      const size_t pi = static_cast<size_t>(y) * W + x;
      const V3& ray_w = rt.ray_w[pi];
      const V3 org = MUT(20) ? V3{{0.0f, 0.0f, 0.0f}} : rt.org_w[pi];
      const double od[3] = {org[0], org[1], org[2]};
      const double dd[3] = {ray_w[0], ray_w[1], ray_w[2]};
      float bt = std::numeric_limits<float>::infinity(), bu = 0.0f, bv = 0.0f;
      int fid = -1;
      for (int f = 0; f < F; f++) {
        const int32_t* fi = m.faces + static_cast<int64_t>(f) * 3;
        double p[3][3];
        bool finite = true;
        for (int k = 0; k < 3; k++)
          for (int c = 0; c < 3; c++) {
            p[k][c] = m.vertices[static_cast<int64_t>(fi[k]) * 3 + c];
            finite = finite && std::isfinite(p[k][c]);
          }
        // a face with a non-finite coordinate can give a NaN t, which the
        // reference's std::sort cannot order: such faces are never hit here
        if (!finite) continue;
        float t, u, v;
        if (moller_trumbore(od, dd, p[0], p[1], p[2], eps, &t, &u, &v) &&
            (MUT(17) ? (fid < 0 || t > bt) : t < bt)) {
          bt = t;
          bu = u;
          bv = v;
          fid = f;
        }
      }
      if (fid < 0) continue;
      if (o.backface_culling && (MUT(18) ? dot3(load3(m.face_normals, fid), ray_w) < 0
                                         : dot3(load3(m.face_normals, fid), ray_w) > 0))
        continue;
      if (face_id) face_id[pi] = fid;
      if (mask) mask[pi] = 255;
      if (depth) {
        V3 hit;
        for (int k = 0; k < 3; k++) hit[k] = org[k] + ray_w[k] * bt;
        const V3 hc = matvec(cam->w2c_R, hit);
        depth[pi] = MUT(19) ? bt * o.depth_scale : (hc[2] + cam->w2c_t[2]) * o.depth_scale;
      }
      V3 sn{{0.0f, 0.0f, 0.0f}};
      if (o.shading_normal == 0) {
        sn = load3(m.face_normals, fid);
      } else {
        const int32_t* ni = m.normal_indices + static_cast<int64_t>(fid) * 3;
        const V3 na = load3(m.normals, ni[0]), nb = load3(m.normals, ni[1]),
                 nc = load3(m.normals, ni[2]);
        const float w = 1.0f - bu - bv;
        for (int k = 0; k < 3; k++) sn[k] = w * na[k] + bu * nb[k] + bv * nc[k];
      }
      if (normal) {
        const V3 snc = matvec(cam->w2c_R, sn);
        for (int k = 0; k < 3; k++) normal[pi * 3 + k] = snc[k];
      }
      if (color) ShadePixel(sc, color + pi * 3, bu, bv, fid, ray_w, ray_w, sn);
    }
  }
  auto t1 = std::chrono::steady_clock::now();
  if (elapsed_ms) *elapsed_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  return kOk;
}

// ---------------------------------------------------------------------------
// VisibilityTester::TestVertex (src/texturing/visibility_tester.cc:390-530)
// for n items against one keyframe; the closest hit of the vertex ray over all
// faces (smallest float t) as in oracle_raytrace.

struct OracleVisOptions {
  float backface_culling_angle_rad_th;
  int32_t use_mask, use_depth, interp;
  float max_viewing_angle, max_distance_from_camera_to_vertex, max_depth_difference, eps;
};

struct OracleVisItem {
  int32_t visible;
  float px, py, angle, dist, c0, c1, c2, intensity;
};

int oracle_visibility(const OracleMesh* mesh, const OracleCamera* cam, const float* c2w_t,
                      int32_t W, int32_t H, const uint8_t* color, const float* depth,
                      const uint8_t* mask, const OracleVisOptions* opt, int32_t n,
                      const float* item_pos, const float* item_vertex, const float* item_normal,
                      OracleVisItem* out) {
  const OracleMesh& m = *mesh;
  const OracleVisOptions& o = *opt;
  const double eps_mt = static_cast<double>(1e-8f);
#pragma omp parallel for schedule(dynamic, 64)
  for (int32_t i = 0; i < n; i++) {
    OracleVisItem r{0, 0.0f, 0.0f, 999.9f, -999.9f, 0.0f, 0.0f, 0.0f, -1.0f};
    out[i] = r;
    const V3 wp = load3(item_pos, i);
    const V3 cp = matvec(cam->w2c_R, wp);
    const float qx = cp[0] + cam->w2c_t[0], qy = cp[1] + cam->w2c_t[1],
                qz = cp[2] + cam->w2c_t[2];
    const float ix = cam->fx / qz * qx + cam->cx;
    const float iy = cam->fy / qz * qy + cam->cy;
    if (qz < 0 || ix < 0 || static_cast<float>(W - 1) < ix || iy < 0 ||
        static_cast<float>(H - 1) < iy)
      continue;
    if (!(ix >= 0.0f && iy >= 0.0f)) continue;  // NaN projection
    const int nx = static_cast<int>(MUT(22) ? std::floor(ix) : std::round(ix));
    const int ny = static_cast<int>(MUT(22) ? std::floor(iy) : std::round(iy));
    const size_t pi = static_cast<size_t>(ny) * W + nx;
    if (o.use_mask && mask[pi] != 255) continue;
    if (o.use_depth && depth[pi] < std::numeric_limits<float>::epsilon()) continue;
    const V3 ray = ray_w_at(*cam, ix, iy);
    const V3 nrm = load3(item_normal, i);
    const float dot = MUT(23) ? dot3(nrm, ray) : -dot3(nrm, ray);
    const float angle = std::acos(dot);
    if (o.backface_culling_angle_rad_th < angle) continue;
    const double od[3] = {c2w_t[0], c2w_t[1], c2w_t[2]};
    const double dd[3] = {ray[0], ray[1], ray[2]};
    float bt = std::numeric_limits<float>::infinity();
    for (int32_t f = 0; f < m.num_faces; f++) {
      const int32_t* fi = m.faces + static_cast<int64_t>(f) * 3;
      double p[3][3];
      bool finite = true;
      for (int k = 0; k < 3; k++)
        for (int c = 0; c < 3; c++) {
          p[k][c] = m.vertices[static_cast<int64_t>(fi[k]) * 3 + c];
          finite = finite && std::isfinite(p[k][c]);
        }
      if (!finite) continue;
      float t, u, v;
      if (moller_trumbore(od, dd, p[0], p[1], p[2], eps_mt, &t, &u, &v) && t < bt) bt = t;
    }
    const V3 vert = load3(item_vertex, i);
    V3 hit = vert;
    const bool has_hit = bt < std::numeric_limits<float>::infinity();
    const float distance = has_hit ? bt : std::numeric_limits<float>::max();
    if (has_hit) {
      for (int k = 0; k < 3; k++) hit[k] = c2w_t[k] + ray[k] * bt;
      const float dx = hit[0] - vert[0], dy = hit[1] - vert[1], dz = hit[2] - vert[2];
      if (!MUT(21) && std::sqrt(sum3(dx * dx, dy * dy, dz * dz)) > o.eps) continue;
    }
    if (o.use_depth) {
      const V3 hc = matvec(cam->w2c_R, hit);
      const float diff = std::fabs((hc[2] + cam->w2c_t[2]) - depth[pi]);
      if (o.max_depth_difference < diff) continue;
    }
    r.px = ix;
    r.py = iy;
    r.angle = angle;
    r.dist = distance;
    out[i] = r;
    if (distance > o.max_distance_from_camera_to_vertex || angle > o.max_viewing_angle) continue;
    float c[3];
    if (o.interp == 0) {
      for (int k = 0; k < 3; k++) c[k] = color[pi * 3 + k];
    } else {
      int x0 = static_cast<int>(std::floor(ix)), y0 = static_cast<int>(std::floor(iy));
      int x1 = x0 + 1, y1 = y0 + 1;
      if (x0 < 0) x0 = 0;
      if (y0 < 0) y0 = 0;
      if (W <= x1) x1 = W - 1;
      if (H <= y1) y1 = H - 1;
      float lu = ix - x0, lv = iy - y0;
      if (MUT(24)) std::swap(lu, lv);
      const uint8_t* c00 = color + (static_cast<size_t>(y0) * W + x0) * 3;
      const uint8_t* c01 = color + (static_cast<size_t>(y0) * W + x1) * 3;
      const uint8_t* c10 = color + (static_cast<size_t>(y1) * W + x0) * 3;
      const uint8_t* c11 = color + (static_cast<size_t>(y1) * W + x1) * 3;
      for (int k = 0; k < 3; k++) {
        const float f = (1.0f - lu) * (1.0f - lv) * c00[k] + lu * (1.0f - lv) * c01[k] +
                        (1.0f - lu) * lv * c10[k] + lu * lv * c11[k];
        c[k] = f2u8_x86(f);
      }
    }
    r.c0 = c[0];
    r.c1 = c[1];
    r.c2 = c[2];
    r.intensity = static_cast<float>(0.2989 * c[0] + 0.5870 * c[1] + 0.1140 * c[2]);
    r.visible = 1;
    out[i] = r;
  }
  return kOk;
}

// ---------------------------------------------------------------------------
// GenerateTextureOnOriginalUv's face loop with RasterizeTriangle and
// BilinearInterpolation, then PaddingSimple (texture_mapper.cc:148-201,
// 234-345; image_util.h:155-193), serially in face order.

struct OracleTexFace {
  int32_t keyframe;
  float src[6];
  float dst[6];
};

static float tm_edge(float ax, float ay, float bx, float by, float cx, float cy) {
  return (cx - ax) * (by - ay) - (cy - ay) * (bx - ax);
}

int oracle_texture_map(const int32_t* img_w, const int32_t* img_h, const uint8_t* const* imgs,
                       int32_t n_faces, const OracleTexFace* faces, int32_t W, int32_t H,
                       int32_t padding_kernel, uint8_t* texture, uint8_t* mask) {
  std::memset(texture, 0, size_t(W) * H * 3);
  std::memset(mask, 0, size_t(W) * H);
  std::vector<int32_t> written_by(size_t(W) * H, -1);
  for (int32_t fi = 0; fi < n_faces; fi++) {
    const OracleTexFace& f = faces[fi];
    if (f.keyframe < 0) continue;
    const int32_t iw = img_w[f.keyframe], ih = img_h[f.keyframe];
    const uint8_t* img = imgs[f.keyframe];
    const float* t = f.dst;
    float area = tm_edge(t[0], t[1], t[2], t[3], t[4], t[5]);
    if (std::fabs(area) < std::numeric_limits<float>::min())
      area = area > 0 ? std::numeric_limits<float>::min() : -std::numeric_limits<float>::min();
    const float inv_area = 1.0f / area;
    int xmin = f2i_x86(std::min({t[0], t[2], t[4]}) - 1);
    xmin = std::max(0, std::min(xmin, W - 1));
    int xmax = f2i_x86(std::max({t[0], t[2], t[4]}) + 1);
    xmax = std::max(0, std::min(xmax, W - 1));
    int ymin = f2i_x86(std::min({t[1], t[3], t[5]}) - 1);
    ymin = std::max(0, std::min(ymin, H - 1));
    int ymax = f2i_x86(std::max({t[1], t[3], t[5]}) + 1);
    ymax = std::max(0, std::min(ymax, H - 1));
    for (int y = ymin; y <= ymax; y++) {
      for (int x = xmin; x <= xmax; x++) {
        const float px = static_cast<float>(x), py = static_cast<float>(y);
        float w0 = tm_edge(t[2], t[3], t[4], t[5], px, py);
        float w1 = tm_edge(t[4], t[5], t[0], t[1], px, py);
        float w2 = tm_edge(t[0], t[1], t[2], t[3], px, py);
        w0 *= inv_area;
        w1 *= inv_area;
        w2 *= inv_area;
        if (w0 < 0 || w1 < 0 || w2 < 0) continue;
        const size_t ti = size_t(y) * W + x;
        if (MUT(25) && written_by[ti] >= 0) continue;  // planted: the first face wins
        written_by[ti] = fi;
        const float sx = w0 * f.src[0] + w1 * f.src[2] + w2 * f.src[4];
        const float sy = w0 * f.src[1] + w1 * f.src[3] + w2 * f.src[5];
        int x0 = static_cast<int>(std::floor(sx)), y0 = static_cast<int>(std::floor(sy));
        int x1 = x0 + 1, y1 = y0 + 1;
        if (x0 < 0) x0 = 0;
        if (y0 < 0) y0 = 0;
        if (iw <= x1) x1 = iw - 1;
        if (ih <= y1) y1 = ih - 1;
        float lu = sx - x0, lv = sy - y0;
        if (MUT(27)) std::swap(lu, lv);
        auto at = [&](int yy, int xx, int c) {
          yy = std::min(std::max(yy, 0), ih - 1);  // in bounds for a visible face
          xx = std::min(std::max(xx, 0), iw - 1);
          return static_cast<float>(img[(size_t(yy) * iw + xx) * 3 + c]);
        };
        for (int c = 0; c < 3; c++) {
          const float v = (1.0f - lu) * (1.0f - lv) * at(y0, x0, c) + lu * (1.0f - lv) * at(y0, x1, c) +
                          (1.0f - lu) * lv * at(y1, x0, c) + lu * lv * at(y1, x1, c);
          texture[ti * 3 + c] = f2u8_x86(v);
        }
        mask[ti] = 255;
      }
    }
  }
  // PaddingSimple on copies of the texture and mask
  const std::vector<uint8_t> t0(texture, texture + size_t(W) * H * 3);
  const std::vector<uint8_t> m0(mask, mask + size_t(W) * H);
  const int hk = padding_kernel / 2;
  for (int j = hk; j < H - hk; j++) {
    for (int i = hk; i < W - hk; i++) {
      // planted mistake 26: read the texture being padded instead of its copy
      const uint8_t* mm = MUT(26) ? mask : m0.data();
      if (m0[size_t(j) * W + i] == 255) continue;
      float a[3] = {0, 0, 0};
      size_t cnt = 0;
      for (int jj = -hk; jj <= hk; jj++)
        for (int ii = -hk; ii <= hk; ii++) {
          const size_t k = size_t(j + jj) * W + (i + ii);
          if (mm[k] == 255) {
            for (int c = 0; c < 3; c++) a[c] += (MUT(26) ? texture : t0.data())[k * 3 + c];
            cnt++;
          }
        }
      if (cnt == 0) continue;
      mask[size_t(j) * W + i] = 255;
      for (int c = 0; c < 3; c++) {
        a[c] /= cnt;
        texture[(size_t(j) * W + i) * 3 + c] = f2u8_x86(a[c]);
      }
    }
  }
  return kOk;
}

// RenderW / RenderDepthW (rasterizer.cc:329-348) depth conversion:
// ConvertTo(f_depth, depth) with scale 1 -> static_cast<uint16_t>(1.0 * (double)d)
void oracle_depth_to_u16(const float* depth, int64_t n, uint16_t* out) {
  for (int64_t i = 0; i < n; i++) {
    double val = static_cast<double>(depth[i]);
    double s = static_cast<double>(1.0f) * val;
    out[i] = static_cast<uint16_t>(static_cast<uint32_t>(d2i_x86(s)) & 0xffffu);
  }
}

// Exposed pieces for unit pins.
float oracle_edge_function(const float* a, const float* b, const float* c) {
  V3 A{{a[0], a[1], 0}}, B{{b[0], b[1], 0}}, C{{c[0], c[1], 0}};
  return EdgeFunction(A, B, C);
}

void oracle_ray_w(const OracleCamera* cam, float x, float y, float* out) {
  V3 r = ray_w_at(*cam, x, y);
  out[0] = r[0];
  out[1] = r[1];
  out[2] = r[2];
}

int oracle_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Differentiable rendering reference in float64 (no reference counterpart:
// the image model and its gradients are those documented in
// include/ugu_hip.h for ugu_hip_diff_forward / ugu_hip_diff_backward).
// Visibility is fixed: face_id[p] comes from the forward render; a pixel is
// shaded when face_id > 0.

namespace {

struct DiffCtx {
  int W, H;
  double fx, fy, cx, cy, R[9], t[3];
  const int32_t* faces;
  const float* uv;
  const int32_t* uvidx;
  const double* verts;
  const double* colors;
  const double* tex;
  int tw, th;
};

struct PixelEval {
  double X[3], Y[3], Z[3], Qx[3], Qy[3];
  double e[3], A, q[3], s, b0, b1, b2;
  double c[3], tt[3], dcdu[3], dcdv[3], dtdu[3], dtdv[3];
  int tx0, ty0, tx1, ty1;
  double lu, lv;
  double uv0[2], uv1[2], uv2[2];
};

void eval_pixel(const DiffCtx& d, int fid, int px, int py, PixelEval& e) {
  const int32_t* f = d.faces + int64_t(fid) * 3;
  const double x = px, y = py;
  for (int i = 0; i < 3; i++) {
    const double* P = d.verts + int64_t(f[i]) * 3;
    double Q[3];
    for (int r = 0; r < 3; r++)
      Q[r] = d.R[r * 3 + 0] * P[0] + d.R[r * 3 + 1] * P[1] + d.R[r * 3 + 2] * P[2] + d.t[r];
    e.Qx[i] = Q[0];
    e.Qy[i] = Q[1];
    e.Z[i] = Q[2];
    e.X[i] = d.fx * Q[0] / Q[2] + d.cx;
    e.Y[i] = d.fy * Q[1] / Q[2] + d.cy;
  }
  auto E = [](double ax, double ay, double bx, double by, double cx_, double cy_) {
    return (cx_ - ax) * (by - ay) - (cy_ - ay) * (bx - ax);
  };
  e.A = E(e.X[0], e.Y[0], e.X[1], e.Y[1], e.X[2], e.Y[2]);
  e.e[0] = E(e.X[1], e.Y[1], e.X[2], e.Y[2], x, y);
  e.e[1] = E(e.X[2], e.Y[2], e.X[0], e.Y[0], x, y);
  e.e[2] = E(e.X[0], e.Y[0], e.X[1], e.Y[1], x, y);
  for (int i = 0; i < 3; i++) e.q[i] = MUT(28) ? e.e[i] / e.A : e.e[i] / e.A / e.Z[i];
  e.s = e.q[0] + e.q[1] + e.q[2];
  e.b1 = e.q[1] / e.s;
  e.b2 = e.q[2] / e.s;
  if (MUT(29)) std::swap(e.b1, e.b2);
  e.b0 = 1.0 - e.b1 - e.b2;
  for (int k = 0; k < 3; k++) {
    e.c[k] = 1.0; e.tt[k] = 1.0;
    e.dcdu[k] = e.dcdv[k] = e.dtdu[k] = e.dtdv[k] = 0.0;
  }
  if (d.colors) {
    const double* C0 = d.colors + int64_t(f[0]) * 3;
    const double* C1 = d.colors + int64_t(f[1]) * 3;
    const double* C2 = d.colors + int64_t(f[2]) * 3;
    for (int k = 0; k < 3; k++) {
      e.c[k] = e.b0 * C0[k] + e.b1 * C1[k] + e.b2 * C2[k];
      e.dcdu[k] = C1[k] - C0[k];
      e.dcdv[k] = C2[k] - C0[k];
    }
  }
  if (d.tex) {
    const int32_t* ui = d.uvidx + int64_t(fid) * 3;
    for (int k = 0; k < 2; k++) {
      e.uv0[k] = d.uv[int64_t(ui[0]) * 2 + k];
      e.uv1[k] = d.uv[int64_t(ui[1]) * 2 + k];
      e.uv2[k] = d.uv[int64_t(ui[2]) * 2 + k];
    }
    const double iu = e.b0 * e.uv0[0] + e.b1 * e.uv1[0] + e.b2 * e.uv2[0];
    const double iv = e.b0 * e.uv0[1] + e.b1 * e.uv1[1] + e.b2 * e.uv2[1];
    const double fxp = iu * (d.tw - 1), fyp = (MUT(30) ? iv : 1.0 - iv) * (d.th - 1);
    const double flx = std::floor(fxp), fly = std::floor(fyp);
    e.lu = fxp - flx;
    e.lv = fyp - fly;
    if (MUT(31)) std::swap(e.lu, e.lv);
    auto clampi = [](double v, int hi) {
      if (!(v >= 0.0)) return 0;
      if (v > hi) return hi;
      return static_cast<int>(v);
    };
    e.tx0 = clampi(flx, d.tw - 1);
    e.ty0 = clampi(fly, d.th - 1);
    e.tx1 = clampi(flx + 1.0, d.tw - 1);
    e.ty1 = clampi(fly + 1.0, d.th - 1);
    const double* p00 = d.tex + (int64_t(e.ty0) * d.tw + e.tx0) * 3;
    const double* p01 = d.tex + (int64_t(e.ty0) * d.tw + e.tx1) * 3;
    const double* p10 = d.tex + (int64_t(e.ty1) * d.tw + e.tx0) * 3;
    const double* p11 = d.tex + (int64_t(e.ty1) * d.tw + e.tx1) * 3;
    const double sx = d.tw - 1, sy = -(d.th - 1.0);
    const double dlu_du = sx * (e.uv1[0] - e.uv0[0]), dlu_dv = sx * (e.uv2[0] - e.uv0[0]);
    const double dlv_du = sy * (e.uv1[1] - e.uv0[1]), dlv_dv = sy * (e.uv2[1] - e.uv0[1]);
    const double lu = e.lu, lv = e.lv;
    for (int k = 0; k < 3; k++) {
      e.tt[k] = (1 - lu) * (1 - lv) * p00[k] + lu * (1 - lv) * p01[k] +
                (1 - lu) * lv * p10[k] + lu * lv * p11[k];
      const double a = (1 - lv) * (p01[k] - p00[k]) + lv * (p11[k] - p10[k]);
      const double b = (1 - lu) * (p10[k] - p00[k]) + lu * (p11[k] - p01[k]);
      e.dtdu[k] = a * dlu_du + b * dlv_du;
      e.dtdv[k] = a * dlu_dv + b * dlv_dv;
    }
  }
}

DiffCtx make_ctx(const OracleCamera* cam, const int32_t* faces, const float* uv,
                 const int32_t* uvidx, const double* verts, const double* colors,
                 const double* tex, int tw, int th) {
  DiffCtx d;
  d.W = cam->width;
  d.H = cam->height;
  d.fx = cam->fx;
  d.fy = cam->fy;
  d.cx = cam->cx;
  d.cy = cam->cy;
  for (int i = 0; i < 9; i++) d.R[i] = cam->w2c_R[i];
  for (int i = 0; i < 3; i++) d.t[i] = cam->w2c_t[i];
  d.faces = faces;
  d.uv = uv;
  d.uvidx = uvidx;
  d.verts = verts;
  d.colors = colors;
  d.tex = tex;
  d.tw = tw;
  d.th = th;
  return d;
}

}  // namespace

extern "C" {

// image[p] (H x W x 3) at fixed visibility face_id[p].
int oracle_diff_forward_f64(const OracleCamera* cam, const int32_t* faces, const float* uv,
                            const int32_t* uvidx, const double* verts, const double* colors,
                            const double* tex, int tw, int th, const int32_t* face_id,
                            double* image) {
  DiffCtx d = make_ctx(cam, faces, uv, uvidx, verts, colors, tex, tw, th);
  for (int py = 0; py < d.H; py++)
    for (int px = 0; px < d.W; px++) {
      const int64_t p = int64_t(py) * d.W + px;
      double* o = image + p * 3;
      o[0] = o[1] = o[2] = 0.0;
      const int fid = face_id[p];
      if (fid <= 0) continue;
      PixelEval e;
      eval_pixel(d, fid, px, py, e);
      for (int k = 0; k < 3; k++) o[k] = e.c[k] * e.tt[k];
    }
  return 0;
}

// Accumulates (+=) the gradients of sum(grad_image * image) at fixed visibility.
int oracle_diff_backward_f64(const OracleCamera* cam, const int32_t* faces, const float* uv,
                             const int32_t* uvidx, const double* verts, const double* colors,
                             const double* tex, int tw, int th, const int32_t* face_id,
                             const double* grad_image, double* g_verts, double* g_colors,
                             double* g_tex) {
  DiffCtx d = make_ctx(cam, faces, uv, uvidx, verts, colors, tex, tw, th);
  for (int py = 0; py < d.H; py++)
    for (int px = 0; px < d.W; px++) {
      const int64_t p = int64_t(py) * d.W + px;
      const int fid = face_id[p];
      if (fid <= 0) continue;
      PixelEval e;
      eval_pixel(d, fid, px, py, e);
      const double* G = grad_image + p * 3;
      const int32_t* f = faces + int64_t(fid) * 3;
      const double beta[3] = {e.b0, e.b1, e.b2};
      double gu = 0, gv = 0;
      for (int k = 0; k < 3; k++) {
        const double gc = G[k] * e.tt[k];
        const double gt = MUT(33) ? G[k] : G[k] * e.c[k];
        if (g_colors && colors)
          for (int i = 0; i < 3; i++) g_colors[int64_t(f[i]) * 3 + k] += gc * beta[i];
        if (g_tex && tex) {
          const double w[4] = {(1 - e.lu) * (1 - e.lv), e.lu * (1 - e.lv), (1 - e.lu) * e.lv,
                               e.lu * e.lv};
          const int64_t idx[4] = {int64_t(e.ty0) * tw + e.tx0, int64_t(e.ty0) * tw + e.tx1,
                                  int64_t(e.ty1) * tw + e.tx0, int64_t(e.ty1) * tw + e.tx1};
          for (int j = 0; j < 4; j++) g_tex[idx[j] * 3 + k] += gt * w[j];
        }
        gu += gc * e.dcdu[k] + gt * e.dtdu[k];
        gv += gc * e.dcdv[k] + gt * e.dtdv[k];
      }
      if (!g_verts) continue;
      // b1 = q1/s, b2 = q2/s; dL/dq_j = (dL/db_j - (gu b1 + gv b2)) / s
      const double S = gu * e.b1 + gv * e.b2;
      const double gq[3] = {-S / e.s, (gu - S) / e.s, (gv - S) / e.s};
      double gX[3] = {0, 0, 0}, gY[3] = {0, 0, 0}, gZ[3], gA = 0;
      double ge[3];
      for (int j = 0; j < 3; j++) {
        ge[j] = gq[j] / (e.A * e.Z[j]);
        gZ[j] = MUT(32) ? 0.0 : -gq[j] * e.q[j] / e.Z[j];
        gA -= gq[j] * e.q[j] / e.A;
      }
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
      for (int i = 0; i < 3; i++) {
        const double iz = 1.0 / e.Z[i];
        const double gQ[3] = {gX[i] * d.fx * iz, gY[i] * d.fy * iz,
                              gZ[i] - gX[i] * d.fx * e.Qx[i] * iz * iz -
                                  gY[i] * d.fy * e.Qy[i] * iz * iz};
        for (int c = 0; c < 3; c++)
          g_verts[int64_t(f[i]) * 3 + c] +=
              d.R[0 * 3 + c] * gQ[0] + d.R[1 * 3 + c] * gQ[1] + d.R[2 * 3 + c] * gQ[2];
      }
    }
  return 0;
}

}  // extern "C"
