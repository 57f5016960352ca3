// HIP kernels of the texture mapper's kUseOriginalMeshUv path
// (src/texturing/texture_mapper.cc:234-345 GenerateTextureOnOriginalUv /
// RasterizeTriangle, :148-201 PaddingSimple).
//
// Every face copies its best keyframe's projected triangle into its uv
// triangle of the texture. The reference loops over faces in order, so where
// uv triangles overlap the face with the largest index wins: k_tex_owner
// records, per texel, the largest face that covers it (atomicMax), and
// k_tex_write lets only that face write. One wave per face; its lanes stride
// over the texels of the face's bounding box. k_tex_pad is PaddingSimple, one
// thread per texel, reading the unpadded texture and mask.

#include "device_common.h"

#pragma clang fp contract(off)

namespace ugu_hip {

struct DevTexFace {
  int32_t keyframe;  // index of the source image, -1: face not textured
  float src[6];      // projected triangle in the source image
  float dst[6];      // target triangle in texel coordinates
};

struct DevImage3b {
  int32_t w, h;
  const uint8_t* data;
};

// EdgeFunction (texture_mapper.cc:143-146)
UGU_DEV float tm_edge(float ax, float ay, float bx, float by, float cx, float cy) {
  return (cx - ax) * (by - ay) - (cy - ay) * (bx - ax);
}

// The set-up of RasterizeTriangle (:238-258): inverse area and the clamped
// bounding box of the target triangle.
UGU_DEV void tm_setup(const DevTexFace& f, int cols, int rows, float& inv_area, int& x0, int& x1,
                      int& y0, int& y1) {
  const float* t = f.dst;
  float area = tm_edge(t[0], t[1], t[2], t[3], t[4], t[5]);
  if (fabsf(area) < FLT_MIN) area = area > 0 ? FLT_MIN : -FLT_MIN;
  inv_area = 1.0f / area;
  x0 = f2i_x86(min3_std(t[0], t[2], t[4]) - 1);
  x0 = max(0, min(x0, cols - 1));
  x1 = f2i_x86(max3_std(t[0], t[2], t[4]) + 1);
  x1 = max(0, min(x1, cols - 1));
  y0 = f2i_x86(min3_std(t[1], t[3], t[5]) - 1);
  y0 = max(0, min(y0, rows - 1));
  y1 = f2i_x86(max3_std(t[1], t[3], t[5]) + 1);
  y1 = max(0, min(y1, rows - 1));
}

// barycentric weights of texel (x, y) in the target triangle; false outside
UGU_DEV bool tm_weights(const DevTexFace& f, float inv_area, int x, int y, float& w0, float& w1,
                        float& w2) {
  const float* t = f.dst;
  const float px = static_cast<float>(x), py = static_cast<float>(y);
  w0 = tm_edge(t[2], t[3], t[4], t[5], px, py) * inv_area;
  w1 = tm_edge(t[4], t[5], t[0], t[1], px, py) * inv_area;
  w2 = tm_edge(t[0], t[1], t[2], t[3], px, py) * inv_area;
  return !(w0 < 0 || w1 < 0 || w2 < 0);
}

__global__ void k_tex_fill(int32_t* owner, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) owner[i] = -1;
}

__global__ __launch_bounds__(64) void k_tex_owner(const DevTexFace* __restrict__ faces, int n,
                                                  int cols, int rows, int32_t* __restrict__ owner) {
  const int fi = blockIdx.x;
  if (fi >= n) return;
  const DevTexFace f = faces[fi];
  if (f.keyframe < 0) return;
  float inv_area;
  int x0, x1, y0, y1;
  tm_setup(f, cols, rows, inv_area, x0, x1, y0, y1);
  const int bw = x1 - x0 + 1;
  const int64_t cnt = int64_t(bw) * (y1 - y0 + 1);
  for (int64_t k = threadIdx.x; k < cnt; k += 64) {
    const int x = x0 + int(k % bw), y = y0 + int(k / bw);
    float w0, w1, w2;
    if (tm_weights(f, inv_area, x, y, w0, w1, w2)) atomicMax(&owner[int64_t(y) * cols + x], fi);
  }
}

__global__ __launch_bounds__(64) void k_tex_write(const DevTexFace* __restrict__ faces, int n,
                                                  const DevImage3b* __restrict__ images, int cols,
                                                  int rows, const int32_t* __restrict__ owner,
                                                  uint8_t* __restrict__ tex,
                                                  uint8_t* __restrict__ mask) {
  const int fi = blockIdx.x;
  if (fi >= n) return;
  const DevTexFace f = faces[fi];
  if (f.keyframe < 0) return;
  const DevImage3b img = images[f.keyframe];
  float inv_area;
  int x0, x1, y0, y1;
  tm_setup(f, cols, rows, inv_area, x0, x1, y0, y1);
  const int bw = x1 - x0 + 1;
  const int64_t cnt = int64_t(bw) * (y1 - y0 + 1);
  for (int64_t k = threadIdx.x; k < cnt; k += 64) {
    const int x = x0 + int(k % bw), y = y0 + int(k / bw);
    const int64_t ti = int64_t(y) * cols + x;
    if (owner[ti] != fi) continue;
    float w0, w1, w2;
    tm_weights(f, inv_area, x, y, w0, w1, w2);
    // src_pos = w0 s0 + w1 s1 + w2 s2, then BilinearInterpolation
    // (include/ugu/util/image_util.h:155-193) on the keyframe colour
    const float sx = w0 * f.src[0] + w1 * f.src[2] + w2 * f.src[4];
    const float sy = w0 * f.src[1] + w1 * f.src[3] + w2 * f.src[5];
    int px0 = f2i_x86(floorf(sx)), py0 = f2i_x86(floorf(sy));
    int px1 = px0 + 1, py1 = py0 + 1;
    if (px0 < 0) px0 = 0;
    if (py0 < 0) py0 = 0;
    if (img.w <= px1) px1 = img.w - 1;
    if (img.h <= py1) py1 = img.h - 1;
    const float lu = sx - static_cast<float>(px0), lv = sy - static_cast<float>(py0);
    // reads stay inside the image (the reference reads outside when the
    // projected triangle leaves it, which a visible face's cannot)
    const int cx0 = min(max(px0, 0), img.w - 1), cx1 = min(max(px1, 0), img.w - 1);
    const int cy0 = min(max(py0, 0), img.h - 1), cy1 = min(max(py1, 0), img.h - 1);
    const uint8_t* c00 = img.data + (int64_t(cy0) * img.w + cx0) * 3;
    const uint8_t* c01 = img.data + (int64_t(cy0) * img.w + cx1) * 3;
    const uint8_t* c10 = img.data + (int64_t(cy1) * img.w + cx0) * 3;
    const uint8_t* c11 = img.data + (int64_t(cy1) * img.w + cx1) * 3;
    for (int c = 0; c < 3; c++) {
      const float v = (1.0f - lu) * (1.0f - lv) * static_cast<float>(c00[c]) +
                      lu * (1.0f - lv) * static_cast<float>(c01[c]) +
                      (1.0f - lu) * lv * static_cast<float>(c10[c]) +
                      lu * lv * static_cast<float>(c11[c]);
      tex[ti * 3 + c] = f2u8_x86(v);
    }
    mask[ti] = 255;
  }
}

// PaddingSimple (texture_mapper.cc:148-201): an unset texel inside the
// kernel margin takes the truncated average of the set texels of its
// kernel x kernel window (read from the unpadded texture).
__global__ void k_tex_pad(const uint8_t* __restrict__ tex0, const uint8_t* __restrict__ mask0,
                          int cols, int rows, int hk, uint8_t* __restrict__ tex,
                          uint8_t* __restrict__ mask) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(cols) * rows) return;
  const int x = int(i % cols), y = int(i / cols);
  tex[i * 3] = tex0[i * 3];
  tex[i * 3 + 1] = tex0[i * 3 + 1];
  tex[i * 3 + 2] = tex0[i * 3 + 2];
  mask[i] = mask0[i];
  if (x < hk || x >= cols - hk || y < hk || y >= rows - hk || mask0[i] == 255) return;
  float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
  int cnt = 0;
  for (int jj = -hk; jj <= hk; jj++)
    for (int ii = -hk; ii <= hk; ii++) {
      const int64_t k = int64_t(y + jj) * cols + (x + ii);
      if (mask0[k] == 255) {
        s0 += tex0[k * 3];
        s1 += tex0[k * 3 + 1];
        s2 += tex0[k * 3 + 2];
        cnt++;
      }
    }
  if (cnt == 0) return;
  // average_color[c] /= valid_pixels.size(): float / size_t -> float division
  const float n = static_cast<float>(cnt);
  mask[i] = 255;
  tex[i * 3] = f2u8_x86(s0 / n);
  tex[i * 3 + 1] = f2u8_x86(s1 / n);
  tex[i * 3 + 2] = f2u8_x86(s2 / n);
}

int launch_texture_map(hipStream_t st, const void* faces, int n, const void* images, int cols,
                       int rows, int hk, int32_t* owner, uint8_t* tex0, uint8_t* mask0,
                       uint8_t* tex, uint8_t* mask) {
  const int64_t texels = int64_t(cols) * rows;
  const DevTexFace* f = static_cast<const DevTexFace*>(faces);
  hipLaunchKernelGGL(k_tex_fill, dim3(unsigned((texels + 255) / 256)), dim3(256), 0, st, owner,
                     texels);
  if (n > 0) {
    hipLaunchKernelGGL(k_tex_owner, dim3(n), dim3(64), 0, st, f, n, cols, rows, owner);
    hipLaunchKernelGGL(k_tex_write, dim3(n), dim3(64), 0, st, f, n,
                       static_cast<const DevImage3b*>(images), cols, rows, owner, tex0, mask0);
  }
  hipLaunchKernelGGL(k_tex_pad, dim3(unsigned((texels + 255) / 256)), dim3(256), 0, st, tex0,
                     mask0, cols, rows, hk, tex, mask);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace ugu_hip
