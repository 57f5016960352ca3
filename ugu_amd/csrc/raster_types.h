// Device-side descriptors shared by the HIP kernels and the host launcher.
// Plain structs, passed by pointer (views) or by value (mesh, bins, options).
#pragma once

#include <stdint.h>

namespace ugu_hip {

// Per-view constants. The camera part is the UguCamera of ugu_hip.h; the rest
// places the view's tiles in the global tile array and its outputs.
struct DevView {
  int32_t W, H;
  float fx, fy, cx, cy;
  float w2c_R[9];
  float w2c_t[3];
  float c2w_R[9];
  float c2w_t[3];     // camera position: the raytracer's ray origin
  int32_t tiles_x, tiles_y;
  int32_t tile_base;  // first global tile of this view
  int32_t pad0;
  uint8_t* color;
  float* depth;
  float* normal;
  uint8_t* mask;
  int32_t* face_id;
  float* image;        // differentiable forward: H x W x 3 f32
  const float* grad;   // differentiable backward: H x W x 3 f32
  // G-buffer planes (null: not written)
  float* bary;
  float* uvo;
  float* pos_wld;
  float* pos_cam;
  float* nrm_wld;
  int32_t* geo_id;
};

struct TexInfo {
  int64_t offset;  // byte offset (u8 textures) into the texture pool
  int32_t w, h;
};

struct DevMesh {
  int32_t V, F;
  const float* verts;       // V x 3
  const int32_t* faces;     // F x 3
  const float* fnormals;    // F x 3 or null
  const float* normals;     // N x 3 or null
  const int32_t* nidx;      // F x 3 or null
  const float* vcolors;     // V x 3 or null
  const float* uv;          // U x 2 or null
  const int32_t* uvidx;     // F x 3 or null
  const int32_t* matid;     // F or null
  const TexInfo* tex;       // per material
  TexInfo tex0;             // the first material's entry (by value: a one-material
  int32_t n_tex;            // mesh needs no per-pixel load of it); materials
  const uint8_t* texdata;   // u8 RGB texture pool
  const float4* soup;       // F x 3: the face's three vertex positions and its
                            // face normal (zero when the mesh has none),
                            // packed (x0 y0 z0 x1)(y1 z1 x2 y2)(z2 nx ny nz)
  const int32_t* blk_off;   // per 256-face setup block: first entry of its
                            // distinct vertices in blk_vid (nblocks + 1), or null
  const int32_t* blk_vid;   // distinct vertex ids of each block
  const uint32_t* face_loc; // per face: its corners' indices in the block's
                            // list, 10 bits each
  const float4* shade;      // F x 4 or null: per-face shading record, the corners'
                            // uvs and vertex normals and the material id, packed
                            // (u0 v0 u1 v1)(u2 v2 n0x n0y)(n0z n1x n1y n1z)
                            // (n2x n2y n2z material-id bits)
  const float* ftex;        // differentiable path: f32 RGB texture or null
  int32_t ftex_w, ftex_h;
};

struct DevOptions {
  int32_t diffuse_color;    // 0 none, 1 texture, 2 vertex
  int32_t interp;           // 0 nn, 1 bilinear
  int32_t shading_normal;   // 0 face, 1 vertex
  int32_t diffuse_shading;  // 0 none, 1 lambertian, 2 oren-nayar
  int32_t culling;
  float on_A, on_B;         // Oren-Nayar constants for sigma 0.3
  float depth_scale;        // raytracer depth factor (the rasterizer ignores it)
};

// Bounding volume hierarchy of the raytracer. Every inner node stores the
// boxes of its two children; a child reference is an inner node index (>= 0,
// always greater than the parent's), ~leaf for a leaf (a run of triangle
// records) or kBvhEmpty. Boxes are widened outward so that the float slab test
// never rejects a box that holds a float64 Moller-Trumbore hit.
struct BvhNode {
  float4 a_lo;  // child A box minimum, w = child A reference (int bits)
  float4 a_hi;  // child A box maximum, w unused
  float4 b_lo;  // child B
  float4 b_hi;
};
constexpr int32_t kBvhEmpty = INT32_MIN;
constexpr int kBvhStack = 64;  // traversal stack entries; the builder keeps depth below

struct DevBvh {
  const BvhNode* nodes;
  const int2* leaves;   // (first triangle record, count)
  const float4* tris;   // 3 per record: (v0, face id bits) (v1, 0) (v2, 0)
};

// Visibility test options (UguVisibilityOptions)
struct DevVisOpt {
  float th;  // backface_culling_angle_rad_th
  int32_t use_mask, use_depth, interp;
  float max_angle, max_dist, max_ddiff, eps;
};

struct DevBins {
  int32_t* counts;   // per global tile: front / back segment counters, minimum
                     // skip depth of the back segment (float bits), view index
  int2* bins;        // global tile t owns bins[t*cap .. t*cap+cap): front
                     // segment from the start, back segment from the end;
                     // entry = (face id, bits of its skip depth)
  int32_t cap;
  int32_t total_tiles;
  int32_t* status;   // [0] max bin count seen; [kQueueStride (q + 1)] head of
                     // work queue q (kStatusBytes in all)
  const int32_t* tile_base;  // n_views + 1 prefix of tiles per view
  int32_t n_views;
};

// Work queues of the persistent rasterizer: run r (kQueueRun consecutive
// tiles) belongs to queue r % kQueues; the queue heads sit 128 B apart so
// their atomics spread over the memory channels instead of queueing on one
// address.
constexpr int kQueues = 8;
constexpr int kQueueStride = 32;  // int32 words between queue heads
constexpr int kStatusBytes = 4 * kQueueStride * (kQueues + 1);

// Gradient targets of the differentiable backward pass.
struct DevGrads {
  float* verts;    // V x 3 or null
  float* colors;   // V x 3 or null
  float* texture;  // Th x Tw x 3 or null
};

}  // namespace ugu_hip
