/*
 * ugu_hip.h -- C ABI of the MI355X (gfx950) rasterizer.
 *
 * Plain C, POD descriptors and pointers only: no Eigen, STL or torch types
 * cross this boundary. The C++ drop-in `ugu::Rasterizer`
 * (include/ugu/renderer/cpu/rasterizer.h) and the Python package `ugu_amd`
 * (ctypes) are both written on top of it.
 *
 * Reference interface each entry point replaces (UGU, unclearness/ugu):
 *   ugu_hip_create / ugu_hip_destroy
 *       Rasterizer::Rasterizer(const RendererCpuOption&) / ~Rasterizer
 *       (src/renderer/cpu/rasterizer.cc:351-356)
 *   ugu_hip_prepare_mesh
 *       Rasterizer::set_mesh + Rasterizer::PrepareMesh
 *       (src/renderer/cpu/rasterizer.cc:77-99; include/ugu/renderer/cpu/renderer.h:49-54)
 *   ugu_hip_render
 *       Rasterizer::Render(color, depth, normal, mask, face_id)
 *       (src/renderer/cpu/rasterizer.cc:105-307, :372-375)
 *   ugu_hip_render_batch
 *       no single reference call: replaces the per-view loop of Render calls
 *       (examples/ex02_renderer.cc:135-169)
 *   ugu_hip_render_gbuffer
 *       no CPU-rasterizer counterpart: the fields of ugu::GBuffer
 *       (include/ugu/renderer/base.h:34-77, filled by the OpenGL renderer in
 *       src/renderer/gl/renderer.cc:501-647) that the rasterizer's resolve
 *       can produce, next to the Render outputs
 *   ugu_hip_prepare_raytracer / ugu_hip_raytrace
 *       Raytracer::PrepareMesh / Raytracer::Render
 *       (src/renderer/cpu/raytracer.cc:88-108, :117-225)
 *   ugu_hip_visibility_test
 *       VisibilityTester::TestVertex (src/texturing/visibility_tester.cc:390-530)
 *       for a batch of vertices against one keyframe
 *   ugu_hip_texture_map
 *       GenerateTextureOnOriginalUv + PaddingSimple of TextureMapping with
 *       kUseOriginalMeshUv (src/texturing/texture_mapper.cc:148-201, :234-345)
 *   ugu_hip_depth_to_u16
 *       the ConvertTo step of Rasterizer::RenderW (src/renderer/cpu/rasterizer.cc:329-344,
 *       include/ugu/image_proc.h:453-520)
 *   ugu_hip_diff_forward / ugu_hip_diff_backward
 *       no reference counterpart: differentiable rendering (gradients of an
 *       image loss with respect to vertex colours, texture and positions)
 *
 * Error convention: every int-returning call returns UGU_HIP_OK (0) or a
 * negative code; ugu_hip_last_error() gives the message. The validation codes
 * -1..-10 correspond one-to-one to the reference's LOGE branches in
 * ValidateAndInitBeforeRender (src/renderer/cpu/util_private.cc:19-66) and
 * RenderW (src/renderer/cpu/rasterizer.cc:331-334).
 *
 * Threading: one context = one HIP stream; calls on one context must not
 * overlap (the reference's Render is likewise not reentrant). Every call is
 * synchronous: outputs are complete when it returns, and it starts only after
 * the work queued on the caller's stream (ugu_hip_set_stream).
 */
#ifndef UGU_HIP_H_
#define UGU_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UGU_HIP_ABI_VERSION 4

typedef struct ugu_hip_ctx ugu_hip_ctx;

enum UguHipStatus {
  UGU_HIP_OK = 0,
  UGU_HIP_ERR_NO_CAMERA = -1,               /* util_private.cc:19-22 */
  UGU_HIP_ERR_MESH_NOT_PREPARED = -2,       /* util_private.cc:23-26 */
  UGU_HIP_ERR_CULLING_NO_FACE_NORMAL = -3,  /* util_private.cc:27-30 */
  UGU_HIP_ERR_TEXTURE_EMPTY = -4,           /* util_private.cc:31-35 */
  UGU_HIP_ERR_TEXTURE_NTH_EMPTY = -5,       /* util_private.cc:36-44 */
  UGU_HIP_ERR_VERTEX_COLOR_EMPTY = -6,      /* util_private.cc:45-50 */
  UGU_HIP_ERR_FACE_NORMAL_EMPTY = -7,       /* util_private.cc:51-55 */
  UGU_HIP_ERR_VERTEX_NORMAL_EMPTY = -8,     /* util_private.cc:56-62 */
  UGU_HIP_ERR_ALL_NULL = -9,                /* util_private.cc:63-66 */
  UGU_HIP_ERR_DEPTH_NULL = -10,             /* rasterizer.cc:331-334 */
  UGU_HIP_ERR_INVALID_ARGUMENT = -20,
  UGU_HIP_ERR_HIP = -21,
  UGU_HIP_ERR_OUT_OF_MEMORY = -22,
  UGU_HIP_ERR_NO_MESH = -23                 /* rasterizer.cc:91-94 */
};

/* Where output (and differentiable-input) pointers live. */
enum UguHipMemory { UGU_HIP_MEMORY_HOST = 0, UGU_HIP_MEMORY_DEVICE = 1 };

/* One RGB8 texture (ObjMaterial::diffuse_tex, an Image3b: rows x cols x 3). */
typedef struct {
  int32_t width;
  int32_t height;
  const uint8_t* data; /* host, height*width*3 bytes, row-major HWC */
} UguTexture;

/* Host-side mesh, the arrays of ugu::Mesh (include/ugu/mesh.h:83-102,137-151).
 * Absent attributes: count 0 and null pointer. Indices are validated. */
typedef struct {
  int32_t num_vertices;
  const float* vertices; /* V x 3 */
  int32_t num_faces;
  const int32_t* faces;        /* F x 3 (vertex_indices) */
  const float* face_normals;   /* F x 3 or null */
  int32_t num_normals;
  const float* normals;        /* N x 3 or null */
  const int32_t* normal_indices; /* F x 3 or null */
  int32_t num_vertex_colors;
  const float* vertex_colors;  /* V x 3, 0..255 scale, or null */
  int32_t num_uv;
  const float* uv;             /* U x 2 or null */
  const int32_t* uv_indices;   /* F x 3 or null */
  const int32_t* material_ids; /* F or null */
  int32_t num_materials;
  const UguTexture* textures;  /* num_materials entries or null */
} UguMeshDesc;

/* Pinhole camera, everything already reduced to float on the host:
 *   w2c_R = camera->w2c().rotation().cast<float>()   (rasterizer.cc:118)
 *   w2c_t = camera->w2c().translation().cast<float>() (rasterizer.cc:119)
 *   c2w_R = c2w.matrix().block<3,3>(0,0).cast<float>() (camera.h:322)
 *   c2w_t = c2w.translation().cast<float>(), the camera position: the ray
 *           origin of the raytracer (camera.h:322, :431-436); unused by the
 *           rasterizer
 * Matrices are row-major. (cx, cy) is the principal point, (fx, fy) the
 * focal length, both in pixels (camera.h:380-387). */
typedef struct {
  int32_t width;
  int32_t height;
  float fx, fy, cx, cy;
  float w2c_R[9];
  float w2c_t[3];
  float c2w_R[9];
  float c2w_t[3];
} UguCamera;

/* RendererCpuOption (include/ugu/renderer/cpu/renderer.h:17-38) with the
 * enum values of include/ugu/renderer/base.h:11-32 and common.h:50-53.
 * As in the reference, oren_nayar_sigma has no effect (RendererCpuOption::
 * CopyTo never copies it) and depth_scale none on the rasterizer's Render;
 * the raytracer multiplies its depth by depth_scale (raytracer.cc:185). */
typedef struct {
  int32_t diffuse_color;   /* 0 kNone, 1 kTexture, 2 kVertex */
  int32_t interp;          /* 0 kNn, 1 kBilinear */
  int32_t shading_normal;  /* 0 kFace, 1 kVertex */
  int32_t diffuse_shading; /* 0 kNone, 1 kLambertian, 2 kOrenNayar */
  float depth_scale;
  int32_t backface_culling;
  float oren_nayar_sigma;
} UguOptions;

/* Output images for one view; null = not requested (as the reference's
 * nullptr arguments). Layouts are interleaved HWC, row-major:
 * color W*H*3 u8 (RGB), depth W*H f32, normal W*H*3 f32, mask W*H u8,
 * face_id W*H i32 (-1 = no face). */
typedef struct {
  uint8_t* color;
  float* depth;
  float* normal;
  uint8_t* mask;
  int32_t* face_id;
} UguOutputs;

/* Tile geometry of the binning rasterizer. tile_w x tile_h must be one of
 * 8x8, 16x16, 32x8. bin_capacity = face slots per tile (grown automatically
 * when a render overflows it). */
typedef struct {
  int32_t tile_w;
  int32_t tile_h;
  int32_t bin_capacity;
  int32_t persistent; /* 1: batched raster pulls (view, tile) items from a queue */
} UguTiling;

/* Per-kernel GPU time of the last render call, from hipEvents recorded on the
 * context's stream (filled only when profiling is enabled). */
typedef struct {
  float clear_ms;
  float setup_bin_ms;
  float raster_ms;
  float total_ms;
  int32_t max_bin_count;
  int32_t retries;
  /* last ugu_hip_diff_backward: the per-face item sort (count, scan,
   * scatter) and the gradient kernel, and the number of (face, view, pixel)
   * items it processed */
  float bwd_sort_ms;
  float bwd_items_ms;
  int32_t bwd_items;
} UguTimings;

/* Context lifetime. device = HIP device ordinal. Returns null on failure. */
ugu_hip_ctx* ugu_hip_create(int device);
void ugu_hip_destroy(ugu_hip_ctx* ctx);
const char* ugu_hip_last_error(const ugu_hip_ctx* ctx);
int ugu_hip_abi_version(void);

/* Stream ordering with the caller: every later call first waits for the
 * work already queued on hip_stream (e.g. torch's current stream; null = the
 * legacy default stream, which is also the initial setting). The kernels run
 * on the context's own stream; calls return when their work is complete. */
int ugu_hip_set_stream(ugu_hip_ctx* ctx, void* hip_stream);

/* set_mesh + PrepareMesh: validates indices and uploads the mesh. The device
 * copy stays valid until the next call (the reference's contract "don't
 * modify mesh after PrepareMesh", renderer.h:52-53). */
int ugu_hip_prepare_mesh(ugu_hip_ctx* ctx, const UguMeshDesc* mesh);

/* Render one view (Rasterizer::Render). memory = where the output pointers
 * live. */
int ugu_hip_render(ugu_hip_ctx* ctx, const UguCamera* camera,
                   const UguOptions* options, const UguOutputs* outputs,
                   int memory);

/* Render n views of possibly different resolutions in one batched pipeline
 * (clear, setup+bin, raster launched back to back, captured once in a
 * hipGraph and replayed). outputs[i] belongs to cameras[i]. Results are
 * bitwise identical to n separate ugu_hip_render calls. */
int ugu_hip_render_batch(ugu_hip_ctx* ctx, const UguCamera* cameras, int n,
                         const UguOptions* options, const UguOutputs* outputs,
                         int memory);

/* Same as ugu_hip_render_batch without the graph: the per-kernel launches are
 * issued directly on the stream (used for per-kernel timing). */
int ugu_hip_render_batch_eager(ugu_hip_ctx* ctx, const UguCamera* cameras,
                               int n, const UguOptions* options,
                               const UguOutputs* outputs, int memory);

/* Allocation-free ugu_hip_render_batch. No reference counterpart: for
 * callers that cannot allocate device memory during the call (code between
 * graph-capture regions, tight serving loops, harnesses that forbid
 * hipMalloc). Outputs are device pointers (UGU_HIP_MEMORY_DEVICE); results
 * are bitwise those of ugu_hip_render_batch. The call allocates nothing:
 * the work memory (bin counters, bins, view descriptors) is one 256 MiB
 * device array per GPU, reserved when the library's code is loaded on that
 * GPU and held until the process ends; the mesh stays the context's prepared
 * copy. Views are rendered in groups whose bins fit that array; a call is
 * rejected (UGU_HIP_ERR_INVALID_ARGUMENT) above 1024 views, and
 * (UGU_HIP_ERR_OUT_OF_MEMORY, nothing launched) when one view's screen tiles
 * need more than the array holds, or when one screen tile of a view holds
 * more faces than the array can give each tile of that view. Calls on one
 * GPU share the array: the library serialises them within a process; any
 * other overlapping use on the GPU would race. The context's tiling is used
 * as by ugu_hip_render_batch; kernels are launched directly (no graph). */
int ugu_hip_render_batch_noalloc(ugu_hip_ctx* ctx, const UguCamera* cameras,
                                 int n, const UguOptions* options,
                                 const UguOutputs* outputs);

/* G-buffer planes of a view (any may be null; memory kind as the outputs).
 * A pixel is covered where the mask output is 255 (face id > 0 and not
 * culled); with u = w1, v = w2 the perspective-correct weights of the face's
 * second and third vertex (the colourizers' u, v) and w = (1 - u) - v:
 *   bary       (u, v, (1 - u) - v)                       as GBuffer::bary
 *   uv         (s, t, (1 - s) - t), (s, t) = w uv0 + u uv1 + v uv2, the
 *              texture coordinate the texture colourizers sample (zero if the
 *              mesh has no uv)
 *   pos_wld    w P0 + u P1 + v P2 (world position on the face)
 *   pos_cam    w2c_t + w2c_R pos_wld
 *   normal_wld the world-space shading normal (face normal, or the
 *              interpolated vertex normal as the reference's Render uses it;
 *              the normal output is w2c_R times this)
 *   geo_id     0 (one geometry)
 * Every other pixel is 0 (geo_id -1). Float planes are H x W x 3. */
typedef struct {
  float* bary;
  float* uv;
  float* pos_wld;
  float* pos_cam;
  float* normal_wld;
  int32_t* geo_id;
} UguGBuffer;

/* ugu_hip_render_batch plus G-buffer planes: gbuffers[i] belongs to
 * cameras[i]; the Render outputs are bitwise the same as without them. */
int ugu_hip_render_gbuffer(ugu_hip_ctx* ctx, const UguCamera* cameras, int n,
                           const UguOptions* options, const UguOutputs* outputs,
                           const UguGBuffer* gbuffers, int memory);

/* ---- raytracer (the reference's second RendererCpu) --------------------
 * ugu_hip_prepare_raytracer builds a bounding volume hierarchy over the
 * prepared mesh (call after ugu_hip_prepare_mesh; the mesh must not change).
 * ugu_hip_raytrace renders like Raytracer::Render: per pixel a ray from c2w_t
 * along ray_w(x, y) finds its closest face with the reference's float64
 * Moller-Trumbore test (src/util/geom_util.cc:60-115, epsilon 1e-8, standard
 * barycentric conversion), closest = smallest float t, equal t -> lowest face
 * id; then culling (face normal . ray > 0 -> background), face_id, mask,
 * depth = z of the camera-space hit * depth_scale, normal and colour with
 * the raytracer's (u, v) conventions (raytracer.cc:160-218). outputs[i]
 * belongs to cameras[i]. */
int ugu_hip_prepare_raytracer(ugu_hip_ctx* ctx);
int ugu_hip_raytrace(ugu_hip_ctx* ctx, const UguCamera* cameras, int n,
                     const UguOptions* options, const UguOutputs* outputs, int memory);

/* Tiling control and the autotuner. ugu_hip_autotune renders the given views
 * with every tile size, sizes the bin capacity from the measured maximum bin
 * occupancy, keeps the fastest and returns it in *chosen. */
int ugu_hip_set_tiling(ugu_hip_ctx* ctx, const UguTiling* tiling);
int ugu_hip_get_tiling(const ugu_hip_ctx* ctx, UguTiling* tiling);
int ugu_hip_autotune(ugu_hip_ctx* ctx, const UguCamera* cameras, int n,
                     const UguOptions* options, int repeats,
                     UguTiling* chosen);

/* ugu_hip_autotune with a persistent choice: the tiling is looked up in the
 * text file cache_path under a key of the device (architecture, CU count),
 * the prepared mesh's face count, the options that change the work and the
 * views' sizes; on a hit it is applied without rendering (*from_cache = 1),
 * otherwise the views are tuned and one line is appended to the file (an
 * unwritable file only loses the persistence). Tiling changes speed only:
 * results are the same for every tiling, and a bin capacity that a later
 * render outgrows is grown as usual. No reference counterpart. */
int ugu_hip_autotune_cached(ugu_hip_ctx* ctx, const UguCamera* cameras, int n,
                            const UguOptions* options, int repeats, const char* cache_path,
                            UguTiling* chosen, int* from_cache);

/* Profiling (hipEvents around each kernel on the context's stream). */
int ugu_hip_enable_profiling(ugu_hip_ctx* ctx, int enable);
int ugu_hip_get_timings(const ugu_hip_ctx* ctx, UguTimings* timings);

/* RenderW's depth conversion: out[i] = uint16(1.0 * (double)depth[i]),
 * truncating, as ConvertTo (image_proc.h:453-520). memory as above. */
int ugu_hip_depth_to_u16(ugu_hip_ctx* ctx, const float* depth, int64_t n,
                         uint16_t* out, int memory);

/* ---- visibility test (texturing consumer of the raytracer) -------------
 * VisibilityTesterOption (include/ugu/texturing/visibility_tester.h:38-57),
 * the fields TestVertex reads. interp: 0 kNn, 1 kBilinear. */
typedef struct {
  float backface_culling_angle_rad_th;
  int32_t use_mask;
  int32_t use_depth;
  int32_t interp;
  float max_viewing_angle;
  float max_distance_from_camera_to_vertex;
  float max_depth_difference;
  float eps;
} UguVisibilityOptions;

/* A keyframe (visibility_tester.h:229-246): the camera (w2c_R / w2c_t = the
 * float cast of w2c()'s linear part and translation; c2w_R, c2w_t, intrinsics
 * as for Render), its colour image (width x height x 3 RGB8, the size the
 * projection is bounded by) and optional depth (float) and mask (u8) images
 * of the same size. Host pointers. */
typedef struct {
  UguCamera camera;
  int32_t width, height;
  const uint8_t* color;
  const float* depth;
  const uint8_t* mask;
} UguKeyframe;

/* TestVertex's result for one item (VertexInfoPerKeyframe, :61-77):
 * visible = TestVertex's return value; the other fields as TestVertex sets
 * them when visible (projected position, viewing angle in radians, distance
 * = the ray parameter of the closest hit or FLT_MAX, colour as floats,
 * intensity = Color2Gray(colour)). */
typedef struct {
  int32_t visible;
  float projected[2];
  float viewing_angle;
  float distance;
  float color[3];
  float intensity;
} UguVertexVisibility;

/* TestVertex for n items against one keyframe, on the hierarchy built by
 * ugu_hip_prepare_raytracer over the context's mesh (the tester's vertices
 * and faces). Item i is the point item_pos[i] (= vertex + offset; the
 * facewise test moves each corner 1e-3 towards its face centre), tested with
 * the vertex position item_vertex[i] (occlusion distance) and the vertex
 * normal item_normal[i]. Host arrays of n x 3 floats; out has n entries.
 * Items are independent: the caller assembles VisibilityInfo from them. */
int ugu_hip_visibility_test(ugu_hip_ctx* ctx, const UguKeyframe* keyframe,
                            const UguVisibilityOptions* options, int n,
                            const float* item_pos, const float* item_vertex,
                            const float* item_normal, UguVertexVisibility* out);

/* ---- texture mapping on the mesh's own uv (texturing consumer) --------
 * One entry per face: keyframe = index into the colour images (-1: the face
 * has no visible keyframe and is skipped), src = its projected triangle in
 * that image (FaceInfoPerKeyframe::projected_tri), dst = its uv triangle in
 * texel coordinates (u * tex_w - 0.5, (1 - v) * tex_h - 0.5). */
typedef struct {
  int32_t keyframe;
  float src[6];
  float dst[6];
} UguTexFace;

/* Rasterizes every face's dst triangle into a tex_w x tex_h RGB8 texture,
 * sampling its src triangle of the keyframe colour bilinearly; where dst
 * triangles overlap the face with the larger index wins (the reference's
 * face loop order); then PaddingSimple with padding_kernel. texture
 * (tex_h x tex_w x 3) and mask (tex_h x tex_w, 255 = set) are host outputs;
 * images are RGB8 host images (UguTexture). */
int ugu_hip_texture_map(ugu_hip_ctx* ctx, const UguTexture* images, int n_images,
                        const UguTexFace* faces, int n_faces, int tex_w, int tex_h,
                        int padding_kernel, uint8_t* texture, uint8_t* mask);

/* ---- differentiable rendering (no reference counterpart) ----------------
 * Uses the prepared mesh's topology (faces, uv, uv_indices) with parameters
 * given as device float32 arrays:
 *   vertices      V x 3   positions (visibility is computed from these)
 *   vertex_colors V x 3   or null
 *   texture       Th x Tw x 3 (row-major, same orientation as an Image3b)
 *                 or null
 * Image model per covered pixel (fid > 0 and not culled, as the reference's
 * shading rule), with u = w1, v = w2 the perspective-correct barycentrics:
 *   c   = ((1-u)-v)*C0 + u*C1 + v*C2                     (vertex colour term)
 *   t   = bilinear(texture, uv(u, v))                    (texture term)
 *   img = c*t if both present, else whichever is present; 0 elsewhere.
 * Face normals for the backface test are recomputed from `vertices`. */
typedef struct {
  const float* vertices;
  const float* vertex_colors;
  const float* texture;
  int32_t tex_width;
  int32_t tex_height;
  int32_t backface_culling;
} UguDiffParams;

/* images[i]: device H_i x W_i x 3 f32; face_ids[i]: device H_i x W_i i32 (the
 * visibility kept for the backward pass). */
int ugu_hip_diff_forward(ugu_hip_ctx* ctx, const UguCamera* cameras, int n,
                         const UguDiffParams* params, float* const* images,
                         int32_t* const* face_ids);

/* Accumulates (+=) dL/d(vertices), dL/d(vertex_colors), dL/d(texture) given
 * dL/d(images) (device, same layout as images) and the face_ids of the
 * forward pass, at fixed visibility. Null gradient pointers are skipped. */
int ugu_hip_diff_backward(ugu_hip_ctx* ctx, const UguCamera* cameras, int n,
                          const UguDiffParams* params,
                          const float* const* grad_images,
                          const int32_t* const* face_ids, float* grad_vertices,
                          float* grad_vertex_colors, float* grad_texture);

#ifdef __cplusplus
}
#endif

#endif /* UGU_HIP_H_ */
