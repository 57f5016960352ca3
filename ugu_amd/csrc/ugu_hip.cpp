// Host side of the C ABI declared in include/ugu_hip.h.
//
// One context owns one HIP stream, the device copy of the prepared mesh and
// the work buffers of the binning pipeline (projected vertices, per-tile bin
// counters and bins, view descriptors). Every call is synchronous.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ugu_hip.h"
#include "bvh_build.h"
#include "raster_types.h"

namespace ugu_hip {
int launch_clear(hipStream_t st, int32_t* counts, int32_t n, int32_t* status,
                 const int32_t* tile_base, int32_t n_views);
int launch_shade_rec(hipStream_t st, const DevMesh& m, float4* out);
int launch_face_soup(hipStream_t st, const float* verts, const int32_t* faces,
                     const float* fnormals, int32_t F, int compute_normals, float4* soup,
                     float* fn_out);
int launch_setup_bin(hipStream_t st, int tw, int th, const DevView* views,
                     int n_views, const DevMesh& mesh,
                     const DevBins& bins);
int launch_raster(hipStream_t st, int tw, int th, int mode, const DevView* views,
                  const DevMesh& mesh, const DevBins& bins,
                  const DevOptions& opt, int persistent_blocks);
int raster_blocks_per_cu(int tw, int th, int mode);
int64_t backward_scratch_ints(int32_t F);
int launch_backward(hipStream_t st, const DevView* views, int n_views,
                    int max_pixels, int64_t total_pixels, const DevMesh& mesh,
                    const DevGrads& g, int4* items, int32_t* scratch, hipEvent_t before_items);
int launch_depth_to_u16(hipStream_t st, const float* d, int64_t n, uint16_t* out);
int launch_raytrace(hipStream_t st, const DevView* views, int n_views, int total_tiles,
                    const DevBvh& bvh, const DevMesh& m, const DevOptions& o);
int launch_texture_map(hipStream_t st, const void* faces, int n, const void* images, int cols,
                       int rows, int hk, int32_t* owner, uint8_t* tex0, uint8_t* mask0,
                       uint8_t* tex, uint8_t* mask);
int launch_visibility(hipStream_t st, const DevView& v, const uint8_t* color, const float* depth,
                      const uint8_t* mask, const DevVisOpt& o, int n, const float* pos,
                      const float* vert, const float* nrm, const DevBvh& bvh, void* out);
}  // namespace ugu_hip

using namespace ugu_hip;

// Work memory of ugu_hip_render_batch_noalloc: one copy per GPU, allocated
// with the code object and held until the process ends. Host code never
// passes this name to a kernel or a copy: its host-side value is not a device
// address. It uses the address hipGetSymbolAddress returns for the current
// GPU (looked up once per GPU in ugu_hip_create, kept in g_scratch).
constexpr size_t kScratchBytes = size_t(256) << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kNoAllocMaxViews = 1024;
constexpr int kMaxDevices = 64;
// Layout of the scratch array (offsets 256-B aligned): status words (max bin
// count, work-queue heads), the view descriptors, the views' first global
// tiles, then per render group the tiles' counters and the bins.
constexpr size_t kNaViewsOff = (size_t(kStatusBytes) + 255) / 256 * 256;
constexpr size_t kNaTileBaseOff =
    (kNaViewsOff + size_t(kNoAllocMaxViews) * sizeof(DevView) + 255) / 256 * 256;
constexpr size_t kNaCountsOff =
    (kNaTileBaseOff + size_t(kNoAllocMaxViews + 1) * 4 + 255) / 256 * 256;
static_assert(kNaCountsOff < kScratchBytes / 2, "descriptor area fits the scratch");

std::mutex g_scratch_mu;                          // guards the two tables below
uint8_t* g_scratch[kMaxDevices] = {};             // device address of `scratch` per GPU
hipError_t g_scratch_err[kMaxDevices] = {};       // lookup result per GPU
std::mutex g_scratch_use[kMaxDevices];            // one allocation-free render per GPU at a time

// With `device` current: looks up this GPU's copy of `scratch` (each GPU
// holds its own copy at its own address). Repeated calls keep the first
// successful result.
void init_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  if (e == hipSuccess && p != nullptr) g_scratch[device] = static_cast<uint8_t*>(p);
}

uint8_t* scratch_of(int device, hipError_t* err) {
  if (device < 0 || device >= kMaxDevices) {
    *err = hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  *err = g_scratch_err[device];
  return g_scratch[device];
}

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  // returns -1 on allocation failure, 1 when (re)allocated, 0 when unchanged
  int ensure(size_t n) {
    if (n == 0) n = 16;
    if (n <= bytes) return 0;
    release();
    if (hipMalloc(&p, n) != hipSuccess) {
      p = nullptr;
      bytes = 0;
      return -1;
    }
    bytes = n;
    return 1;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

struct GraphKey {
  int n, total_tiles, V, F, tw, th, persistent_blocks, mode, cap;
  int64_t buf_gen, mesh_gen;
  DevOptions opt;
  const void* verts;
  const void* vcolors;
  const void* ftex;
  int ftex_w, ftex_h;
  bool operator==(const GraphKey& o) const { return std::memcmp(this, &o, sizeof(*this)) == 0; }
};

struct GraphEntry {
  GraphKey key;
  hipGraphExec_t exec;
};

constexpr size_t kMaxGraphs = 16;

}  // namespace

struct ugu_hip_ctx {
  int device = 0;
  hipStream_t own = nullptr;     // every kernel and copy of the context runs here
  hipStream_t stream = nullptr;  // == own
  hipStream_t user = nullptr;    // caller's stream the calls are ordered after (null = legacy default)
  hipEvent_t ev_user = nullptr;
  std::string err;

  // prepared mesh
  bool prepared = false;
  int64_t mesh_gen = 0;
  int32_t V = 0, F = 0, N = 0, U = 0, M = 0;
  bool has_fn = false, has_normals = false, has_nidx = false, has_vc = false,
       has_uv = false, has_uvidx = false, has_matid = false;
  std::vector<int> tex_empty;
  DevBuf verts, faces, faces_orig, fnormals, normals, nidx, vcolors, uv, uvidx, matid,
      texinfo, texdata;

  // work buffers
  DevBuf counts, bins, status, views, tile_base, stage_out, diff_fn;
  DevBuf bwd_items;  // (face, view, pixel) items of the backward pass, 16 B each
  DevBuf bwd_scratch;  // per-face counts and offsets of the backward pass
  DevBuf soup, diff_soup;  // 48-byte face records (prepared mesh / differentiable path)
  DevBuf shade;            // 64-byte per-face shading records (DevMesh::shade)
  TexInfo tex0{};          // host copy of the first material's TexInfo
  DevBuf blk_off, blk_vid, face_loc;           // setup blocks' vertex lists (prepared mesh)
  DevBuf blk_off_o, blk_vid_o, face_loc_o;     // the same for the original face array
  DevBuf bvh_nodes, bvh_leaves, bvh_tris;      // raytracer hierarchy
  DevBuf vis_items, vis_images, vis_out;       // visibility test inputs / results
  DevBuf tm_faces, tm_images, tm_out;          // texture mapping inputs / work / outputs
  bool rt_ready = false;
  int64_t rt_mesh_gen = -1;  // mesh generation the hierarchy was built for
  int64_t buf_gen = 0;
  void* h_views = nullptr;
  size_t h_views_bytes = 0;
  int32_t* h_status = nullptr;
  std::vector<int32_t> h_tile_base;

  UguTiling tiling{16, 16, 1024, 1};
  int cu_count = 0;
  std::string arch;  // gcnArchName of the device (tuning-cache key)
  std::vector<GraphEntry> graphs;

  bool profile = false;
  UguTimings tm{};
  hipEvent_t ev[6] = {};

  // allocation-free render: host copies of one group's view descriptors and
  // first tiles (sized at create), and the per-tile bin capacity the last
  // calls needed (groups are formed so every tile gets at least this much)
  std::vector<DevView> na_views;
  std::vector<int32_t> na_tile_base;
  int64_t na_cap_hint = 0;
};

namespace {

int fail(ugu_hip_ctx* c, int code, const char* msg) {
  c->err = msg;
  return code;
}

int hip_fail(ugu_hip_ctx* c, hipError_t e, const char* where) {
  c->err = std::string(where) + ": " + hipGetErrorString(e);
  return UGU_HIP_ERR_HIP;
}

#define UGU_HIP_CHECK(ctx, expr)                                   \
  do {                                                             \
    hipError_t e_ = (expr);                                        \
    if (e_ != hipSuccess) return hip_fail((ctx), e_, #expr);       \
  } while (0)

// Orders the context's stream after everything already queued on the
// caller's stream (torch's current stream, or the legacy default stream):
// without it, kernels of this library could read tensors whose producing
// kernels have not run yet. Calls are host-synchronous, so work the caller
// queues afterwards sees the results.
int order_after_user(ugu_hip_ctx* c) {
  UGU_HIP_CHECK(c, hipEventRecord(c->ev_user, c->user));
  UGU_HIP_CHECK(c, hipStreamWaitEvent(c->stream, c->ev_user, 0));
  return 0;
}

int ensure(ugu_hip_ctx* c, DevBuf& b, size_t n) {
  int r = b.ensure(n);
  if (r < 0) return fail(c, UGU_HIP_ERR_OUT_OF_MEMORY, "device allocation failed");
  if (r > 0) c->buf_gen++;
  return 0;
}

void drop_graphs(ugu_hip_ctx* c) {
  for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
  c->graphs.clear();
}

int upload(ugu_hip_ctx* c, DevBuf& b, const void* src, size_t n) {
  if (src == nullptr || n == 0) {
    b.release();
    return 0;
  }
  if (b.ensure(n) < 0) return fail(c, UGU_HIP_ERR_OUT_OF_MEMORY, "device allocation failed");
  UGU_HIP_CHECK(c, hipMemcpy(b.p, src, n, hipMemcpyHostToDevice));
  return 0;
}

// Renumbers the `count` rows referenced by idx[0..n) in order of first use
// (unreferenced rows last): *remapped receives the renumbered indices and
// *order the old row id of each new id.
void relabel(const int32_t* idx, int64_t n, int32_t count, std::vector<int32_t>* remapped,
             std::vector<int32_t>* order) {
  std::vector<int32_t> new_of_old(size_t(count), -1);
  order->clear();
  order->reserve(size_t(count));
  remapped->resize(size_t(n));
  for (int64_t i = 0; i < n; i++) {
    int32_t& id = new_of_old[size_t(idx[i])];
    if (id < 0) {
      id = static_cast<int32_t>(order->size());
      order->push_back(idx[i]);
    }
    (*remapped)[size_t(i)] = id;
  }
  for (int32_t v = 0; v < count; v++)
    if (new_of_old[size_t(v)] < 0) order->push_back(v);
}

std::vector<float> gather_rows(const float* src, int32_t rows, int cols,
                               const std::vector<int32_t>& order) {
  std::vector<float> out(size_t(rows) * cols);
  for (int32_t r = 0; r < rows; r++)
    for (int k = 0; k < cols; k++)
      out[size_t(r) * cols + k] = src[size_t(order[size_t(r)]) * cols + k];
  return out;
}

// Distinct vertices of every block of 256 faces (the setup kernel's blocks)
// and each face's corners as 10-bit indices into its block's list.
int upload_block_lists(ugu_hip_ctx* c, const int32_t* faces, int64_t F, DevBuf* off,
                       DevBuf* vid, DevBuf* loc) {
  if (F <= 0) {
    off->release();
    vid->release();
    loc->release();
    return 0;
  }
  const int64_t nb = (F + 255) / 256;
  std::vector<int32_t> o(size_t(nb + 1)), ids;
  std::vector<uint32_t> l(static_cast<size_t>(F));
  ids.reserve(size_t(F));
  std::vector<int32_t> u;
  for (int64_t b = 0; b < nb; b++) {
    const int64_t f0 = b * 256, f1 = std::min<int64_t>(f0 + 256, F);
    u.assign(faces + f0 * 3, faces + f1 * 3);
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    o[size_t(b)] = static_cast<int32_t>(ids.size());
    ids.insert(ids.end(), u.begin(), u.end());
    for (int64_t f = f0; f < f1; f++) {
      uint32_t packed = 0;
      for (int k = 0; k < 3; k++) {
        const auto it = std::lower_bound(u.begin(), u.end(), faces[f * 3 + k]);
        packed |= static_cast<uint32_t>(it - u.begin()) << (10 * k);
      }
      l[size_t(f)] = packed;
    }
  }
  o[size_t(nb)] = static_cast<int32_t>(ids.size());
  int err;
  if ((err = upload(c, *off, o.data(), o.size() * 4))) return err;
  if ((err = upload(c, *vid, ids.data(), ids.size() * 4))) return err;
  return upload(c, *loc, l.data(), l.size() * 4);
}

bool check_idx(const int32_t* idx, int64_t n, int32_t limit) {
  for (int64_t i = 0; i < n; i++)
    if (idx[i] < 0 || idx[i] >= limit) return false;
  return true;
}

// Validation of ValidateAndInitBeforeRender (util_private.cc:12-66), same
// order and messages, plus the array presence the device path needs.
int validate(ugu_hip_ctx* c, const UguCamera* cam, const UguOptions* o,
             bool any_output) {
  if (cam == nullptr) return fail(c, UGU_HIP_ERR_NO_CAMERA, "camera has not been set\n");
  if (!c->prepared) return fail(c, UGU_HIP_ERR_MESH_NOT_PREPARED, "mesh has not been initialized\n");
  if (o == nullptr) return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "options are null");
  if (o->backface_culling && !c->has_fn)
    return fail(c, UGU_HIP_ERR_CULLING_NO_FACE_NORMAL,
                "specified back-face culling but face normal is empty.\n");
  if (o->diffuse_color == 1 && c->M == 0)
    return fail(c, UGU_HIP_ERR_TEXTURE_EMPTY,
                "specified texture as diffuse color but texture is empty.\n");
  if (o->diffuse_color == 1) {
    for (int i = 0; i < c->M; i++) {
      if (c->tex_empty[i]) {
        char buf[128];
        std::snprintf(buf, sizeof(buf),
                      "specified texture as diffuse color but %d th texture is empty.\n", i);
        return fail(c, UGU_HIP_ERR_TEXTURE_NTH_EMPTY, buf);
      }
    }
  }
  if (o->diffuse_color == 2 && !c->has_vc)
    return fail(c, UGU_HIP_ERR_VERTEX_COLOR_EMPTY,
                "specified vertex color as diffuse color but vertex color is empty.\n");
  if (o->shading_normal == 0 && !c->has_fn)
    return fail(c, UGU_HIP_ERR_FACE_NORMAL_EMPTY,
                "specified face normal as shading normal but face normal is empty.\n");
  if (o->shading_normal == 1 && !c->has_normals)
    return fail(c, UGU_HIP_ERR_VERTEX_NORMAL_EMPTY,
                "specified vertex normal as shading normal but vertex normal is empty.\n");
  if (!any_output)
    return fail(c, UGU_HIP_ERR_ALL_NULL, "all arguments are nullptr. nothing to do\n");
  // arrays the reference would read out of bounds without
  if (o->diffuse_color == 1 && (!c->has_uv || !c->has_uvidx || !c->has_matid))
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT,
                "texture colouring needs uv, uv_indices and material_ids");
  if (o->shading_normal == 1 && !c->has_nidx)
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT,
                "vertex shading normals need normal_indices");
  if (o->diffuse_color < 0 || o->diffuse_color > 2 || o->interp < 0 || o->interp > 1 ||
      o->shading_normal < 0 || o->shading_normal > 1 || o->diffuse_shading < 0 ||
      o->diffuse_shading > 2)
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "option value out of range");
  return UGU_HIP_OK;
}

int check_camera(ugu_hip_ctx* c, const UguCamera& cam) {
  if (cam.width <= 0 || cam.height <= 0 || cam.width > 32768 || cam.height > 32768)
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "camera size must be in [1, 32768]");
  return 0;
}

bool valid_tile(int tw, int th) {
  return (tw == 8 && th == 8) || (tw == 16 && th == 16) || (tw == 32 && th == 8);
}

DevOptions dev_options(const UguOptions& o) {
  DevOptions d;
  d.diffuse_color = o.diffuse_color;
  d.interp = o.interp;
  d.shading_normal = o.shading_normal;
  d.diffuse_shading = o.diffuse_shading;
  d.culling = o.backface_culling ? 1 : 0;
  // OrenNayarParam(0.3f) (pixel_shader.h:143-148): RendererCpuOption::CopyTo
  // never copies the sigma, so the renderer always sees the default 0.3.
  const float sigma = 0.3f;
  const float sigma_2 = sigma * sigma;
  d.on_A = 1.0f - (0.5f * sigma_2 / (sigma_2 + 0.33f));
  d.on_B = 0.45f * sigma_2 / (sigma_2 + 0.09f);
  d.depth_scale = o.depth_scale;
  return d;
}

DevMesh dev_mesh(const ugu_hip_ctx* c) {
  DevMesh m;
  m.V = c->V;
  m.F = c->F;
  m.verts = c->verts.as<float>();
  m.faces = c->faces.as<int32_t>();
  m.fnormals = c->has_fn ? c->fnormals.as<float>() : nullptr;
  m.soup = c->soup.as<float4>();
  m.shade = c->F > 0 ? c->shade.as<float4>() : nullptr;
  m.blk_off = c->F > 0 ? c->blk_off.as<int32_t>() : nullptr;
  m.blk_vid = c->blk_vid.as<int32_t>();
  m.face_loc = c->face_loc.as<uint32_t>();
  m.normals = c->has_normals ? c->normals.as<float>() : nullptr;
  m.nidx = c->has_nidx ? c->nidx.as<int32_t>() : nullptr;
  m.vcolors = c->has_vc ? c->vcolors.as<float>() : nullptr;
  m.uv = c->has_uv ? c->uv.as<float>() : nullptr;
  m.uvidx = c->has_uvidx ? c->uvidx.as<int32_t>() : nullptr;
  m.matid = c->has_matid ? c->matid.as<int32_t>() : nullptr;
  m.tex = c->M > 0 ? c->texinfo.as<TexInfo>() : nullptr;
  m.tex0 = c->tex0;
  m.n_tex = c->M;
  m.texdata = c->M > 0 ? c->texdata.as<uint8_t>() : nullptr;
  m.ftex = nullptr;
  m.ftex_w = m.ftex_h = 0;
  return m;
}

struct OutReq {
  bool color, depth, normal, mask, fid;
};

// Runs the pipeline for n views. mode 0: reference outputs; mode 1:
// differentiable forward (outputs[i] unused, images/face_ids used).
// Output planes of a render call: the five Render outputs, then the G-buffer
// planes. kPlaneBytes = bytes per pixel.
constexpr int kPlanes = 11;
constexpr int kPlaneBytes[kPlanes] = {3, 4, 12, 1, 4, 12, 12, 12, 12, 12, 4};

void plane_ptrs(const UguOutputs& o, const UguGBuffer* g, void* p[kPlanes]) {
  p[0] = o.color;
  p[1] = o.depth;
  p[2] = o.normal;
  p[3] = o.mask;
  p[4] = o.face_id;
  p[5] = g ? g->bary : nullptr;
  p[6] = g ? g->uv : nullptr;
  p[7] = g ? g->pos_wld : nullptr;
  p[8] = g ? g->pos_cam : nullptr;
  p[9] = g ? g->normal_wld : nullptr;
  p[10] = g ? g->geo_id : nullptr;
}

void set_view_planes(DevView& v, void* const p[kPlanes]) {
  v.color = static_cast<uint8_t*>(p[0]);
  v.depth = static_cast<float*>(p[1]);
  v.normal = static_cast<float*>(p[2]);
  v.mask = static_cast<uint8_t*>(p[3]);
  v.face_id = static_cast<int32_t*>(p[4]);
  v.bary = static_cast<float*>(p[5]);
  v.uvo = static_cast<float*>(p[6]);
  v.pos_wld = static_cast<float*>(p[7]);
  v.pos_cam = static_cast<float*>(p[8]);
  v.nrm_wld = static_cast<float*>(p[9]);
  v.geo_id = static_cast<int32_t*>(p[10]);
}

// Camera floats and tile grid of one view descriptor; output planes null.
void fill_view_camera(DevView& v, const UguCamera& cam, int tw, int th, int32_t tile_base) {
  std::memset(&v, 0, sizeof(v));
  v.W = cam.width;
  v.H = cam.height;
  v.fx = cam.fx;
  v.fy = cam.fy;
  v.cx = cam.cx;
  v.cy = cam.cy;
  std::memcpy(v.w2c_R, cam.w2c_R, sizeof(v.w2c_R));
  std::memcpy(v.w2c_t, cam.w2c_t, sizeof(v.w2c_t));
  std::memcpy(v.c2w_R, cam.c2w_R, sizeof(v.c2w_R));
  std::memcpy(v.c2w_t, cam.c2w_t, sizeof(v.c2w_t));
  v.tiles_x = (v.W + tw - 1) / tw;
  v.tiles_y = (v.H + th - 1) / th;
  v.tile_base = tile_base;
}

int run_pipeline(ugu_hip_ctx* c, const UguCamera* cams, int n, const DevMesh& dm,
                 const DevOptions& dopt, int mode, const UguOutputs* outs,
                 int memory, float* const* images, int32_t* const* fids,
                 bool use_graph, const UguGBuffer* gbufs = nullptr) {
  const int tw = c->tiling.tile_w, th = c->tiling.tile_h;
  // layout
  c->h_tile_base.resize(n + 1);
  int64_t total = 0;
  int64_t stage_bytes = 0;
  std::vector<int64_t> stage_off(size_t(n) * kPlanes, -1);
  for (int i = 0; i < n; i++) {
    int err = check_camera(c, cams[i]);
    if (err) return err;
    const int tx = (cams[i].width + tw - 1) / tw, ty = (cams[i].height + th - 1) / th;
    c->h_tile_base[i] = static_cast<int32_t>(total);
    total += int64_t(tx) * ty;
    if (mode == 0 && memory == UGU_HIP_MEMORY_HOST) {
      const int64_t px = int64_t(cams[i].width) * cams[i].height;
      void* req[kPlanes];
      plane_ptrs(outs[i], gbufs ? &gbufs[i] : nullptr, req);
      for (int k = 0; k < kPlanes; k++) {
        if (req[k]) {
          stage_off[size_t(i) * kPlanes + k] = stage_bytes;
          stage_bytes += (kPlaneBytes[k] * px + 255) / 256 * 256;
        }
      }
    }
  }
  if (total > INT32_MAX / 2) return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "too many tiles");
  c->h_tile_base[n] = static_cast<int32_t>(total);
  const int total_tiles = static_cast<int>(total);

  int err;
  if ((err = ensure(c, c->counts, size_t(total_tiles) * 16))) return err;  // 4 per tile
  if ((err = ensure(c, c->status, kStatusBytes))) return err;
  if ((err = ensure(c, c->views, size_t(n) * sizeof(DevView)))) return err;
  if ((err = ensure(c, c->tile_base, size_t(n + 1) * 4))) return err;
  if (stage_bytes > 0 && (err = ensure(c, c->stage_out, size_t(stage_bytes)))) return err;
  const size_t view_bytes = size_t(n) * sizeof(DevView) + size_t(n + 1) * 4;
  if (c->h_views_bytes < view_bytes) {
    if (c->h_views) (void)hipHostFree(c->h_views);
    c->h_views = nullptr;
    c->h_views_bytes = 0;
    UGU_HIP_CHECK(c, hipHostMalloc(&c->h_views, view_bytes, hipHostMallocDefault));
    c->h_views_bytes = view_bytes;
  }

  // view descriptors
  DevView* hv = static_cast<DevView*>(c->h_views);
  for (int i = 0; i < n; i++) {
    DevView& v = hv[i];
    fill_view_camera(v, cams[i], tw, th, c->h_tile_base[i]);
    if (mode == 0) {
      void* p[kPlanes];
      plane_ptrs(outs[i], gbufs ? &gbufs[i] : nullptr, p);
      if (memory == UGU_HIP_MEMORY_HOST) {
        uint8_t* sb = c->stage_out.as<uint8_t>();
        const int64_t* so = &stage_off[size_t(i) * kPlanes];
        for (int k = 0; k < kPlanes; k++) p[k] = so[k] >= 0 ? sb + so[k] : nullptr;
      }
      set_view_planes(v, p);
    } else {
      v.image = images[i];
      v.face_id = fids ? fids[i] : nullptr;
    }
  }
  std::memcpy(reinterpret_cast<uint8_t*>(c->h_views) + size_t(n) * sizeof(DevView),
              c->h_tile_base.data(), size_t(n + 1) * 4);

  int persistent_blocks = 0;
  if (c->tiling.persistent) {
    const int per_cu = raster_blocks_per_cu(tw, th, mode);
    persistent_blocks = std::max(1, per_cu) * std::max(1, c->cu_count);
    if (persistent_blocks > total_tiles) persistent_blocks = 0;  // plain grid is enough
  }

  c->tm = UguTimings{};
  for (int attempt = 0; attempt < 8; attempt++) {
    if ((err = ensure(c, c->bins, size_t(total_tiles) * size_t(c->tiling.bin_capacity) * 8))) return err;
    DevBins b;
    b.counts = c->counts.as<int32_t>();
    b.bins = c->bins.as<int2>();
    b.cap = c->tiling.bin_capacity;
    b.total_tiles = total_tiles;
    b.status = c->status.as<int32_t>();
    b.tile_base = c->tile_base.as<int32_t>();
    b.n_views = n;
    const DevView* dv = c->views.as<DevView>();

    UGU_HIP_CHECK(c, hipMemcpyAsync(c->views.p, c->h_views, size_t(n) * sizeof(DevView),
                                    hipMemcpyHostToDevice, c->stream));
    UGU_HIP_CHECK(c, hipMemcpyAsync(c->tile_base.p,
                                    reinterpret_cast<uint8_t*>(c->h_views) + size_t(n) * sizeof(DevView),
                                    size_t(n + 1) * 4, hipMemcpyHostToDevice, c->stream));

    auto launch_all = [&](bool timed) -> int {
      if (timed) (void)hipEventRecord(c->ev[0], c->stream);
      if (launch_clear(c->stream, b.counts, total_tiles, b.status, b.tile_base, n))
        return -1;
      if (timed) (void)hipEventRecord(c->ev[1], c->stream);
      if (launch_setup_bin(c->stream, tw, th, dv, n, dm, b)) return -1;
      if (timed) (void)hipEventRecord(c->ev[2], c->stream);
      if (launch_raster(c->stream, tw, th, mode, dv, dm, b, dopt, persistent_blocks)) return -1;
      if (timed) (void)hipEventRecord(c->ev[3], c->stream);
      return 0;
    };

    if (use_graph && !c->profile) {
      GraphKey key;
      std::memset(&key, 0, sizeof(key));
      key.n = n;
      key.total_tiles = total_tiles;
      key.V = dm.V;
      key.F = dm.F;
      key.tw = tw;
      key.th = th;
      key.persistent_blocks = persistent_blocks;
      key.mode = mode;
      key.cap = b.cap;  // the capacity is baked into the captured launches
      key.buf_gen = c->buf_gen;
      key.mesh_gen = c->mesh_gen;
      key.opt = dopt;
      key.verts = dm.verts;
      key.vcolors = dm.vcolors;
      key.ftex = dm.ftex;
      key.ftex_w = dm.ftex_w;
      key.ftex_h = dm.ftex_h;
      hipGraphExec_t exec = nullptr;
      for (auto& g : c->graphs)
        if (g.key == key) exec = g.exec;
      if (!exec) {
        hipGraph_t graph = nullptr;
        UGU_HIP_CHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int le = launch_all(false);
        hipError_t ce = hipStreamEndCapture(c->stream, &graph);
        if (le || ce != hipSuccess) {
          if (graph) (void)hipGraphDestroy(graph);
          return fail(c, UGU_HIP_ERR_HIP, "graph capture failed");
        }
        hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) return hip_fail(c, ie, "hipGraphInstantiate");
        if (c->graphs.size() >= kMaxGraphs) {
          (void)hipGraphExecDestroy(c->graphs.front().exec);
          c->graphs.erase(c->graphs.begin());
        }
        c->graphs.push_back(GraphEntry{key, exec});
      }
      UGU_HIP_CHECK(c, hipGraphLaunch(exec, c->stream));
    } else {
      if (launch_all(c->profile)) return fail(c, UGU_HIP_ERR_HIP, "kernel launch failed");
    }
    UGU_HIP_CHECK(c, hipMemcpyAsync(c->h_status, c->status.p, 16, hipMemcpyDeviceToHost, c->stream));
    UGU_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    const int max_count = c->h_status[0];
    c->tm.max_bin_count = std::max(c->tm.max_bin_count, max_count);
    if (c->profile) {
      float ms[3] = {0, 0, 0};
      for (int k = 0; k < 3; k++) (void)hipEventElapsedTime(&ms[k], c->ev[k], c->ev[k + 1]);
      c->tm.clear_ms += ms[0];
      c->tm.setup_bin_ms += ms[1];
      c->tm.raster_ms += ms[2];
      c->tm.total_ms += ms[0] + ms[1] + ms[2];
    }
    if (max_count <= b.cap) break;
    // a bin overflowed: grow to the measured occupancy and render again
    c->tm.retries++;
    int64_t cap = int64_t(max_count) + max_count / 8 + 64;
    cap = (cap + 63) / 64 * 64;
    if (cap * int64_t(total_tiles) * 8 > (int64_t(64) << 30))
      return fail(c, UGU_HIP_ERR_OUT_OF_MEMORY, "bins would exceed 64 GiB");
    c->tiling.bin_capacity = static_cast<int32_t>(cap);
    if (attempt == 7) return fail(c, UGU_HIP_ERR_OUT_OF_MEMORY, "bin capacity did not converge");
  }

  if (mode == 0 && memory == UGU_HIP_MEMORY_HOST) {
    const uint8_t* sb = c->stage_out.as<uint8_t>();
    for (int i = 0; i < n; i++) {
      const int64_t px = int64_t(cams[i].width) * cams[i].height;
      void* dst[kPlanes];
      plane_ptrs(outs[i], gbufs ? &gbufs[i] : nullptr, dst);
      for (int k = 0; k < kPlanes; k++) {
        const int64_t off = stage_off[size_t(i) * kPlanes + k];
        if (off >= 0)
          UGU_HIP_CHECK(c, hipMemcpyAsync(dst[k], sb + off, size_t(kPlaneBytes[k] * px),
                                          hipMemcpyDeviceToHost, c->stream));
      }
    }
    UGU_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  }
  return UGU_HIP_OK;
}

int render_common(ugu_hip_ctx* c, const UguCamera* cams, int n, const UguOptions* o,
                  const UguOutputs* outs, int memory, bool use_graph,
                  const UguGBuffer* gbufs = nullptr) {
  if (!c) return UGU_HIP_ERR_INVALID_ARGUMENT;
  if (n < 1 || cams == nullptr || outs == nullptr)
    return fail(c, n < 1 ? UGU_HIP_ERR_INVALID_ARGUMENT : UGU_HIP_ERR_NO_CAMERA,
                cams == nullptr ? "camera has not been set\n" : "invalid view count or outputs");
  if (memory != UGU_HIP_MEMORY_HOST && memory != UGU_HIP_MEMORY_DEVICE)
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "unknown memory kind");
  for (int i = 0; i < n; i++) {
    void* p[kPlanes];
    plane_ptrs(outs[i], gbufs ? &gbufs[i] : nullptr, p);
    bool any = false;
    for (int k = 0; k < kPlanes; k++) any = any || p[k] != nullptr;
    int err = validate(c, &cams[i], o, any);
    if (err) return err;
  }
  UGU_HIP_CHECK(c, hipSetDevice(c->device));
  if (int oe = order_after_user(c)) return oe;
  DevMesh dm = dev_mesh(c);
  return run_pipeline(c, cams, n, dm, dev_options(*o), 0, outs, memory, nullptr, nullptr, use_graph,
                      gbufs);
}

// The pipeline of run_pipeline (clear, setup+bin, raster; eager launches, no
// graph) over the scratch array at `base` instead of the context's buffers.
// Views are rendered in consecutive groups whose tiles all get at least
// `want` bin entries; a group in which a tile overflows is rendered again as
// smaller groups with the measured occupancy as the new `want` (the tiles of
// one view do not depend on the other views of a batch, so results equal a
// single batched render). A single view whose largest tile does not fit is
// an error. Nothing is allocated: the descriptors are staged in the
// context's host arrays (sized at create) and the status word is read back
// into its pinned word.
int run_noalloc(ugu_hip_ctx* c, uint8_t* base, const UguCamera* cams, int n, const DevMesh& dm,
                const DevOptions& dopt, const UguOutputs* outs) {
  const int tw = c->tiling.tile_w, th = c->tiling.tile_h;
  int32_t* status = reinterpret_cast<int32_t*>(base);
  DevView* dviews = reinterpret_cast<DevView*>(base + kNaViewsOff);
  int32_t* dtile_base = reinterpret_cast<int32_t*>(base + kNaTileBaseOff);
  const size_t avail = kScratchBytes - kNaCountsOff;
  auto tiles_of = [&](const UguCamera& cam) {
    return int64_t((cam.width + tw - 1) / tw) * int64_t((cam.height + th - 1) / th);
  };
  auto counter_bytes = [](int64_t tiles) { return (size_t(tiles) * 16 + 255) / 256 * 256; };
  int persistent_blocks_per_cu = 0;
  if (c->tiling.persistent) persistent_blocks_per_cu = std::max(1, raster_blocks_per_cu(tw, th, 0));

  int64_t want = std::max<int64_t>({c->na_cap_hint, int64_t(c->tiling.bin_capacity), 64});
  int s = 0;
  while (s < n) {
    // the longest run of views from s whose tiles all get `want` entries
    int e = s;
    int64_t tiles = 0;
    while (e < n) {
      const int64_t t = tiles + tiles_of(cams[e]);
      if (e > s && counter_bytes(t) + size_t(t) * 8 * size_t(want) > avail) break;
      tiles = t;
      e++;
    }
    const size_t cbytes = counter_bytes(tiles);
    const int64_t cap_fit = cbytes < avail ? int64_t((avail - cbytes) / (size_t(tiles) * 8)) : 0;
    const int64_t cap = std::min<int64_t>(cap_fit, INT32_MAX / 2) / 8 * 8;
    if (cap < 8) {
      char msg[200];
      std::snprintf(msg, sizeof(msg),
                    "view %d (%dx%d, %lld tiles of %dx%d) is too large for the %zu MiB "
                    "allocation-free scratch",
                    s, cams[s].width, cams[s].height, static_cast<long long>(tiles), tw, th,
                    kScratchBytes >> 20);
      return fail(c, UGU_HIP_ERR_OUT_OF_MEMORY, msg);
    }
    // descriptors of the group, tiles numbered from 0
    const int g = e - s;
    int32_t tb = 0;
    for (int i = 0; i < g; i++) {
      DevView& v = c->na_views[size_t(i)];
      fill_view_camera(v, cams[s + i], tw, th, tb);
      void* p[kPlanes];
      plane_ptrs(outs[s + i], nullptr, p);
      set_view_planes(v, p);
      c->na_tile_base[size_t(i)] = tb;
      tb += static_cast<int32_t>(tiles_of(cams[s + i]));
    }
    c->na_tile_base[size_t(g)] = tb;
    const int total_tiles = static_cast<int>(tiles);
    DevBins b;
    b.counts = reinterpret_cast<int32_t*>(base + kNaCountsOff);
    b.bins = reinterpret_cast<int2*>(base + kNaCountsOff + cbytes);
    b.cap = static_cast<int32_t>(cap);
    b.total_tiles = total_tiles;
    b.status = status;
    b.tile_base = dtile_base;
    b.n_views = g;
    int persistent_blocks = persistent_blocks_per_cu * std::max(1, c->cu_count);
    if (!c->tiling.persistent || persistent_blocks > total_tiles) persistent_blocks = 0;
    UGU_HIP_CHECK(c, hipMemcpyAsync(dviews, c->na_views.data(), size_t(g) * sizeof(DevView),
                                    hipMemcpyHostToDevice, c->stream));
    UGU_HIP_CHECK(c, hipMemcpyAsync(dtile_base, c->na_tile_base.data(), size_t(g + 1) * 4,
                                    hipMemcpyHostToDevice, c->stream));
    if (launch_clear(c->stream, b.counts, total_tiles, b.status, b.tile_base, g) ||
        launch_setup_bin(c->stream, tw, th, dviews, g, dm, b) ||
        launch_raster(c->stream, tw, th, 0, dviews, dm, b, dopt, persistent_blocks))
      return fail(c, UGU_HIP_ERR_HIP, "kernel launch failed");
    UGU_HIP_CHECK(c, hipMemcpyAsync(c->h_status, status, 16, hipMemcpyDeviceToHost, c->stream));
    UGU_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    const int64_t max_count = c->h_status[0];
    if (max_count <= cap) {
      s = e;
      continue;
    }
    if (g == 1) {
      char msg[200];
      std::snprintf(msg, sizeof(msg),
                    "view %d: a screen tile holds %lld faces; the %zu MiB allocation-free "
                    "scratch holds %lld per tile of this view",
                    s, static_cast<long long>(max_count), kScratchBytes >> 20,
                    static_cast<long long>(cap));
      return fail(c, UGU_HIP_ERR_OUT_OF_MEMORY, msg);
    }
    // `want` grows past the measured occupancy, so the next groups are
    // smaller and this run of views fits
    want = std::max(want, max_count + max_count / 8 + 64);
  }
  c->na_cap_hint = want;
  return UGU_HIP_OK;
}

// Raytracer::Render for n views (outputs[i] belongs to cams[i]).
int raytrace_common(ugu_hip_ctx* c, const UguCamera* cams, int n, const UguOptions* o,
                    const UguOutputs* outs, int memory) {
  if (!c) return UGU_HIP_ERR_INVALID_ARGUMENT;
  if (n < 1 || cams == nullptr || outs == nullptr)
    return fail(c, n < 1 ? UGU_HIP_ERR_INVALID_ARGUMENT : UGU_HIP_ERR_NO_CAMERA,
                cams == nullptr ? "camera has not been set\n" : "invalid view count or outputs");
  if (memory != UGU_HIP_MEMORY_HOST && memory != UGU_HIP_MEMORY_DEVICE)
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "unknown memory kind");
  for (int i = 0; i < n; i++) {
    void* p[kPlanes];
    plane_ptrs(outs[i], nullptr, p);
    bool any = false;
    for (int k = 0; k < kPlanes; k++) any = any || p[k] != nullptr;
    // mesh_initialized_ of the raytracer is set by its own PrepareMesh
    if (c->prepared && !(c->rt_ready && c->rt_mesh_gen == c->mesh_gen) && cams != nullptr)
      return fail(c, UGU_HIP_ERR_MESH_NOT_PREPARED, "mesh has not been initialized\n");
    int err = validate(c, &cams[i], o, any);
    if (err) return err;
    if ((err = check_camera(c, cams[i]))) return err;
  }
  UGU_HIP_CHECK(c, hipSetDevice(c->device));
  if (int oe = order_after_user(c)) return oe;

  // output staging for host memory, view descriptors
  std::vector<int64_t> stage_off(size_t(n) * kPlanes, -1);
  std::vector<int32_t> rt_tile_base(size_t(n) + 1, 0);
  int64_t stage_bytes = 0;
  for (int i = 0; i < n; i++) {
    const int64_t px = int64_t(cams[i].width) * cams[i].height;
    const int64_t tiles = int64_t((cams[i].width + 7) / 8) * ((cams[i].height + 7) / 8);
    const int64_t next = int64_t(rt_tile_base[size_t(i)]) + tiles;
    if (next > INT32_MAX / 2) return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "too many tiles");
    rt_tile_base[size_t(i) + 1] = static_cast<int32_t>(next);
    if (memory == UGU_HIP_MEMORY_HOST) {
      void* req[kPlanes];
      plane_ptrs(outs[i], nullptr, req);
      for (int k = 0; k < kPlanes; k++) {
        if (req[k]) {
          stage_off[size_t(i) * kPlanes + k] = stage_bytes;
          stage_bytes += (kPlaneBytes[k] * px + 255) / 256 * 256;
        }
      }
    }
  }
  int err;
  if ((err = ensure(c, c->views, size_t(n) * sizeof(DevView)))) return err;
  if (stage_bytes > 0 && (err = ensure(c, c->stage_out, size_t(stage_bytes)))) return err;
  const size_t view_bytes = size_t(n) * sizeof(DevView);
  if (c->h_views_bytes < view_bytes) {
    if (c->h_views) (void)hipHostFree(c->h_views);
    c->h_views = nullptr;
    c->h_views_bytes = 0;
    UGU_HIP_CHECK(c, hipHostMalloc(&c->h_views, view_bytes, hipHostMallocDefault));
    c->h_views_bytes = view_bytes;
  }
  DevView* hv = static_cast<DevView*>(c->h_views);
  for (int i = 0; i < n; i++) {
    DevView& v = hv[i];
    std::memset(&v, 0, sizeof(v));
    v.W = cams[i].width;
    v.H = cams[i].height;
    v.fx = cams[i].fx;
    v.fy = cams[i].fy;
    v.cx = cams[i].cx;
    v.cy = cams[i].cy;
    std::memcpy(v.w2c_R, cams[i].w2c_R, sizeof(v.w2c_R));
    std::memcpy(v.w2c_t, cams[i].w2c_t, sizeof(v.w2c_t));
    std::memcpy(v.c2w_R, cams[i].c2w_R, sizeof(v.c2w_R));
    std::memcpy(v.c2w_t, cams[i].c2w_t, sizeof(v.c2w_t));
    v.tile_base = rt_tile_base[size_t(i)];
    void* p[kPlanes];
    plane_ptrs(outs[i], nullptr, p);
    if (memory == UGU_HIP_MEMORY_HOST) {
      uint8_t* sb = c->stage_out.as<uint8_t>();
      const int64_t* so = &stage_off[size_t(i) * kPlanes];
      for (int k = 0; k < kPlanes; k++) p[k] = so[k] >= 0 ? sb + so[k] : nullptr;
    }
    set_view_planes(v, p);
  }
  UGU_HIP_CHECK(c, hipMemcpyAsync(c->views.p, c->h_views, view_bytes, hipMemcpyHostToDevice,
                                  c->stream));
  DevBvh bvh;
  bvh.nodes = c->bvh_nodes.as<BvhNode>();
  bvh.leaves = c->bvh_leaves.as<int2>();
  bvh.tris = c->bvh_tris.as<float4>();
  const DevMesh dm = dev_mesh(c);
  const DevOptions dopt = dev_options(*o);
  c->tm = UguTimings{};
  if (c->profile) (void)hipEventRecord(c->ev[0], c->stream);
  if (launch_raytrace(c->stream, c->views.as<DevView>(), n, rt_tile_base[size_t(n)], bvh, dm, dopt))
    return fail(c, UGU_HIP_ERR_HIP, "kernel launch failed");
  if (c->profile) (void)hipEventRecord(c->ev[1], c->stream);
  if (memory == UGU_HIP_MEMORY_HOST) {
    const uint8_t* sb = c->stage_out.as<uint8_t>();
    for (int i = 0; i < n; i++) {
      const int64_t px = int64_t(cams[i].width) * cams[i].height;
      void* dst[kPlanes];
      plane_ptrs(outs[i], nullptr, dst);
      for (int k = 0; k < kPlanes; k++) {
        const int64_t off = stage_off[size_t(i) * kPlanes + k];
        if (off >= 0)
          UGU_HIP_CHECK(c, hipMemcpyAsync(dst[k], sb + off, size_t(kPlaneBytes[k] * px),
                                          hipMemcpyDeviceToHost, c->stream));
      }
    }
  }
  UGU_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  if (c->profile) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
    c->tm.raster_ms = ms;
    c->tm.total_ms = ms;
  }
  return UGU_HIP_OK;
}

}  // namespace

extern "C" {

int ugu_hip_abi_version(void) { return UGU_HIP_ABI_VERSION; }

ugu_hip_ctx* ugu_hip_create(int device) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return nullptr;
  if (hipSetDevice(device) != hipSuccess) return nullptr;
  ugu_hip_ctx* c = new ugu_hip_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return nullptr;
  }
  c->stream = c->own;
  if (hipEventCreateWithFlags(&c->ev_user, hipEventDisableTiming) != hipSuccess) {
    (void)hipStreamDestroy(c->own);
    delete c;
    return nullptr;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    c->cu_count = prop.multiProcessorCount;
    c->arch = prop.gcnArchName;
    for (char& ch : c->arch)
      if (ch == ' ' || ch == '\n' || ch == '\t') ch = '_';
  }
  if (hipHostMalloc(reinterpret_cast<void**>(&c->h_status), 64, hipHostMallocDefault) != hipSuccess) {
    (void)hipStreamDestroy(c->own);
    delete c;
    return nullptr;
  }
  for (auto& e : c->ev) (void)hipEventCreate(&e);
  init_scratch(device);
  c->na_views.resize(kNoAllocMaxViews);
  c->na_tile_base.resize(kNoAllocMaxViews + 1);
  return c;
}

void ugu_hip_destroy(ugu_hip_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graphs(c);
  DevBuf* bufs[] = {&c->verts, &c->faces, &c->faces_orig, &c->fnormals, &c->normals, &c->nidx,
                    &c->vcolors, &c->uv, &c->uvidx, &c->matid, &c->texinfo,
                    &c->texdata, &c->counts, &c->bins, &c->status,
                    &c->views, &c->tile_base, &c->stage_out, &c->diff_fn,
                    &c->bwd_items, &c->bwd_scratch, &c->soup, &c->diff_soup, &c->shade, &c->blk_off,
                    &c->blk_vid, &c->face_loc, &c->blk_off_o, &c->blk_vid_o, &c->face_loc_o,
                    &c->bvh_nodes, &c->bvh_leaves, &c->bvh_tris, &c->vis_items,
                    &c->vis_images, &c->vis_out, &c->tm_faces, &c->tm_images, &c->tm_out};
  for (DevBuf* b : bufs) b->release();
  if (c->h_views) (void)hipHostFree(c->h_views);
  if (c->h_status) (void)hipHostFree(c->h_status);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->ev_user) (void)hipEventDestroy(c->ev_user);
  if (c->own) (void)hipStreamDestroy(c->own);
  delete c;
}

const char* ugu_hip_last_error(const ugu_hip_ctx* c) {
  return c ? c->err.c_str() : "null context";
}

int ugu_hip_set_stream(ugu_hip_ctx* c, void* s) {
  if (!c) return UGU_HIP_ERR_INVALID_ARGUMENT;
  c->user = static_cast<hipStream_t>(s);
  return UGU_HIP_OK;
}

int ugu_hip_prepare_mesh(ugu_hip_ctx* c, const UguMeshDesc* m) {
  if (!c) return UGU_HIP_ERR_INVALID_ARGUMENT;
  c->prepared = false;
  if (m == nullptr) return fail(c, UGU_HIP_ERR_NO_MESH, "mesh has not been set\n");
  if (m->num_vertices < 0 || m->num_faces < 0 || (m->num_vertices > 0 && !m->vertices) ||
      (m->num_faces > 0 && !m->faces))
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "vertices/faces missing");
  const int64_t F = m->num_faces;
  if (!check_idx(m->faces, F * 3, m->num_vertices))
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "face index out of range");
  const bool has_normals = m->normals && m->num_normals > 0;
  if (m->normal_indices && !check_idx(m->normal_indices, F * 3, has_normals ? m->num_normals : 0))
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "normal index out of range");
  const bool has_vc = m->vertex_colors && m->num_vertex_colors > 0;
  if (has_vc && m->num_vertex_colors < m->num_vertices)
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "fewer vertex colours than vertices");
  const bool has_uv = m->uv && m->num_uv > 0;
  if (m->uv_indices && !check_idx(m->uv_indices, F * 3, has_uv ? m->num_uv : 0))
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "uv index out of range");
  if (m->num_materials < 0 || (m->num_materials > 0 && !m->textures))
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "materials missing");
  if (m->material_ids && !check_idx(m->material_ids, F, m->num_materials))
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "material id out of range");
  UGU_HIP_CHECK(c, hipSetDevice(c->device));
  if (int oe = order_after_user(c)) return oe;
  UGU_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);

  // Relabel vertices, normals and uvs in first-touch order of their index
  // arrays: a screen tile's faces are a compact run of the face array, and
  // after relabelling their vertices are too, so the per-face gathers of the
  // rasterizer hit the same cache lines. Faces keep their order (face ids and
  // the lowest-id tie rule are unchanged) and every value is only moved, so
  // outputs are bitwise identical. The original face array is kept for the
  // differentiable path, whose per-vertex parameters arrive in user order.
  std::vector<int32_t> rv, rn, ru, fr, nr, ur;
  std::vector<float> pv, pn, pc, pu;
  relabel(m->faces, F * 3, m->num_vertices, &fr, &rv);
  const int32_t* faces_r = fr.data();
  pv = gather_rows(m->vertices, m->num_vertices, 3, rv);
  if (has_vc) pc = gather_rows(m->vertex_colors, m->num_vertices, 3, rv);
  const int32_t* nidx_r = nullptr;
  if (m->normal_indices && has_normals) {
    relabel(m->normal_indices, F * 3, m->num_normals, &nr, &rn);
    nidx_r = nr.data();
    pn = gather_rows(m->normals, m->num_normals, 3, rn);
  }
  const int32_t* uvidx_r = nullptr;
  if (m->uv_indices && has_uv) {
    relabel(m->uv_indices, F * 3, m->num_uv, &ur, &ru);
    uvidx_r = ur.data();
    pu = gather_rows(m->uv, m->num_uv, 2, ru);
  }

  int err;
  if ((err = upload_block_lists(c, faces_r, F, &c->blk_off, &c->blk_vid, &c->face_loc))) return err;
  if ((err = upload_block_lists(c, m->faces, F, &c->blk_off_o, &c->blk_vid_o, &c->face_loc_o)))
    return err;
  if ((err = upload(c, c->verts, pv.data(), size_t(m->num_vertices) * 12))) return err;
  if ((err = upload(c, c->faces, faces_r, size_t(F) * 12))) return err;
  if ((err = upload(c, c->faces_orig, m->faces, size_t(F) * 12))) return err;
  if ((err = upload(c, c->fnormals, F > 0 ? m->face_normals : nullptr, size_t(F) * 12))) return err;
  if ((err = upload(c, c->normals, has_normals ? (nidx_r ? pn.data() : m->normals) : nullptr,
                    size_t(m->num_normals) * 12))) return err;
  if ((err = upload(c, c->nidx, nidx_r ? nidx_r : m->normal_indices, size_t(F) * 12))) return err;
  if ((err = upload(c, c->vcolors, has_vc ? pc.data() : nullptr, size_t(m->num_vertices) * 12))) return err;
  if ((err = upload(c, c->uv, has_uv ? (uvidx_r ? pu.data() : m->uv) : nullptr, size_t(m->num_uv) * 8))) return err;
  if ((err = upload(c, c->uvidx, uvidx_r ? uvidx_r : m->uv_indices, size_t(F) * 12))) return err;
  if ((err = upload(c, c->matid, m->material_ids, size_t(F) * 4))) return err;

  c->M = m->num_materials;
  c->tex_empty.assign(size_t(c->M), 0);
  std::vector<TexInfo> ti(size_t(c->M));
  int64_t pool = 0;
  for (int i = 0; i < c->M; i++) {
    const UguTexture& t = m->textures[i];
    const bool empty = t.width < 1 || t.height < 1 || t.data == nullptr;
    c->tex_empty[i] = empty ? 1 : 0;
    ti[i].offset = pool;
    ti[i].w = empty ? 1 : t.width;
    ti[i].h = empty ? 1 : t.height;
    pool += empty ? 256 : (int64_t(t.width) * t.height * 4 + 255) / 256 * 256;
  }
  c->tex0 = c->M > 0 ? ti[0] : TexInfo{};
  if (c->M > 0) {
    if ((err = upload(c, c->texinfo, ti.data(), ti.size() * sizeof(TexInfo)))) return err;
    if (c->texdata.ensure(size_t(pool)) < 0) return fail(c, UGU_HIP_ERR_OUT_OF_MEMORY, "device allocation failed");
    std::vector<uint8_t> rgba;
    for (int i = 0; i < c->M; i++) {
      const UguTexture& t = m->textures[i];
      if (c->tex_empty[i]) continue;
      // RGB8 -> RGBA8 (alpha unused): one aligned 32-bit load per texel
      const size_t n = size_t(t.width) * t.height;
      rgba.assign(n * 4, 0);
      for (size_t k = 0; k < n; k++) {
        rgba[k * 4 + 0] = t.data[k * 3 + 0];
        rgba[k * 4 + 1] = t.data[k * 3 + 1];
        rgba[k * 4 + 2] = t.data[k * 3 + 2];
      }
      UGU_HIP_CHECK(c, hipMemcpy(c->texdata.as<uint8_t>() + ti[i].offset, rgba.data(), n * 4,
                                 hipMemcpyHostToDevice));
    }
  } else {
    c->texinfo.release();
    c->texdata.release();
  }

  c->V = m->num_vertices;
  c->F = m->num_faces;
  c->N = has_normals ? m->num_normals : 0;
  c->U = has_uv ? m->num_uv : 0;
  c->has_fn = m->face_normals != nullptr && F > 0;
  c->has_normals = has_normals;
  c->has_nidx = m->normal_indices != nullptr && F > 0;
  c->has_vc = has_vc;
  c->has_uv = has_uv;
  c->has_uvidx = m->uv_indices != nullptr && F > 0;
  c->has_matid = m->material_ids != nullptr && F > 0;
  if (F > 0) {
    if ((err = ensure(c, c->soup, size_t(F) * 48))) return err;
    if (launch_face_soup(c->stream, c->verts.as<float>(), c->faces.as<int32_t>(),
                         c->has_fn ? c->fnormals.as<float>() : nullptr, F, 0,
                         c->soup.as<float4>(), nullptr))
      return fail(c, UGU_HIP_ERR_HIP, "kernel launch failed");
    if ((err = ensure(c, c->shade, size_t(F) * 64))) return err;
    DevMesh dm = dev_mesh(c);
    dm.shade = nullptr;
    if (launch_shade_rec(c->stream, dm, c->shade.as<float4>()))
      return fail(c, UGU_HIP_ERR_HIP, "kernel launch failed");
    UGU_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  }
  c->mesh_gen++;
  c->prepared = true;
  return UGU_HIP_OK;
}

int ugu_hip_render(ugu_hip_ctx* c, const UguCamera* cam, const UguOptions* o,
                   const UguOutputs* out, int memory) {
  if (c && cam == nullptr) return fail(c, UGU_HIP_ERR_NO_CAMERA, "camera has not been set\n");
  return render_common(c, cam, 1, o, out, memory, true);
}

int ugu_hip_render_batch(ugu_hip_ctx* c, const UguCamera* cams, int n,
                         const UguOptions* o, const UguOutputs* outs, int memory) {
  return render_common(c, cams, n, o, outs, memory, true);
}

int ugu_hip_render_gbuffer(ugu_hip_ctx* c, const UguCamera* cams, int n, const UguOptions* o,
                           const UguOutputs* outs, const UguGBuffer* gbufs, int memory) {
  if (c && !gbufs) return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "gbuffers are required");
  return render_common(c, cams, n, o, outs, memory, true, gbufs);
}

int ugu_hip_render_batch_eager(ugu_hip_ctx* c, const UguCamera* cams, int n,
                               const UguOptions* o, const UguOutputs* outs, int memory) {
  return render_common(c, cams, n, o, outs, memory, false);
}

int ugu_hip_render_batch_noalloc(ugu_hip_ctx* c, const UguCamera* cams, int n,
                                 const UguOptions* o, const UguOutputs* outs) {
  if (!c) return UGU_HIP_ERR_INVALID_ARGUMENT;
  if (n < 1 || cams == nullptr || outs == nullptr)
    return fail(c, n < 1 ? UGU_HIP_ERR_INVALID_ARGUMENT : UGU_HIP_ERR_NO_CAMERA,
                cams == nullptr ? "camera has not been set\n" : "invalid view count or outputs");
  if (n > kNoAllocMaxViews)
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT,
                "the allocation-free render takes at most 1024 views per call");
  for (int i = 0; i < n; i++) {
    void* p[kPlanes];
    plane_ptrs(outs[i], nullptr, p);
    bool any = false;
    for (int k = 0; k < kPlanes; k++) any = any || p[k] != nullptr;
    int err = validate(c, &cams[i], o, any);
    if (err) return err;
    if ((err = check_camera(c, cams[i]))) return err;
  }
  hipError_t se = hipSuccess;
  uint8_t* base = scratch_of(c->device, &se);
  if (base == nullptr)
    return hip_fail(c, se == hipSuccess ? hipErrorInvalidSymbol : se,
                    "scratch of the allocation-free render unavailable on this device");
  UGU_HIP_CHECK(c, hipSetDevice(c->device));
  if (int oe = order_after_user(c)) return oe;
  std::lock_guard<std::mutex> lk(g_scratch_use[c->device]);
  return run_noalloc(c, base, cams, n, dev_mesh(c), dev_options(*o), outs);
}

int ugu_hip_prepare_raytracer(ugu_hip_ctx* c) {
  if (!c) return UGU_HIP_ERR_INVALID_ARGUMENT;
  c->rt_ready = false;
  // Raytracer::Impl::PrepareMesh (raytracer.cc:90-108)
  if (!c->prepared) return fail(c, UGU_HIP_ERR_NO_MESH, "mesh has not been set\n");
  if (c->V == 0 || c->F == 0) return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "mesh is empty\n");
  UGU_HIP_CHECK(c, hipSetDevice(c->device));
  if (int oe = order_after_user(c)) return oe;
  // the face records (positions in face order) are already on the device
  std::vector<float> rec(size_t(c->F) * 12);
  UGU_HIP_CHECK(c, hipMemcpyAsync(rec.data(), c->soup.p, rec.size() * 4, hipMemcpyDeviceToHost,
                                  c->stream));
  UGU_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  BvhBuild b;
  // inner nodes at depth <= kBvhStack - 1: a walk holds at most one deferred
  // child per level, so the stack never exceeds kBvhStack entries
  const int br = build_bvh(rec.data(), c->F, kBvhStack - 1, &b);
  if (br == -1)
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "no face has finite vertex coordinates");
  if (br != 0 || b.max_inner_depth + 1 > kBvhStack)
    return fail(c, UGU_HIP_ERR_HIP, "hierarchy failed its structure checks");
  int err;
  if ((err = upload(c, c->bvh_nodes, b.nodes.data(), b.nodes.size() * 4))) return err;
  if ((err = upload(c, c->bvh_leaves, b.leaves.data(), b.leaves.size() * 4))) return err;
  if ((err = upload(c, c->bvh_tris, b.tris.data(), b.tris.size() * 4))) return err;
  c->rt_mesh_gen = c->mesh_gen;
  c->rt_ready = true;
  return UGU_HIP_OK;
}

int ugu_hip_raytrace(ugu_hip_ctx* c, const UguCamera* cams, int n, const UguOptions* o,
                     const UguOutputs* outs, int memory) {
  return raytrace_common(c, cams, n, o, outs, memory);
}

int ugu_hip_visibility_test(ugu_hip_ctx* c, const UguKeyframe* kf, const UguVisibilityOptions* o,
                            int n, const float* item_pos, const float* item_vertex,
                            const float* item_normal, UguVertexVisibility* out) {
  if (!c) return UGU_HIP_ERR_INVALID_ARGUMENT;
  if (kf == nullptr) return fail(c, UGU_HIP_ERR_NO_CAMERA, "keyframe has not been set\n");
  if (!(c->prepared && c->rt_ready && c->rt_mesh_gen == c->mesh_gen))
    return fail(c, UGU_HIP_ERR_MESH_NOT_PREPARED, "mesh has not been initialized\n");
  if (o == nullptr || n < 0 || (n > 0 && (!item_pos || !item_vertex || !item_normal || !out)))
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "invalid options or items");
  // ValidateAndInitBeforeTest (visibility_tester.cc:331-355)
  if (kf->color == nullptr || kf->width < 1 || kf->height < 1)
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "keyframe color is empty\n");
  if (o->use_mask && kf->mask == nullptr)
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "use_mask is true but keyframe mask is empty\n");
  if (o->use_depth && kf->depth == nullptr)
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "use_depth is true but keyframe depth is empty\n");
  if (o->interp < 0 || o->interp > 1)
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "option value out of range");
  if (int err = check_camera(c, kf->camera)) return err;
  if (n == 0) return UGU_HIP_OK;
  UGU_HIP_CHECK(c, hipSetDevice(c->device));
  if (int oe = order_after_user(c)) return oe;
  const size_t px = size_t(kf->width) * size_t(kf->height);
  const size_t item_bytes = size_t(n) * 12;
  int err;
  if ((err = ensure(c, c->vis_items, item_bytes * 3))) return err;
  if ((err = ensure(c, c->vis_images, px * 8 + 512))) return err;  // colour, depth, mask
  if ((err = ensure(c, c->vis_out, size_t(n) * sizeof(UguVertexVisibility)))) return err;
  uint8_t* items = c->vis_items.as<uint8_t>();
  uint8_t* img = c->vis_images.as<uint8_t>();
  const size_t depth_off = (px * 3 + 255) / 256 * 256, mask_off = depth_off + px * 4;
  UGU_HIP_CHECK(c, hipMemcpyAsync(items, item_pos, item_bytes, hipMemcpyHostToDevice, c->stream));
  UGU_HIP_CHECK(c, hipMemcpyAsync(items + item_bytes, item_vertex, item_bytes,
                                  hipMemcpyHostToDevice, c->stream));
  UGU_HIP_CHECK(c, hipMemcpyAsync(items + 2 * item_bytes, item_normal, item_bytes,
                                  hipMemcpyHostToDevice, c->stream));
  UGU_HIP_CHECK(c, hipMemcpyAsync(img, kf->color, px * 3, hipMemcpyHostToDevice, c->stream));
  if (o->use_depth)
    UGU_HIP_CHECK(c, hipMemcpyAsync(img + depth_off, kf->depth, px * 4, hipMemcpyHostToDevice,
                                    c->stream));
  if (o->use_mask)
    UGU_HIP_CHECK(c, hipMemcpyAsync(img + mask_off, kf->mask, px, hipMemcpyHostToDevice,
                                    c->stream));
  DevView v;
  std::memset(&v, 0, sizeof(v));
  v.W = kf->width;
  v.H = kf->height;
  v.fx = kf->camera.fx;
  v.fy = kf->camera.fy;
  v.cx = kf->camera.cx;
  v.cy = kf->camera.cy;
  std::memcpy(v.w2c_R, kf->camera.w2c_R, sizeof(v.w2c_R));
  std::memcpy(v.w2c_t, kf->camera.w2c_t, sizeof(v.w2c_t));
  std::memcpy(v.c2w_R, kf->camera.c2w_R, sizeof(v.c2w_R));
  std::memcpy(v.c2w_t, kf->camera.c2w_t, sizeof(v.c2w_t));
  DevVisOpt d;
  d.th = o->backface_culling_angle_rad_th;
  d.use_mask = o->use_mask ? 1 : 0;
  d.use_depth = o->use_depth ? 1 : 0;
  d.interp = o->interp;
  d.max_angle = o->max_viewing_angle;
  d.max_dist = o->max_distance_from_camera_to_vertex;
  d.max_ddiff = o->max_depth_difference;
  d.eps = o->eps;
  DevBvh bvh;
  bvh.nodes = c->bvh_nodes.as<BvhNode>();
  bvh.leaves = c->bvh_leaves.as<int2>();
  bvh.tris = c->bvh_tris.as<float4>();
  const float* ip = reinterpret_cast<const float*>(items);
  if (launch_visibility(c->stream, v, img, reinterpret_cast<const float*>(img + depth_off),
                        img + mask_off, d, n, ip, ip + size_t(n) * 3, ip + size_t(n) * 6, bvh,
                        c->vis_out.p))
    return fail(c, UGU_HIP_ERR_HIP, "kernel launch failed");
  UGU_HIP_CHECK(c, hipMemcpyAsync(out, c->vis_out.p, size_t(n) * sizeof(UguVertexVisibility),
                                  hipMemcpyDeviceToHost, c->stream));
  UGU_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return UGU_HIP_OK;
}

int ugu_hip_texture_map(ugu_hip_ctx* c, const UguTexture* images, int n_images,
                        const UguTexFace* faces, int n_faces, int tex_w, int tex_h,
                        int padding_kernel, uint8_t* texture, uint8_t* mask) {
  if (!c) return UGU_HIP_ERR_INVALID_ARGUMENT;
  if (tex_w < 1 || tex_h < 1 || tex_w > 32768 || tex_h > 32768 || padding_kernel < 0 ||
      n_images < 0 || n_faces < 0 || (n_faces > 0 && faces == nullptr) ||
      (n_images > 0 && images == nullptr) || texture == nullptr || mask == nullptr)
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "invalid texture mapping arguments");
  for (int i = 0; i < n_images; i++)
    if (images[i].width < 1 || images[i].height < 1 || images[i].data == nullptr)
      return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "keyframe color is empty\n");
  for (int i = 0; i < n_faces; i++)
    if (faces[i].keyframe < -1 || faces[i].keyframe >= n_images)
      return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "face keyframe index out of range");
  UGU_HIP_CHECK(c, hipSetDevice(c->device));
  if (int oe = order_after_user(c)) return oe;
  struct Img {
    int32_t w, h;
    const uint8_t* data;
  };
  std::vector<size_t> off(size_t(n_images) + 1, 0);
  for (int i = 0; i < n_images; i++)
    off[size_t(i) + 1] = off[size_t(i)] +
                         (size_t(images[i].width) * size_t(images[i].height) * 3 + 255) / 256 * 256;
  const size_t texels = size_t(tex_w) * size_t(tex_h);
  const size_t desc_bytes = (size_t(n_images) * sizeof(Img) + 255) / 256 * 256;
  int err;
  if ((err = ensure(c, c->tm_faces, size_t(n_faces) * sizeof(UguTexFace) + 16))) return err;
  if ((err = ensure(c, c->tm_images, desc_bytes + off[size_t(n_images)] + 16))) return err;
  // owner (int32), unpadded texture + mask, padded texture + mask
  const size_t o_tex0 = texels * 4, o_mask0 = o_tex0 + texels * 3,
               o_tex = (o_mask0 + texels + 255) / 256 * 256, o_mask = o_tex + texels * 3;
  if ((err = ensure(c, c->tm_out, o_mask + texels))) return err;
  uint8_t* ib = c->tm_images.as<uint8_t>();
  std::vector<Img> desc(static_cast<size_t>(n_images));
  for (int i = 0; i < n_images; i++) {
    desc[size_t(i)] = Img{images[i].width, images[i].height, ib + desc_bytes + off[size_t(i)]};
    UGU_HIP_CHECK(c, hipMemcpyAsync(ib + desc_bytes + off[size_t(i)], images[i].data,
                                    size_t(images[i].width) * size_t(images[i].height) * 3,
                                    hipMemcpyHostToDevice, c->stream));
  }
  if (n_images > 0)
    UGU_HIP_CHECK(c, hipMemcpyAsync(ib, desc.data(), desc.size() * sizeof(Img),
                                    hipMemcpyHostToDevice, c->stream));
  if (n_faces > 0)
    UGU_HIP_CHECK(c, hipMemcpyAsync(c->tm_faces.p, faces, size_t(n_faces) * sizeof(UguTexFace),
                                    hipMemcpyHostToDevice, c->stream));
  uint8_t* ob = c->tm_out.as<uint8_t>();
  UGU_HIP_CHECK(c, hipMemsetAsync(ob + o_tex0, 0, texels * 4, c->stream));  // texture0 + mask0
  if (launch_texture_map(c->stream, c->tm_faces.p, n_faces, ib, tex_w, tex_h, padding_kernel / 2,
                         reinterpret_cast<int32_t*>(ob), ob + o_tex0, ob + o_mask0, ob + o_tex,
                         ob + o_mask))
    return fail(c, UGU_HIP_ERR_HIP, "kernel launch failed");
  UGU_HIP_CHECK(c, hipMemcpyAsync(texture, ob + o_tex, texels * 3, hipMemcpyDeviceToHost,
                                  c->stream));
  UGU_HIP_CHECK(c, hipMemcpyAsync(mask, ob + o_mask, texels, hipMemcpyDeviceToHost, c->stream));
  UGU_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return UGU_HIP_OK;
}

int ugu_hip_set_tiling(ugu_hip_ctx* c, const UguTiling* t) {
  if (!c || !t) return UGU_HIP_ERR_INVALID_ARGUMENT;
  if (!valid_tile(t->tile_w, t->tile_h))
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "tile must be 8x8, 16x16 or 32x8");
  if (t->bin_capacity < 1) return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "bin capacity must be >= 1");
  c->tiling = *t;
  drop_graphs(c);
  return UGU_HIP_OK;
}

int ugu_hip_get_tiling(const ugu_hip_ctx* c, UguTiling* t) {
  if (!c || !t) return UGU_HIP_ERR_INVALID_ARGUMENT;
  *t = c->tiling;
  return UGU_HIP_OK;
}

int ugu_hip_enable_profiling(ugu_hip_ctx* c, int enable) {
  if (!c) return UGU_HIP_ERR_INVALID_ARGUMENT;
  c->profile = enable != 0;
  return UGU_HIP_OK;
}

int ugu_hip_get_timings(const ugu_hip_ctx* c, UguTimings* t) {
  if (!c || !t) return UGU_HIP_ERR_INVALID_ARGUMENT;
  *t = c->tm;
  return UGU_HIP_OK;
}

int ugu_hip_autotune(ugu_hip_ctx* c, const UguCamera* cams, int n, const UguOptions* o,
                     int repeats, UguTiling* chosen) {
  if (!c) return UGU_HIP_ERR_INVALID_ARGUMENT;
  if (n < 1 || !cams) return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "no views to tune on");
  if (repeats < 1) repeats = 1;
  UguOptions opt = *o;
  int err = validate(c, &cams[0], &opt, true);
  if (err) return err;
  // outputs: device buffers sized for the views (all five requested)
  int64_t px = 0;
  for (int i = 0; i < n; i++) px += int64_t(cams[i].width) * cams[i].height;
  DevBuf outbuf;
  if (outbuf.ensure(size_t(px) * 24 + size_t(n) * 5 * 256) < 0)
    return fail(c, UGU_HIP_ERR_OUT_OF_MEMORY, "device allocation failed");
  std::vector<UguOutputs> outs(static_cast<size_t>(n));
  uint8_t* p = outbuf.as<uint8_t>();
  for (int i = 0; i < n; i++) {
    const int64_t q = int64_t(cams[i].width) * cams[i].height;
    auto take = [&](int64_t bytes) {
      uint8_t* r = p;
      p += (bytes + 255) / 256 * 256;
      return r;
    };
    outs[i].color = take(3 * q);
    outs[i].depth = reinterpret_cast<float*>(take(4 * q));
    outs[i].normal = reinterpret_cast<float*>(take(12 * q));
    outs[i].mask = take(q);
    outs[i].face_id = reinterpret_cast<int32_t*>(take(4 * q));
  }
  // candidates: every tile shape, with the persistent work queue and with a
  // plain one-workgroup-per-tile grid
  const int shapes[6][3] = {{8, 8, 0}, {16, 16, 0}, {32, 8, 0},
                            {8, 8, 1}, {16, 16, 1}, {32, 8, 1}};
  const UguTiling saved = c->tiling;
  UguTiling best = saved;
  double best_ms = 1e30;
  for (auto& sh : shapes) {
    UguTiling t = saved;
    t.tile_w = sh[0];
    t.tile_h = sh[1];
    t.persistent = sh[2];
    t.bin_capacity = 256;
    c->tiling = t;
    drop_graphs(c);
    // first render sizes the bins from the measured occupancy
    err = render_common(c, cams, n, &opt, outs.data(), UGU_HIP_MEMORY_DEVICE, true);
    if (err) break;
    std::vector<double> times;
    for (int r = 0; r < repeats; r++) {
      auto t0 = std::chrono::steady_clock::now();
      err = render_common(c, cams, n, &opt, outs.data(), UGU_HIP_MEMORY_DEVICE, true);
      auto t1 = std::chrono::steady_clock::now();
      if (err) break;
      times.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
    }
    if (err) break;
    std::sort(times.begin(), times.end());
    const double med = times[times.size() / 2];
    if (med < best_ms) {
      best_ms = med;
      best = c->tiling;
    }
  }
  outbuf.release();
  c->buf_gen++;
  if (err) {
    c->tiling = saved;
    drop_graphs(c);
    return err;
  }
  c->tiling = best;
  drop_graphs(c);
  if (chosen) *chosen = best;
  return UGU_HIP_OK;
}

int ugu_hip_autotune_cached(ugu_hip_ctx* c, const UguCamera* cams, int n, const UguOptions* o,
                            int repeats, const char* cache_path, UguTiling* chosen,
                            int* from_cache) {
  if (!c) return UGU_HIP_ERR_INVALID_ARGUMENT;
  if (from_cache) *from_cache = 0;
  if (n < 1 || !cams || !o) return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "no views to tune on");
  // key: device, mesh size, options that change the work, the views' sizes
  uint64_t sizes = 1469598103934665603ull;  // FNV-1a over (W, H) of every view
  int64_t px = 0;
  for (int i = 0; i < n; i++) {
    for (int32_t v : {cams[i].width, cams[i].height}) {
      sizes = (sizes ^ uint64_t(uint32_t(v))) * 1099511628211ull;
    }
    px += int64_t(cams[i].width) * cams[i].height;
  }
  // kTuneRevision changes whenever the kernels' binning or cost changes, so
  // choices tuned against older kernels are not reused
  constexpr int kTuneRevision = 4;
  char key[256];
  std::snprintf(key, sizeof(key), "r%d|%s|cu%d|F%d|n%d|px%lld|s%016llx|o%d%d%d%d", kTuneRevision,
                c->arch.empty() ? "unknown" : c->arch.c_str(), c->cu_count, c->F, n,
                static_cast<long long>(px), static_cast<unsigned long long>(sizes),
                o->diffuse_color, o->interp, o->shading_normal, o->diffuse_shading);
  if (cache_path && *cache_path) {
    // the last line with this key wins
    FILE* f = std::fopen(cache_path, "r");
    if (f) {
      char line[512];
      UguTiling found{};
      bool hit = false;
      while (std::fgets(line, sizeof(line), f)) {
        char k[256];
        UguTiling t{};
        if (std::sscanf(line, "%255s %d %d %d %d", k, &t.tile_w, &t.tile_h, &t.persistent,
                        &t.bin_capacity) == 5 &&
            std::strcmp(k, key) == 0 && valid_tile(t.tile_w, t.tile_h) && t.bin_capacity > 0 &&
            (t.persistent == 0 || t.persistent == 1)) {
          found = t;
          hit = true;
        }
      }
      std::fclose(f);
      if (hit) {
        c->tiling = found;
        drop_graphs(c);
        if (chosen) *chosen = found;
        if (from_cache) *from_cache = 1;
        return UGU_HIP_OK;
      }
    }
  }
  UguTiling t{};
  int err = ugu_hip_autotune(c, cams, n, o, repeats, &t);
  if (err) return err;
  if (chosen) *chosen = t;
  if (cache_path && *cache_path) {
    // one short line per append: concurrent writers do not interleave inside it
    FILE* f = std::fopen(cache_path, "a");
    if (f) {
      std::fprintf(f, "%s %d %d %d %d\n", key, t.tile_w, t.tile_h, t.persistent, t.bin_capacity);
      std::fclose(f);
    }
  }
  return UGU_HIP_OK;
}

int ugu_hip_depth_to_u16(ugu_hip_ctx* c, const float* depth, int64_t n, uint16_t* out,
                         int memory) {
  if (!c || (n > 0 && (!depth || !out))) return UGU_HIP_ERR_INVALID_ARGUMENT;
  if (memory == UGU_HIP_MEMORY_HOST) {
    for (int64_t i = 0; i < n; i++) {
      const double s = static_cast<double>(1.0f) * static_cast<double>(depth[i]);
      const int32_t iv = (s >= -2147483648.0 && s < 2147483648.0) ? static_cast<int32_t>(s) : INT32_MIN;
      out[i] = static_cast<uint16_t>(static_cast<uint32_t>(iv) & 0xffffu);
    }
    return UGU_HIP_OK;
  }
  UGU_HIP_CHECK(c, hipSetDevice(c->device));
  if (int oe = order_after_user(c)) return oe;
  if (launch_depth_to_u16(c->stream, depth, n, out)) return fail(c, UGU_HIP_ERR_HIP, "kernel launch failed");
  UGU_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  return UGU_HIP_OK;
}

static int diff_mesh(ugu_hip_ctx* c, const UguDiffParams* p, DevMesh* dm) {
  if (!c->prepared) return fail(c, UGU_HIP_ERR_MESH_NOT_PREPARED, "mesh has not been initialized\n");
  if (!p || !p->vertices) return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "vertices are required");
  if (p->texture && (!c->has_uv || !c->has_uvidx || p->tex_width < 1 || p->tex_height < 1))
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "texture needs uv, uv_indices and a size");
  if (p->texture && int64_t(p->tex_width) * p->tex_height > INT32_MAX)
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "texture has more than 2^31-1 texels");
  *dm = dev_mesh(c);
  dm->faces = c->faces_orig.as<int32_t>();  // parameters are in user vertex order
  dm->verts = p->vertices;
  dm->vcolors = p->vertex_colors;
  dm->ftex = p->texture;
  dm->ftex_w = p->texture ? p->tex_width : 0;
  dm->ftex_h = p->texture ? p->tex_height : 0;
  int err = ensure(c, c->diff_fn, size_t(std::max(c->F, 1)) * 12);
  if (err) return err;
  if ((err = ensure(c, c->diff_soup, size_t(std::max(c->F, 1)) * 48))) return err;
  dm->fnormals = c->diff_fn.as<float>();
  dm->soup = c->diff_soup.as<float4>();
  // setup block lists of the original face array (user vertex ids)
  dm->blk_off = c->F > 0 ? c->blk_off_o.as<int32_t>() : nullptr;
  dm->blk_vid = c->blk_vid_o.as<int32_t>();
  dm->face_loc = c->face_loc_o.as<uint32_t>();
  return 0;
}

int ugu_hip_diff_forward(ugu_hip_ctx* c, const UguCamera* cams, int n,
                         const UguDiffParams* p, float* const* images,
                         int32_t* const* face_ids) {
  if (!c) return UGU_HIP_ERR_INVALID_ARGUMENT;
  if (n < 1 || !cams || !images || !face_ids)
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "views, images and face_ids are required");
  for (int i = 0; i < n; i++)
    if (!images[i] || !face_ids[i]) return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "null image");
  UGU_HIP_CHECK(c, hipSetDevice(c->device));
  if (int oe = order_after_user(c)) return oe;
  DevMesh dm;
  int err = diff_mesh(c, p, &dm);
  if (err) return err;
  // face records and normals from this call's positions
  if (launch_face_soup(c->stream, dm.verts, dm.faces, nullptr, dm.F, 1, c->diff_soup.as<float4>(),
                       c->diff_fn.as<float>()))
    return fail(c, UGU_HIP_ERR_HIP, "kernel launch failed");
  UguOptions o{};
  o.backface_culling = p->backface_culling;
  DevOptions dopt = dev_options(o);
  return run_pipeline(c, cams, n, dm, dopt, 1, nullptr, UGU_HIP_MEMORY_DEVICE, images,
                      face_ids, true);
}

int ugu_hip_diff_backward(ugu_hip_ctx* c, const UguCamera* cams, int n,
                          const UguDiffParams* p, const float* const* grad_images,
                          const int32_t* const* face_ids, float* grad_vertices,
                          float* grad_vertex_colors, float* grad_texture) {
  if (!c) return UGU_HIP_ERR_INVALID_ARGUMENT;
  if (n < 1 || !cams || !grad_images || !face_ids)
    return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "views, gradients and face_ids are required");
  UGU_HIP_CHECK(c, hipSetDevice(c->device));
  if (int oe = order_after_user(c)) return oe;
  DevMesh dm;
  int err = diff_mesh(c, p, &dm);
  if (err) return err;
  int max_px = 0;
  int64_t total_px = 0;
  if ((err = ensure(c, c->views, size_t(n) * sizeof(DevView)))) return err;
  const size_t view_bytes = size_t(n) * sizeof(DevView) + size_t(n + 1) * 4;
  if (c->h_views_bytes < view_bytes) {
    if (c->h_views) (void)hipHostFree(c->h_views);
    c->h_views = nullptr;
    c->h_views_bytes = 0;
    UGU_HIP_CHECK(c, hipHostMalloc(&c->h_views, view_bytes, hipHostMallocDefault));
    c->h_views_bytes = view_bytes;
  }
  DevView* hv = static_cast<DevView*>(c->h_views);
  for (int i = 0; i < n; i++) {
    if ((err = check_camera(c, cams[i]))) return err;
    if (!grad_images[i] || !face_ids[i]) return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "null gradient");
    DevView& v = hv[i];
    std::memset(&v, 0, sizeof(v));
    v.W = cams[i].width;
    v.H = cams[i].height;
    v.fx = cams[i].fx;
    v.fy = cams[i].fy;
    v.cx = cams[i].cx;
    v.cy = cams[i].cy;
    std::memcpy(v.w2c_R, cams[i].w2c_R, sizeof(v.w2c_R));
    std::memcpy(v.w2c_t, cams[i].w2c_t, sizeof(v.w2c_t));
    std::memcpy(v.c2w_R, cams[i].c2w_R, sizeof(v.c2w_R));
    v.face_id = const_cast<int32_t*>(face_ids[i]);
    v.grad = grad_images[i];
    max_px = std::max(max_px, v.W * v.H);
    total_px += int64_t(v.W) * v.H;
  }
  if (total_px > INT32_MAX) return fail(c, UGU_HIP_ERR_INVALID_ARGUMENT, "too many pixels");
  if ((err = ensure(c, c->bwd_items, size_t(total_px) * 16))) return err;
  if ((err = ensure(c, c->bwd_scratch, size_t(backward_scratch_ints(dm.F)) * 4))) return err;
  UGU_HIP_CHECK(c, hipMemcpyAsync(c->views.p, c->h_views, size_t(n) * sizeof(DevView),
                                  hipMemcpyHostToDevice, c->stream));
  const DevView* dv = c->views.as<DevView>();
  DevGrads g;
  g.verts = grad_vertices;
  g.colors = p->vertex_colors ? grad_vertex_colors : nullptr;
  g.texture = p->texture ? grad_texture : nullptr;
  if (c->profile) (void)hipEventRecord(c->ev[4], c->stream);
  if (launch_backward(c->stream, dv, n, max_px, total_px, dm, g, c->bwd_items.as<int4>(),
                      c->bwd_scratch.as<int32_t>(), c->profile ? c->ev[5] : nullptr))
    return fail(c, UGU_HIP_ERR_HIP, "kernel launch failed");
  if (c->profile) {
    // item total (the scan's last offset) and the two phases' times
    (void)hipEventRecord(c->ev[3], c->stream);
    UGU_HIP_CHECK(c, hipMemcpyAsync(c->h_status + 4,
                                    c->bwd_scratch.as<int32_t>() + (int64_t(dm.F) + 1) + dm.F, 4,
                                    hipMemcpyDeviceToHost, c->stream));
  }
  UGU_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  if (c->profile) {
    float a = 0.0f, b = 0.0f;
    (void)hipEventElapsedTime(&a, c->ev[4], c->ev[5]);
    (void)hipEventElapsedTime(&b, c->ev[5], c->ev[3]);
    c->tm.bwd_sort_ms = a;
    c->tm.bwd_items_ms = b;
    c->tm.bwd_items = c->h_status[4];
  }
  return UGU_HIP_OK;
}

}  // extern "C"
