"""Differentiable rendering on the gfx950 rasterizer (no reference counterpart).

``DiffRenderer(mesh)`` uploads the mesh topology (faces, uv, uv_indices);
``render(cameras, vertices, colors, texture)`` returns one float32 image per
view, differentiable with respect to the vertex positions, vertex colours and
the float32 RGB texture through ``RasterizeFunction`` (a
``torch.autograd.Function`` over ugu_hip_diff_forward / ugu_hip_diff_backward).

Image model per covered pixel (fid > 0 and not culled), u = w1, v = w2 the
perspective-correct barycentrics of the reference rasterizer:
    c = ((1-u)-v) C0 + u C1 + v C2,  t = bilinear(texture, uv(u, v))
    image = c * t  (either factor is 1 when absent), 0 elsewhere.
Gradients are taken at fixed visibility: the face id of every pixel is the
one found by the forward pass.
"""
import ctypes

import torch

from . import _abi
from .renderer import MeshDesc


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _check_param(t, name, shape_tail):
    if t is None:
        return None
    if not (t.is_cuda and t.dtype == torch.float32):
        raise TypeError(f"{name} must be a float32 CUDA tensor")
    if tuple(t.shape[-len(shape_tail):]) != shape_tail:
        raise ValueError(f"{name} has shape {tuple(t.shape)}")
    return t.contiguous()


class DiffRenderer:
    def __init__(self, mesh, device=0, backface_culling=True):
        self._lib = _abi.load_library()
        self._ctx = self._lib.ugu_hip_create(int(device))
        if not self._ctx:
            raise RuntimeError(f"ugu_hip_create({device}) failed: no usable HIP device")
        self.device = torch.device("cuda", device)
        self.backface_culling = bool(backface_culling)
        self.num_vertices = len(mesh.vertices)
        desc = MeshDesc(mesh)
        self._check(self._lib.ugu_hip_prepare_mesh(self._ctx, ctypes.byref(desc.desc)))

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx:
            self._lib.ugu_hip_destroy(ctx)
            self._ctx = None

    def _check(self, rc):
        if rc != 0:
            raise _abi.HipError(rc, self._lib.ugu_hip_last_error(self._ctx).decode())

    def enable_profiling(self, on=True):
        """HIP-event timings of the following calls on the context's stream."""
        self._check(self._lib.ugu_hip_enable_profiling(self._ctx, 1 if on else 0))

    def timings(self):
        """Timings of the last forward (rasterizer kernels) and backward (item
        sort, gradient kernel, item count) while profiling is on."""
        t = _abi.UguTimings()
        self._check(self._lib.ugu_hip_get_timings(self._ctx, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in _abi.UguTimings._fields_}

    def set_tiling(self, tile_w, tile_h, bin_capacity=1024, persistent=True):
        t = _abi.UguTiling(tile_w, tile_h, bin_capacity, 1 if persistent else 0)
        self._check(self._lib.ugu_hip_set_tiling(self._ctx, ctypes.byref(t)))

    def _sync_stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        self._check(self._lib.ugu_hip_set_stream(self._ctx, ctypes.c_void_p(s)))

    def _params(self, vertices, colors, texture):
        p = _abi.UguDiffParams()
        p.vertices = _ptr(vertices)
        p.vertex_colors = _ptr(colors)
        p.texture = _ptr(texture)
        p.tex_height, p.tex_width = (0, 0) if texture is None else texture.shape[:2]
        p.backface_culling = 1 if self.backface_culling else 0
        return p

    def forward_raw(self, cams, vertices, colors, texture):
        sizes = [c.width * c.height for c in cams]
        total = sum(sizes)
        flat = torch.empty(total * 3, dtype=torch.float32, device=vertices.device)
        fids = torch.empty(total, dtype=torch.int32, device=vertices.device)
        n = len(cams)
        ip = (ctypes.c_void_p * n)()
        fp = (ctypes.c_void_p * n)()
        off = 0
        for i, s in enumerate(sizes):
            ip[i] = flat.data_ptr() + off * 12
            fp[i] = fids.data_ptr() + off * 4
            off += s
        arr = (_abi.UguCamera * n)(*cams)
        p = self._params(vertices, colors, texture)
        self._sync_stream()
        self._check(self._lib.ugu_hip_diff_forward(self._ctx, arr, n, ctypes.byref(p), ip, fp))
        return flat, fids

    def backward_raw(self, cams, vertices, colors, texture, fids, gflat, want):
        n = len(cams)
        sizes = [c.width * c.height for c in cams]
        gp = (ctypes.c_void_p * n)()
        fp = (ctypes.c_void_p * n)()
        off = 0
        for i, s in enumerate(sizes):
            gp[i] = gflat.data_ptr() + off * 12
            fp[i] = fids.data_ptr() + off * 4
            off += s
        gv = torch.zeros_like(vertices) if want[0] else None
        gc = torch.zeros_like(colors) if (want[1] and colors is not None) else None
        gt = torch.zeros_like(texture) if (want[2] and texture is not None) else None
        arr = (_abi.UguCamera * n)(*cams)
        p = self._params(vertices, colors, texture)
        self._sync_stream()
        self._check(self._lib.ugu_hip_diff_backward(self._ctx, arr, n, ctypes.byref(p), gp, fp,
                                                     _ptr(gv), _ptr(gc), _ptr(gt)))
        return gv, gc, gt

    def render_flat(self, cameras, vertices, colors=None, texture=None):
        """All views in one flat float32 tensor (view i occupies H_i*W_i*3
        consecutive values, row-major HWC) and the flat face ids; a loss over
        many views is then a few kernels instead of a few per view."""
        cams = [c if isinstance(c, _abi.UguCamera) else c.to_abi() for c in cameras]
        vertices = _check_param(vertices, "vertices", (3,))
        if vertices.shape[0] != self.num_vertices:
            raise ValueError("vertices must match the prepared mesh")
        colors = _check_param(colors, "colors", (3,))
        texture = _check_param(texture, "texture", (3,))
        return RasterizeFunction.apply(vertices, colors, texture, self, cams)

    def render(self, cameras, vertices, colors=None, texture=None, return_face_ids=False):
        """Images (list of (H, W, 3) float32 tensors) for each camera."""
        cams = [c if isinstance(c, _abi.UguCamera) else c.to_abi() for c in cameras]
        vertices = _check_param(vertices, "vertices", (3,))
        if vertices.shape[0] != self.num_vertices:
            raise ValueError("vertices must match the prepared mesh")
        colors = _check_param(colors, "colors", (3,))
        texture = _check_param(texture, "texture", (3,))
        flat, fids = RasterizeFunction.apply(vertices, colors, texture, self, cams)
        imgs, ids, off = [], [], 0
        for c in cams:
            s = c.width * c.height
            imgs.append(flat[off * 3:(off + s) * 3].view(c.height, c.width, 3))
            ids.append(fids[off:off + s].view(c.height, c.width))
            off += s
        return (imgs, ids) if return_face_ids else imgs


class RasterizeFunction(torch.autograd.Function):
    """Forward: rasterize + shade (HIP). Backward: HIP scatter of dL/dimage to
    vertices, vertex colours and texels at fixed visibility."""

    @staticmethod
    def forward(ctx, vertices, colors, texture, renderer, cams):
        flat, fids = renderer.forward_raw(cams, vertices, colors, texture)
        ctx.save_for_backward(vertices, colors, texture, fids)
        ctx.renderer = renderer
        ctx.cams = cams
        ctx.mark_non_differentiable(fids)
        return flat, fids

    @staticmethod
    def backward(ctx, gflat, _gfid):
        vertices, colors, texture, fids = ctx.saved_tensors
        want = ctx.needs_input_grad[:3]
        gv, gc, gt = ctx.renderer.backward_raw(ctx.cams, vertices, colors, texture, fids,
                                               gflat.contiguous(), want)
        return gv, gc, gt, None, None
