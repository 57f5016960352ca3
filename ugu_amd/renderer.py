"""Python mirror of ugu::Rasterizer / ugu::RendererCpu over the C ABI.

Same method names and error behaviour as the reference
(include/ugu/renderer/cpu/renderer.h:42-76, src/renderer/cpu/rasterizer.cc):
``set_*`` never fail, ``PrepareMesh`` and ``Render*`` return bool and print
the reference's LOGE message on failure. ``Render`` fills caller-provided
``Image`` objects (``None`` = not requested), reallocating them to the camera
size as ``ValidateAndInitBeforeRender`` does (util_private.cc:12-89).

Extensions with no reference counterpart: ``render_batch`` (many views in one
pipeline), ``render_device`` (outputs in device memory, torch tensors),
``autotune``.
"""
import ctypes
import sys

import numpy as np

from . import _abi
from .types import RendererCpuOption


def LOGE(msg):
    sys.stderr.write(msg if msg.endswith("\n") else msg + "\n")


def LOGW(msg):
    sys.stderr.write(msg if msg.endswith("\n") else msg + "\n")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def options_to_abi(o):
    opt = _abi.UguOptions()
    opt.diffuse_color = int(o.diffuse_color)
    opt.interp = int(o.interp)
    opt.shading_normal = int(o.shading_normal)
    opt.diffuse_shading = int(o.diffuse_shading)
    opt.depth_scale = float(o.depth_scale)
    opt.backface_culling = 1 if o.backface_culling else 0
    opt.oren_nayar_sigma = float(o.oren_nayar_sigma)
    return opt


class MeshDesc:
    """Keeps the numpy arrays alive while a UguMeshDesc points into them."""

    def __init__(self, mesh):
        m = mesh
        self.keep = []

        def arr(a, dtype):
            if a is None or len(a) == 0:
                return None
            a = np.ascontiguousarray(a, dtype)
            self.keep.append(a)
            return a

        d = _abi.UguMeshDesc()
        v = arr(m.vertices, np.float32)
        f = arr(m.vertex_indices, np.int32)
        d.num_vertices = 0 if v is None else len(v)
        d.vertices = _ptr(v)
        d.num_faces = 0 if f is None else len(f)
        d.faces = _ptr(f)
        fn = arr(m.face_normals, np.float32)
        d.face_normals = _ptr(fn) if fn is not None and len(fn) >= d.num_faces else None
        n = arr(m.normals, np.float32)
        d.num_normals = 0 if n is None else len(n)
        d.normals = _ptr(n)
        ni = arr(m.normal_indices, np.int32)
        d.normal_indices = _ptr(ni) if ni is not None and len(ni) >= d.num_faces else None
        vc = arr(m.vertex_colors, np.float32)
        d.num_vertex_colors = 0 if vc is None else len(vc)
        d.vertex_colors = _ptr(vc)
        uv = arr(m.uv, np.float32)
        d.num_uv = 0 if uv is None else len(uv)
        d.uv = _ptr(uv)
        ui = arr(m.uv_indices, np.int32)
        d.uv_indices = _ptr(ui) if ui is not None and len(ui) >= d.num_faces else None
        mi = arr(m.material_ids, np.int32)
        d.material_ids = _ptr(mi) if mi is not None and len(mi) >= d.num_faces else None
        nm = len(m.materials)
        self.tex = (_abi.UguTexture * max(nm, 1))()
        for i, t in enumerate(m.materials):
            t = None if t is None or np.size(t) == 0 else np.ascontiguousarray(t, np.uint8)
            if t is None:
                self.tex[i].width = 0
                self.tex[i].height = 0
                self.tex[i].data = None
            else:
                self.keep.append(t)
                self.tex[i].width = t.shape[1]
                self.tex[i].height = t.shape[0]
                self.tex[i].data = _ptr(t)
        d.num_materials = nm
        d.textures = self.tex if nm > 0 else None
        self.desc = d


class Rasterizer:
    """ugu::Rasterizer on the gfx950 pipeline (one HIP context per object)."""

    def __init__(self, option=None, device=0):
        self._lib = _abi.load_library()
        self._ctx = self._lib.ugu_hip_create(int(device))
        if not self._ctx:
            raise RuntimeError(f"ugu_hip_create({device}) failed: no usable HIP device")
        self.option_ = RendererCpuOption()
        if option is not None:
            self.set_option(option)
        self.mesh_ = None
        self.camera_ = None
        self.mesh_initialized_ = False

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx:
            self._lib.ugu_hip_destroy(ctx)
            self._ctx = None

    # --- reference interface ------------------------------------------------
    def set_option(self, option):
        option.CopyTo(self.option_)

    def set_mesh(self, mesh):
        """rasterizer.cc:77-88."""
        self.mesh_initialized_ = False
        self.mesh_ = mesh
        if len(mesh.face_normals) == 0:
            LOGW("face normal is empty. culling and shading may not work\n")
        if len(mesh.normals) == 0:
            LOGW("vertex normal is empty. shading may not work\n")

    def PrepareMesh(self):
        """rasterizer.cc:90-99, plus the device upload."""
        if self.mesh_ is None:
            LOGE("mesh has not been set\n")
            return False
        desc = MeshDesc(self.mesh_)
        rc = self._lib.ugu_hip_prepare_mesh(self._ctx, ctypes.byref(desc.desc))
        if rc != 0:
            LOGE(self.last_error())
            return False
        self.mesh_initialized_ = True
        return True

    def set_camera(self, camera):
        self.camera_ = camera

    def Render(self, color=None, depth=None, normal=None, mask=None, face_id=None):
        if self.camera_ is None:
            LOGE("camera has not been set\n")
            return False
        if not self.mesh_initialized_:
            LOGE("mesh has not been initialized\n")
            return False
        w, h = self.camera_.width, self.camera_.height
        arrays = {}
        # validation happens in the library first; images are only reset on success
        for name, img, shape, dt in (("color", color, (h, w, 3), np.uint8),
                                     ("depth", depth, (h, w), np.float32),
                                     ("normal", normal, (h, w, 3), np.float32),
                                     ("mask", mask, (h, w), np.uint8),
                                     ("face_id", face_id, (h, w), np.int32)):
            arrays[name] = None if img is None else np.empty(shape, dt)
        out = _abi.UguOutputs(*[_ptr(arrays[k]) for k in
                                ("color", "depth", "normal", "mask", "face_id")])
        cam = self.camera_.to_abi()
        opt = options_to_abi(self.option_)
        rc = self._render_views(ctypes.byref(cam), 1, ctypes.byref(opt), ctypes.byref(out),
                                _abi.UGU_HIP_MEMORY_HOST, graph=False)
        if rc != 0:
            LOGE(self.last_error())
            return False
        for name, img in (("color", color), ("depth", depth), ("normal", normal),
                          ("mask", mask), ("face_id", face_id)):
            if img is not None:
                img.data = arrays[name]
        return True

    def RenderColor(self, color):
        return self.Render(color, None, None, None, None)

    def RenderDepth(self, depth):
        return self.Render(None, depth, None, None, None)

    def RenderNormal(self, normal):
        return self.Render(None, None, normal, None, None)

    def RenderMask(self, mask):
        return self.Render(None, None, None, mask, None)

    def RenderFaceId(self, face_id):
        return self.Render(None, None, None, None, face_id)

    def RenderW(self, color, depth, normal, mask, face_id):
        """rasterizer.cc:329-344: float depth converted with ConvertTo."""
        from .types import Image1f
        if depth is None:
            LOGE("depth is nullptr")
            return False
        f_depth = Image1f()
        ok = self.Render(color, f_depth, normal, mask, face_id)
        if ok:
            out = np.empty(f_depth.data.shape, np.uint16)
            self._lib.ugu_hip_depth_to_u16(self._ctx, _ptr(f_depth.data), f_depth.data.size,
                                           _ptr(out), _abi.UGU_HIP_MEMORY_HOST)
            depth.data = out
        return ok

    def RenderDepthW(self, depth):
        return self.RenderW(None, depth, None, None, None)

    # --- extensions ---------------------------------------------------------
    def _render_views(self, cams, n, opt, outs, memory, graph=True):
        if n == 1 and not graph:
            return self._lib.ugu_hip_render(self._ctx, cams, opt, outs, memory)
        fn = self._lib.ugu_hip_render_batch if graph else self._lib.ugu_hip_render_batch_eager
        return fn(self._ctx, cams, n, opt, outs, memory)

    def last_error(self):
        return self._lib.ugu_hip_last_error(self._ctx).decode()

    def _check(self, rc):
        if rc != 0:
            raise _abi.HipError(rc, self.last_error())

    def set_tiling(self, tile_w, tile_h, bin_capacity=1024, persistent=True):
        t = _abi.UguTiling(tile_w, tile_h, bin_capacity, 1 if persistent else 0)
        self._check(self._lib.ugu_hip_set_tiling(self._ctx, ctypes.byref(t)))

    def get_tiling(self):
        t = _abi.UguTiling()
        self._check(self._lib.ugu_hip_get_tiling(self._ctx, ctypes.byref(t)))
        return dict(tile_w=t.tile_w, tile_h=t.tile_h, bin_capacity=t.bin_capacity,
                    persistent=bool(t.persistent))

    def set_stream(self, stream_ptr):
        self._check(self._lib.ugu_hip_set_stream(self._ctx, stream_ptr))

    def enable_profiling(self, on=True):
        self._check(self._lib.ugu_hip_enable_profiling(self._ctx, 1 if on else 0))

    def timings(self):
        t = _abi.UguTimings()
        self._check(self._lib.ugu_hip_get_timings(self._ctx, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in _abi.UguTimings._fields_}

    def autotune(self, cameras, repeats=3, cache_path=None):
        """Tile shape and bin capacity for these views (ugu_hip_autotune); with
        cache_path, a choice kept in that file for the same device, mesh size,
        options and view sizes is applied without tuning
        (ugu_hip_autotune_cached). Returns the tiling and whether it came from
        the file."""
        cams = (_abi.UguCamera * len(cameras))(*[c.to_abi() for c in cameras])
        opt = options_to_abi(self.option_)
        t = _abi.UguTiling()
        hit = ctypes.c_int(0)
        if cache_path is None:
            self._check(self._lib.ugu_hip_autotune(self._ctx, cams, len(cameras),
                                                   ctypes.byref(opt), int(repeats),
                                                   ctypes.byref(t)))
        else:
            self._check(self._lib.ugu_hip_autotune_cached(
                self._ctx, cams, len(cameras), ctypes.byref(opt), int(repeats),
                str(cache_path).encode(), ctypes.byref(t), ctypes.byref(hit)))
        return dict(tile_w=t.tile_w, tile_h=t.tile_h, bin_capacity=t.bin_capacity,
                    persistent=bool(t.persistent), from_cache=bool(hit.value))

    def render_batch(self, cameras, outputs=("color", "depth", "normal", "mask", "face_id"),
                     graph=True):
        """Render many views into host numpy arrays; returns a list of dicts."""
        res, outs = [], (_abi.UguOutputs * len(cameras))()
        for i, c in enumerate(cameras):
            w, h = c.width, c.height
            d = {}
            if "color" in outputs:
                d["color"] = np.empty((h, w, 3), np.uint8)
            if "depth" in outputs:
                d["depth"] = np.empty((h, w), np.float32)
            if "normal" in outputs:
                d["normal"] = np.empty((h, w, 3), np.float32)
            if "mask" in outputs:
                d["mask"] = np.empty((h, w), np.uint8)
            if "face_id" in outputs:
                d["face_id"] = np.empty((h, w), np.int32)
            outs[i] = _abi.UguOutputs(*[_ptr(d.get(k)) for k in
                                        ("color", "depth", "normal", "mask", "face_id")])
            res.append(d)
        cams = (_abi.UguCamera * len(cameras))(
            *[c if isinstance(c, _abi.UguCamera) else c.to_abi() for c in cameras])
        opt = options_to_abi(self.option_)
        self._check(self._render_views(cams, len(cameras), ctypes.byref(opt), outs,
                                       _abi.UGU_HIP_MEMORY_HOST, graph=graph))
        return res

    GBUFFER_PLANES = ("bary", "uv", "pos_wld", "pos_cam", "normal_wld", "geo_id")

    def render_gbuffer(self, cameras, outputs=("color", "depth", "normal", "mask", "face_id"),
                       gbuffer=GBUFFER_PLANES):
        """render_batch plus G-buffer planes (ugu_hip_render_gbuffer): the
        fields of ugu::GBuffer the rasterizer's resolve fills (bary, uv,
        pos_wld, pos_cam, normal_wld as H x W x 3 float32, geo_id H x W int32).
        Returns a list of dicts holding both kinds of planes."""
        res = []
        outs = (_abi.UguOutputs * len(cameras))()
        gbs = (_abi.UguGBuffer * len(cameras))()
        shapes = {"color": (3, np.uint8), "depth": (0, np.float32), "normal": (3, np.float32),
                  "mask": (0, np.uint8), "face_id": (0, np.int32), "bary": (3, np.float32),
                  "uv": (3, np.float32), "pos_wld": (3, np.float32), "pos_cam": (3, np.float32),
                  "normal_wld": (3, np.float32), "geo_id": (0, np.int32)}
        for i, c in enumerate(cameras):
            d = {}
            for k in tuple(outputs) + tuple(gbuffer):
                ch, dt = shapes[k]
                d[k] = np.empty((c.height, c.width, ch) if ch else (c.height, c.width), dt)
            outs[i] = _abi.UguOutputs(*[_ptr(d.get(k)) for k in
                                        ("color", "depth", "normal", "mask", "face_id")])
            gbs[i] = _abi.UguGBuffer(*[_ptr(d.get(k)) for k in self.GBUFFER_PLANES])
            res.append(d)
        cams = (_abi.UguCamera * len(cameras))(
            *[c if isinstance(c, _abi.UguCamera) else c.to_abi() for c in cameras])
        opt = options_to_abi(self.option_)
        self._check(self._lib.ugu_hip_render_gbuffer(self._ctx, cams, len(cameras),
                                                     ctypes.byref(opt), outs, gbs,
                                                     _abi.UGU_HIP_MEMORY_HOST))
        return res

    def render_device(self, cameras, buffers, graph=True, batched=True):
        """Render into device memory. `buffers` is a list (one per camera) of
        dicts of torch CUDA tensors (any of color/depth/normal/mask/face_id).
        batched=False issues one ugu_hip_render per view (the per-view loop)."""
        cams = [c if isinstance(c, _abi.UguCamera) else c.to_abi() for c in cameras]
        opt = options_to_abi(self.option_)

        def outs_of(b):
            return _abi.UguOutputs(*[(b[k].data_ptr() if k in b and b[k] is not None else None)
                                     for k in ("color", "depth", "normal", "mask", "face_id")])
        if batched:
            arr = (_abi.UguCamera * len(cams))(*cams)
            outs = (_abi.UguOutputs * len(cams))(*[outs_of(b) for b in buffers])
            self._check(self._render_views(arr, len(cams), ctypes.byref(opt), outs,
                                           _abi.UGU_HIP_MEMORY_DEVICE, graph=graph))
        else:
            for c, b in zip(cams, buffers):
                o = outs_of(b)
                self._check(self._render_views(ctypes.byref(c), 1, ctypes.byref(opt),
                                               ctypes.byref(o), _abi.UGU_HIP_MEMORY_DEVICE,
                                               graph=False))


    def bind_device(self, cameras, buffers):
        """Marshal `cameras` and the device `buffers` of render_device once,
        for a loop that renders into the same tensors again and again (a
        serving loop): render_bound(binding) then only crosses into the
        library. set_camera(i, camera) on the binding changes a pose; the
        tensors must stay alive and unmoved while the binding is used."""
        return DeviceBinding(self, cameras, buffers)

    def render_bound(self, binding, graph=True):
        """render_device for a DeviceBinding (same results, no per-call
        marshalling of arguments)."""
        self._check(self._render_views(binding.cams, binding.n, ctypes.byref(binding.opt),
                                       binding.outs, _abi.UGU_HIP_MEMORY_DEVICE, graph=graph))


class DeviceBinding:
    """Camera, option and output-pointer arrays of one render_device call,
    built once (Rasterizer.bind_device). The option struct is taken when the
    binding is made."""

    def __init__(self, renderer, cameras, buffers):
        if len(cameras) != len(buffers) or not cameras:
            raise ValueError("one buffer dict per camera, at least one camera")
        self.n = len(cameras)
        self.cams = (_abi.UguCamera * self.n)(
            *[c if isinstance(c, _abi.UguCamera) else c.to_abi() for c in cameras])
        self.sizes = [(c.width, c.height) for c in self.cams]
        for (w, h), b in zip(self.sizes, buffers):
            for k, t in b.items():
                if t is not None and tuple(t.shape[:2]) != (h, w):
                    raise ValueError(f"buffer {k} has shape {tuple(t.shape)} for a {w}x{h} view")
        self.opt = options_to_abi(renderer.option_)
        self.outs = (_abi.UguOutputs * self.n)(*[
            _abi.UguOutputs(*[(b[k].data_ptr() if k in b and b[k] is not None else None)
                              for k in ("color", "depth", "normal", "mask", "face_id")])
            for b in buffers])
        self._buffers = buffers  # keeps the tensors alive

    def set_camera(self, i, camera):
        c = camera if isinstance(camera, _abi.UguCamera) else camera.to_abi()
        if (c.width, c.height) != self.sizes[i]:
            raise ValueError(f"view {i} is bound to {self.sizes[i][0]}x{self.sizes[i][1]} "
                             f"buffers, camera is {c.width}x{c.height}")
        self.cams[i] = c


class Raytracer(Rasterizer):
    """ugu::Raytracer (src/renderer/cpu/raytracer.cc) on the gfx950 raytracer
    kernel: PrepareMesh also builds the bounding volume hierarchy
    (raytracer.cc:90-108), Render traces one ray per pixel from the camera
    position (raytracer.cc:117-225). Every pixel gets the closest hit over all
    faces (smallest float t, equal t -> lowest face id); the reference's
    std::sort leaves equal t unordered. render_batch / render_device trace many
    views in one launch; the rasterizer-only extensions (tiling, autotune,
    G-buffer) are not available."""

    def PrepareMesh(self):
        """raytracer.cc:90-108: set_mesh is the rasterizer's (raytracer.cc:78-88)."""
        if self.mesh_ is None:
            LOGE("mesh has not been set\n")
            return False
        if len(self.mesh_.vertices) == 0 or len(self.mesh_.vertex_indices) == 0:
            LOGE("mesh is empty\n")
            return False
        if not super().PrepareMesh():
            return False
        rc = self._lib.ugu_hip_prepare_raytracer(self._ctx)
        if rc != 0:
            LOGE(self.last_error())
            self.mesh_initialized_ = False
            return False
        return True

    def _render_views(self, cams, n, opt, outs, memory, graph=True):
        return self._lib.ugu_hip_raytrace(self._ctx, cams, n, opt, outs, memory)

    def _unsupported(self, *a, **k):
        raise NotImplementedError("rasterizer-only extension")

    set_tiling = get_tiling = autotune = render_gbuffer = _unsupported
