"""Host-side mirror of the reference types the rasterizer reads.

* ``RendererCpuOption`` and the enums: include/ugu/renderer/cpu/renderer.h:17-38,
  include/ugu/renderer/base.h:11-32, include/ugu/common.h:50-53.
* ``Mesh``: the arrays of ugu::Mesh (include/ugu/mesh.h:81-199) plus
  ``CalcFaceNormal`` / ``CalcNormal`` (src/mesh.cc:297-349).
* ``PinholeCamera``: include/ugu/camera.h:77-154,257-331 (pose, intrinsics,
  the fov constructor) and the float reductions the renderer takes from it
  (src/renderer/cpu/rasterizer.cc:118-119, camera.h:317-331).
* ``Image``: a numpy-backed stand-in for ugu::Image<T> (include/ugu/image.h:934-985):
  interleaved HWC, row-major, contiguous.
"""
import enum
import math

import numpy as np


class DiffuseColor(enum.IntEnum):
    kNone = 0
    kTexture = 1
    kVertex = 2


class ShadingNormal(enum.IntEnum):
    kFace = 0
    kVertex = 1


class DiffuseShading(enum.IntEnum):
    kNone = 0
    kLambertian = 1
    kOrenNayar = 2


class ColorInterpolation(enum.IntEnum):
    kNn = 0
    kBilinear = 1


class RendererCpuOption:
    """renderer.h:17-38 with the same defaults."""

    def __init__(self, diffuse_color=DiffuseColor.kNone, interp=ColorInterpolation.kBilinear,
                 shading_normal=ShadingNormal.kVertex, diffuse_shading=DiffuseShading.kNone,
                 depth_scale=1.0, backface_culling=True, oren_nayar_sigma=0.3):
        self.diffuse_color = DiffuseColor(diffuse_color)
        self.interp = ColorInterpolation(interp)
        self.shading_normal = ShadingNormal(shading_normal)
        self.diffuse_shading = DiffuseShading(diffuse_shading)
        self.depth_scale = float(depth_scale)
        self.backface_culling = bool(backface_culling)
        self.oren_nayar_sigma = float(oren_nayar_sigma)

    def CopyTo(self, dst):
        """renderer.h:30-37: copies six fields; oren_nayar_sigma is NOT copied."""
        dst.diffuse_color = self.diffuse_color
        dst.depth_scale = self.depth_scale
        dst.interp = self.interp
        dst.shading_normal = self.shading_normal
        dst.diffuse_shading = self.diffuse_shading
        dst.backface_culling = self.backface_culling


class Image:
    """Minimal ugu::Image<T>: ``data`` is an (H, W[, C]) numpy array."""

    def __init__(self, dtype, channels):
        self.dtype = np.dtype(dtype)
        self.channels = channels
        self.data = np.zeros((0, 0, channels) if channels > 1 else (0, 0), self.dtype)

    @property
    def rows(self):
        return self.data.shape[0]

    @property
    def cols(self):
        return self.data.shape[1]

    def empty(self):
        return self.rows < 1 or self.cols < 1

    def init(self, width, height, value):
        """ugu::Init (image.h:964-985): reallocate when the size differs, then fill."""
        if self.cols != width or self.rows != height:
            shape = (height, width, self.channels) if self.channels > 1 else (height, width)
            self.data = np.zeros(shape, self.dtype)
        self.data[...] = value


def Image3b():
    return Image(np.uint8, 3)


def Image1b():
    return Image(np.uint8, 1)


def Image1f():
    return Image(np.float32, 1)


def Image3f():
    return Image(np.float32, 3)


def Image1i():
    return Image(np.int32, 1)


def Image1w():
    return Image(np.uint16, 1)


def _sum3(a, b, c):
    # Eigen's unrolled fixed-size reduction order: x0 + (x1 + x2)
    return a + (b + c)


def _normalize_rows_f32(v):
    """Eigen normalized() on float32 rows: z = squaredNorm; v / sqrt(z) if z > 0."""
    v = v.astype(np.float32)
    z = _sum3(v[:, 0] * v[:, 0], v[:, 1] * v[:, 1], v[:, 2] * v[:, 2])
    s = np.sqrt(z, dtype=np.float32)
    out = v.copy()
    nz = z > 0
    out[nz] = v[nz] / s[nz, None]
    return out


def _cross_f32(a, b):
    """Eigen cross() of 3-vectors, per component."""
    return np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1],
                     a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                     a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], axis=1).astype(np.float32)


class Mesh:
    """Arrays of ugu::Mesh. All float arrays float32, index arrays int32."""

    def __init__(self, vertices=None, faces=None):
        self.vertices = _f32(vertices, 3)
        self.vertex_indices = _i32(faces, 3)
        self.vertex_colors = np.zeros((0, 3), np.float32)
        self.normals = np.zeros((0, 3), np.float32)
        self.face_normals = np.zeros((0, 3), np.float32)
        self.normal_indices = np.zeros((0, 3), np.int32)
        self.uv = np.zeros((0, 2), np.float32)
        self.uv_indices = np.zeros((0, 3), np.int32)
        self.material_ids = np.zeros((0,), np.int32)
        self.materials = []  # list of (H, W, 3) uint8 diffuse textures

    def set_vertices(self, v):
        self.vertices = _f32(v, 3)

    def set_vertex_indices(self, f):
        self.vertex_indices = _i32(f, 3)

    def set_vertex_colors(self, c):
        self.vertex_colors = _f32(c, 3)

    def set_normals(self, n):
        self.normals = _f32(n, 3)

    def set_face_normals(self, n):
        self.face_normals = _f32(n, 3)

    def set_normal_indices(self, i):
        self.normal_indices = _i32(i, 3)

    def set_uv(self, uv):
        self.uv = _f32(uv, 2)

    def set_uv_indices(self, i):
        self.uv_indices = _i32(i, 3)

    def set_material_ids(self, ids):
        self.material_ids = np.ascontiguousarray(ids, np.int32).reshape(-1)

    def set_single_material(self, texture):
        self.materials = [np.ascontiguousarray(texture, np.uint8)]
        self.material_ids = np.zeros(len(self.vertex_indices), np.int32)

    def CalcFaceNormal(self):
        """src/mesh.cc:339-349: normalize(normalize(p1-p0) x normalize(p2-p0))."""
        f = self.vertex_indices
        p0, p1, p2 = self.vertices[f[:, 0]], self.vertices[f[:, 1]], self.vertices[f[:, 2]]
        v1 = _normalize_rows_f32(p1 - p0)
        v2 = _normalize_rows_f32(p2 - p0)
        self.face_normals = _normalize_rows_f32(_cross_f32(v1, v2))

    def CalcNormal(self):
        """src/mesh.cc:297-337: per-vertex average of face normals, normalized,
        accumulated in face order; normal_indices = vertex_indices."""
        if len(self.vertex_indices) == 0:
            return
        self.CalcFaceNormal()
        self.normal_indices = self.vertex_indices.copy()
        acc = np.zeros((len(self.vertices), 3), np.float32)
        cnt = np.zeros(len(self.vertices), np.int64)
        # np.add.at is unbuffered: additions happen in face-major, corner-minor
        # order, the loop order of the reference
        order = self.vertex_indices.reshape(-1)
        fn_rep = np.repeat(self.face_normals, 3, axis=0)
        np.add.at(acc, order, fn_rep)
        np.add.at(cnt, order, 1)
        ok = cnt > 0
        acc[ok] = acc[ok] / cnt[ok, None].astype(np.float32)
        acc[ok] = _normalize_rows_f32(acc[ok])
        acc[~ok] = 0.0
        self.normals = acc


def _f32(a, cols):
    if a is None:
        return np.zeros((0, cols), np.float32)
    return np.ascontiguousarray(a, np.float32).reshape(-1, cols)


def _i32(a, cols):
    if a is None:
        return np.zeros((0, cols), np.int32)
    return np.ascontiguousarray(a, np.int32).reshape(-1, cols)


# ---------------------------------------------------------------------------
# camera


def _cofactor(m, i, j):
    i1, i2, j1, j2 = (i + 1) % 3, (i + 2) % 3, (j + 1) % 3, (j + 2) % 3
    return m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1]


def inverse3x3(m):
    """3x3 inverse by cofactors in double, in the order Eigen's fixed-size
    compute_inverse<3> evaluates it (det = c00*m00 + (c10*m10 + c20*m20))."""
    m = [[float(m[r][c]) for c in range(3)] for r in range(3)]
    c0 = [_cofactor(m, 0, 0), _cofactor(m, 1, 0), _cofactor(m, 2, 0)]
    det = _sum3(c0[0] * m[0][0], c0[1] * m[1][0], c0[2] * m[2][0])
    inv = 1.0 / det
    r = [[0.0] * 3 for _ in range(3)]
    for i in range(3):
        r[0][i] = c0[i] * inv
        r[1][i] = _cofactor(m, i, 1) * inv
        r[2][i] = _cofactor(m, i, 2) * inv
    return r


def affine_inverse(c2w):
    """Eigen Transform<double,3,Affine>::inverse(): linear^-1, -linear^-1 * t."""
    c2w = np.asarray(c2w, np.float64)
    li = inverse3x3(c2w[:3, :3])
    t = [float(x) for x in c2w[:3, 3]]
    out = np.eye(4)
    for r in range(3):
        for c in range(3):
            out[r, c] = li[r][c]
        out[r, 3] = -_sum3(li[r][0] * t[0], li[r][1] * t[1], li[r][2] * t[2])
    return out


def polar_rotation(m):
    """Transform::rotation(): the rotation factor of the polar decomposition
    (U * diag(1,1,det) * V^T from an SVD, in double)."""
    u, _, vt = np.linalg.svd(np.asarray(m, np.float64))
    x = -1.0 if np.linalg.det(u @ vt) < 0 else 1.0
    u = u.copy()
    u[:, 2] *= x
    return u @ vt


def radians_f32(deg):
    """ugu::radians(float) (common.h): float arithmetic."""
    return np.float32(np.float32(deg) * np.float32(math.pi) / np.float32(180.0))


class PinholeCamera:
    """camera.h:77-154,257-331. Pose c2w is a 4x4 double matrix (Affine3d)."""

    def __init__(self, width=-1, height=-1, c2w=None, principal_point=None,
                 focal_length=None, fov_y_deg=None):
        self.width = int(width)
        self.height = int(height)
        if fov_y_deg is not None:
            # PinholeCamera(w, h, [c2w,] fov_y): pp = size*0.5f - 0.5f, set_fov_y
            self.principal_point = np.array([np.float32(self.width) * np.float32(0.5) - np.float32(0.5),
                                             np.float32(self.height) * np.float32(0.5) - np.float32(0.5)],
                                            np.float32)
            self.set_fov_y(fov_y_deg)
        else:
            self.principal_point = np.asarray(
                principal_point if principal_point is not None else (-1, -1), np.float32)
            self.focal_length = np.asarray(
                focal_length if focal_length is not None else (-1, -1), np.float32)
        self.set_c2w(np.eye(4) if c2w is None else c2w)

    def set_fov_y(self, fov_y_deg):
        # camera.h set_fov_y: height*0.5f / float(tan(radians(fov)*0.5)) (double tan)
        t = np.float32(math.tan(float(radians_f32(fov_y_deg)) * 0.5))
        f = np.float32(np.float32(self.height) * np.float32(0.5) / t)
        self.focal_length = np.array([f, f], np.float32)

    def set_size(self, width, height):
        self.width, self.height = int(width), int(height)

    def set_c2w(self, c2w):
        self.c2w = np.array(c2w, np.float64).reshape(4, 4)
        self.w2c = affine_inverse(self.c2w)

    def to_abi(self):
        """The float reductions of rasterizer.cc:118-119 and camera.h:322
        (c2w_t: the camera position, the raytracer's ray origin)."""
        from ._abi import UguCamera
        cam = UguCamera()
        cam.width, cam.height = self.width, self.height
        cam.fx, cam.fy = float(self.focal_length[0]), float(self.focal_length[1])
        cam.cx, cam.cy = float(self.principal_point[0]), float(self.principal_point[1])
        w2c_R = polar_rotation(self.w2c[:3, :3]).astype(np.float32)
        w2c_t = self.w2c[:3, 3].astype(np.float32)
        c2w_R = self.c2w[:3, :3].astype(np.float32)
        for i in range(9):
            cam.w2c_R[i] = float(w2c_R.reshape(-1)[i])
            cam.c2w_R[i] = float(c2w_R.reshape(-1)[i])
        c2w_t = self.c2w[:3, 3].astype(np.float32)
        for i in range(3):
            cam.w2c_t[i] = float(w2c_t[i])
            cam.c2w_t[i] = float(c2w_t[i])
        return cam

    def arrays(self):
        """(w2c_R, w2c_t, c2w_R, c2w_t) float32, as the renderer consumes them."""
        return (polar_rotation(self.w2c[:3, :3]).astype(np.float32),
                self.w2c[:3, 3].astype(np.float32),
                self.c2w[:3, :3].astype(np.float32),
                self.c2w[:3, 3].astype(np.float32))


def look_at_c2w(position, target, up):
    """c2wImpl (src/util/camera_util.cc:19-50), OpenCV axes, in double:
    z = normalize(target - pos), x = normalize(z x up), y = normalize(z x x)."""
    p = np.asarray(position, np.float64)
    t = np.asarray(target, np.float64)
    u = np.asarray(up, np.float64)

    def nrm(v):
        z = _sum3(v[0] * v[0], v[1] * v[1], v[2] * v[2])
        return v / math.sqrt(z) if z > 0 else v

    def cross(a, b):
        return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                         a[0] * b[1] - a[1] * b[0]])

    z = nrm(t - p)
    x = nrm(cross(z, u))
    y = nrm(cross(z, x))
    T = np.eye(4)
    T[:3, 0], T[:3, 1], T[:3, 2], T[:3, 3] = x, y, z, p
    return T
