"""Seeded synthetic scenes of the benchmark workloads.

* UV sphere with MakeUvSphere semantics (src/util/geom_util.cc:643-706).
* Icosphere by recursive 4-way subdivision (subdivision level L: 20*4^L faces;
  level 8 = 1,310,720 faces, 655,362 vertices), children of face f stored at
  4f..4f+3, so the face order is spatially coherent.
* Geodesic icosphere of frequency n (20*n^2 faces; n = 71 gives 100,820).
* Radius perturbation r = 1 + 0.05 sin(6 theta) cos(8 phi), theta the polar
  angle from +y, phi = atan2(z, x); spherical UVs u = phi/(2 pi) + 0.5,
  v = 1 - theta/pi.
* Seeded RGB8 / float textures (uniform noise + 64-px checker).
* Cameras: look-at poses (c2wImpl, src/util/camera_util.cc:19-34) and the
  PinholeCamera fov constructor (include/ugu/camera.h:273-284).
"""
import math

import numpy as np

from .types import Mesh, PinholeCamera, look_at_c2w

RES_CHOICES = [(320, 240), (640, 480), (1280, 720), (1920, 1080)]


def make_uv_sphere(n_stacks, n_slices):
    """MakeUvSphere(n_stacks, n_slices) (geom_util.cc:643-706): vertices in
    double then cast to float; faces in the reference's order. (The
    reference's uv array is misaligned with uv_indices; it is not produced.)"""
    verts = [(0.0, 1.0, 0.0)]
    for i in range(n_stacks - 1):
        phi = math.pi * float(i + 1) / float(n_stacks)
        for j in range(n_slices):
            theta = 2.0 * math.pi * float(j) / float(n_slices)
            verts.append((math.sin(phi) * math.cos(theta), math.cos(phi),
                          math.sin(phi) * math.sin(theta)))
    v1 = len(verts)
    verts.append((0.0, -1.0, 0.0))
    faces = []
    for i in range(n_slices):
        i0 = i + 1
        i1 = (i + 1) % n_slices + 1
        faces.append((0, i1, i0))
        i0 = i + n_slices * (n_stacks - 2) + 1
        i1 = (i + 1) % n_slices + n_slices * (n_stacks - 2) + 1
        faces.append((v1, i0, i1))
    for j in range(n_stacks - 2):
        j0 = j * n_slices + 1
        j1 = (j + 1) * n_slices + 1
        for i in range(n_slices):
            a = j0 + i
            b = j0 + (i + 1) % n_slices
            c = j1 + (i + 1) % n_slices
            d = j1 + i
            faces.append((a, b, c))
            faces.append((a, c, d))
    return Mesh(np.array(verts, np.float64).astype(np.float32), np.array(faces, np.int32))


def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
                  (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
                  (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
                  (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
                  (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], np.int64)
    return v, f


def icosphere_unit(level):
    """Unit icosphere (float64 vertices, int64 faces) by `level` recursive
    midpoint subdivisions with re-projection onto the sphere."""
    v, f = _icosahedron()
    for _ in range(level):
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        e = np.stack([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)], 1)
        e = np.sort(e.reshape(-1, 2), axis=1)
        key = e[:, 0] * (len(v) + 1) + e[:, 1]
        uk, first, inv = np.unique(key, return_index=True, return_inverse=True)
        ue = e[first]
        mid = v[ue[:, 0]] + v[ue[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        mid_id = (len(v) + inv).reshape(-1, 3)
        ab, bc, ca = mid_id[:, 0], mid_id[:, 1], mid_id[:, 2]
        v = np.concatenate([v, mid], 0)
        nf = np.stack([np.stack([a, ab, ca], 1), np.stack([ab, b, bc], 1),
                       np.stack([ca, bc, c], 1), np.stack([ab, bc, ca], 1)], 1)
        f = nf.reshape(-1, 3)
    return v, f


def geodesic_unit(freq):
    """Class-I geodesic sphere of frequency `freq`: each icosahedron face split
    into freq^2 triangles, shared edge/corner vertices deduplicated."""
    v0, f0 = _icosahedron()
    n = int(freq)
    verts = [tuple(x) for x in v0]
    edge_pts = {}

    def edge_vertex(a, b, k):
        # k-th interior point from a towards b (1..n-1), shared by both faces
        if a < b:
            key, kk = (a, b), k
        else:
            key, kk = (b, a), n - k
        lst = edge_pts.get(key)
        if lst is None:
            lo, hi = key
            lst = []
            for s in range(1, n):
                p = v0[lo] + (v0[hi] - v0[lo]) * (s / n)
                lst.append(len(verts))
                verts.append(tuple(p))
            edge_pts[key] = lst
        return lst[kk - 1]

    faces = []
    for (A, B, C) in f0:
        idx = {}
        for i in range(n + 1):
            for j in range(n + 1 - i):
                # point = A + (B-A) i/n + (C-A) j/n
                if i == 0 and j == 0:
                    vid = A
                elif i == n:
                    vid = B
                elif j == n:
                    vid = C
                elif j == 0:
                    vid = edge_vertex(A, B, i)
                elif i == 0:
                    vid = edge_vertex(A, C, j)
                elif i + j == n:
                    vid = edge_vertex(B, C, j)
                else:
                    p = v0[A] + (v0[B] - v0[A]) * (i / n) + (v0[C] - v0[A]) * (j / n)
                    vid = len(verts)
                    verts.append(tuple(p))
                idx[(i, j)] = vid
        for i in range(n):
            for j in range(n - i):
                faces.append((idx[(i, j)], idx[(i + 1, j)], idx[(i, j + 1)]))
                if i + j + 1 < n:
                    faces.append((idx[(i + 1, j)], idx[(i + 1, j + 1)], idx[(i, j + 1)]))
    v = np.array(verts, np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v, np.array(faces, np.int64)


def perturbed_sphere(unit_v, faces, amp=0.05, texture=None, with_normals=True):
    """Radius 1 + amp*sin(6 theta)*cos(8 phi); spherical UVs; face and vertex
    normals (CalcNormal semantics)."""
    y = np.clip(unit_v[:, 1], -1.0, 1.0)
    theta = np.arccos(y)
    phi = np.arctan2(unit_v[:, 2], unit_v[:, 0])
    r = 1.0 + amp * np.sin(6.0 * theta) * np.cos(8.0 * phi)
    m = Mesh((unit_v * r[:, None]).astype(np.float32), faces.astype(np.int32))
    uv = np.stack([phi / (2.0 * math.pi) + 0.5, 1.0 - theta / math.pi], 1)
    m.set_uv(uv.astype(np.float32))
    m.set_uv_indices(m.vertex_indices)
    if with_normals:
        m.CalcNormal()
    else:
        m.CalcFaceNormal()
    if texture is not None:
        m.set_single_material(texture)
    return m


def checker_noise_texture(size, seed, dtype=np.uint8, block=64):
    """Seeded uniform noise plus a `block`-pixel checker: 0.5*noise + 0.5*checker."""
    rng = np.random.default_rng(seed)
    noise = rng.random((size, size, 3), dtype=np.float32)
    yy, xx = np.meshgrid(np.arange(size), np.arange(size), indexing="ij")
    checker = (((yy // block) + (xx // block)) % 2).astype(np.float32)[..., None]
    t = 0.5 * noise + 0.5 * checker
    if dtype == np.uint8:
        return np.clip(t * 255.0, 0, 255).astype(np.uint8)
    return t.astype(np.float32)


def front_camera(width, height, f, distance):
    """Camera on -z looking at the origin with identity rotation (OpenCV axes),
    principal point at the image centre (W/2-0.5, H/2-0.5)."""
    c2w = np.eye(4)
    c2w[2, 3] = -float(distance)
    return PinholeCamera(width, height, c2w,
                         (width * 0.5 - 0.5, height * 0.5 - 0.5), (f, f))


def ring_cameras(n, radius, seed_elev, seed_res, res_choices=None, fov_y_deg=45.0,
                 elev_deg=30.0):
    """n cameras on a ring around the origin, azimuth 2 pi k / n, elevation
    uniform in [-elev_deg, elev_deg] (seed_elev), resolution drawn from
    res_choices (seed_res), PinholeCamera(w, h, c2w, fov_y)."""
    res_choices = res_choices or RES_CHOICES
    re = np.random.default_rng(seed_elev)
    rr = np.random.default_rng(seed_res)
    elev = np.radians(re.uniform(-elev_deg, elev_deg, n))
    res = [res_choices[i] for i in rr.integers(0, len(res_choices), n)]
    cams = []
    for k in range(n):
        az = 2.0 * math.pi * k / n
        pos = radius * np.array([math.cos(elev[k]) * math.sin(az), math.sin(elev[k]),
                                 -math.cos(elev[k]) * math.cos(az)])
        c2w = look_at_c2w(pos, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
        w, h = res[k]
        cams.append(PinholeCamera(w, h, c2w, fov_y_deg=fov_y_deg))
    return cams


# ---------------------------------------------------------------------------
# benchmark workloads


def scene_uv_sphere_128(seed=0):
    """UV sphere 128x128 (32,512 faces), analytic vertex normals, vertex
    colours U[0,1) x 255 from `seed`; camera 640x480, f=500, 3 units away."""
    m = make_uv_sphere(128, 128)
    n = m.vertices / np.linalg.norm(m.vertices, axis=1, keepdims=True)
    m.set_normals(n.astype(np.float32))
    m.set_normal_indices(m.vertex_indices)
    m.CalcFaceNormal()
    rng = np.random.default_rng(seed)
    m.set_vertex_colors((rng.random((len(m.vertices), 3), dtype=np.float32) * 255.0))
    cam = front_camera(640, 480, 500.0, 3.0)
    return m, cam


def scene_icosphere8(level=8, tex_size=2048, seed=0):
    """Icosphere subdivided `level` times, perturbed radius, spherical UVs,
    seeded RGB8 texture; camera 1920x1080, f=1400 at distance 2.5."""
    v, f = icosphere_unit(level)
    tex = checker_noise_texture(tex_size, seed)
    m = perturbed_sphere(v, f, texture=tex)
    cam = front_camera(1920, 1080, 1400.0, 2.5)
    return m, cam


def multiview_cameras(n=64):
    """n cameras on a ring of radius 3, elevations U[-30,30] deg (seed 1),
    resolutions from {320x240, 640x480, 1280x720, 1920x1080} (seed 2), fov_y 45."""
    return ring_cameras(n, 3.0, 1, 2)


def inverse_texturing_cameras():
    """The 32 views of the inverse-texturing workload: a ring of radius 3
    (elevations seed 41), views 0-15 at 512x512 and 16-31 at 1024x768, fov_y
    45 deg, principal points at the image centres."""
    cams = ring_cameras(32, 3.0, 41, 42, res_choices=[(512, 512)])
    for k in range(16, 32):
        cams[k].set_size(1024, 768)
        cams[k].principal_point[:] = (511.5, 383.5)
        cams[k].set_fov_y(45.0)
    return cams
