"""Pins of the CPU oracle against facts that do not come from the oracle.

Each pin returns (ok, message). test_oracle_pins.py requires every pin to
pass on the faithful oracle; test_oracle_mutations.py rebuilds the oracle
with planted mistakes and requires at least one pin to fail on each.

Sources of truth used here:
* an independent float64 brute-force rasterizer (per pixel, per face,
  exact-sign coverage, the reference's sequential depth rule);
* closed forms of a sphere seen by a pinhole camera (silhouette radius
  f R / sqrt(d^2 - R^2), near-point depth d - R, analytic normals);
* ray-plane intersections for perspective-correct attribute interpolation;
* hand-computed values of the truncating conversions.
"""
import math

import numpy as np

import oracle as O
import tiny_scenes as TS
from ugu_amd import (ColorInterpolation, DiffuseColor, DiffuseShading, Mesh, PinholeCamera,
                     RendererCpuOption, ShadingNormal, look_at_c2w, scenes)

FLT_MIN = float(np.finfo(np.float32).tiny)


def render(mesh, cam, opt):
    rc, out, _ = O.render(mesh, O.camera_from_abi(cam.to_abi()), opt, c2w_t=cam.c2w[:3, 3])
    assert rc == 0, rc
    return out


# ---------------------------------------------------------------------------
# float64 brute force


def brute_force(mesh, cam, culling, all_edges_ambiguous=False):
    """Independent float64 rasterizer: returns (face_id, depth, ambiguous).
    all_edges_ambiguous: pixels within 1e-6 of any edge count as ambiguous,
    covered or not (for checks of a different coverage rule)."""
    a = cam.to_abi()
    R = np.array(a.w2c_R[:], np.float64).reshape(3, 3)
    t = np.array(a.w2c_t[:], np.float64)
    C = np.array(a.c2w_R[:], np.float64).reshape(3, 3)
    fx, fy, cx, cy = float(a.fx), float(a.fy), float(a.cx), float(a.cy)
    W, H = cam.width, cam.height
    q = mesh.vertices.astype(np.float64) @ R.T + t
    X = fx * q[:, 0] / q[:, 2] + cx
    Y = fy * q[:, 1] / q[:, 2] + cy
    Z = q[:, 2]
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    d = np.stack([(xs - cx) / fx, (ys - cy) / fy, np.ones_like(xs)], -1)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    ray = d @ C.T
    depth = np.zeros((H, W))
    fid = -np.ones((H, W), np.int64)
    bfimg = np.zeros((H, W), bool)
    amb = np.zeros((H, W), bool)
    fn = mesh.face_normals.astype(np.float64)
    for i, (i0, i1, i2) in enumerate(mesh.vertex_indices):
        if Z[i0] < 0 or Z[i1] < 0 or Z[i2] < 0:
            continue
        x0, y0, x1, y1, x2, y2 = X[i0], Y[i0], X[i1], Y[i1], X[i2], Y[i2]
        A = (x2 - x0) * (y1 - y0) - (y2 - y0) * (x1 - x0)
        if abs(A) < FLT_MIN:
            continue
        e0 = (xs - x1) * (y2 - y1) - (ys - y1) * (x2 - x1)
        e1 = (xs - x2) * (y0 - y2) - (ys - y2) * (x0 - x2)
        e2 = (xs - x0) * (y1 - y0) - (ys - y0) * (x1 - x0)
        bf = (ray @ fn[i]) > 0
        cov = np.where(bf, (e0 <= 0) & (e1 <= 0) & (e2 <= 0), (e0 >= 0) & (e1 >= 0) & (e2 >= 0))
        # the reference only visits its integer bbox
        bx0, bx1 = max(0, math.ceil(min(x0, x1, x2))), min(W - 1, math.floor(max(x0, x1, x2)))
        by0, by1 = max(0, math.ceil(min(y0, y1, y2))), min(H - 1, math.floor(max(y0, y1, y2)))
        inbox = (xs >= bx0) & (xs <= bx1) & (ys >= by0) & (ys <= by1)
        cov &= inbox
        near_edge = inbox & (np.minimum(np.minimum(abs(e0), abs(e1)), abs(e2)) < 1e-6 * abs(A))
        amb |= near_edge if all_edges_ambiguous else (near_edge & ~cov)
        with np.errstate(divide="ignore", invalid="ignore"):
            z = 1.0 / (e0 / A / Z[i0] + e1 / A / Z[i1] + e2 / A / Z[i2])
        for yy, xx in zip(*np.nonzero(cov)):
            zz, dd = z[yy, xx], depth[yy, xx]
            if dd > 0 and abs(zz - dd) <= 1e-5 * dd:
                amb[yy, xx] = True
            if dd < FLT_MIN or zz < dd:
                depth[yy, xx] = zz
                fid[yy, xx] = i
                bfimg[yy, xx] = bf[yy, xx]
    if culling:
        fid[bfimg] = -1
        depth[bfimg] = 0.0
    return fid, depth, amb


def pin_brute_force(seeds=range(8)):
    for seed in seeds:
        w, h = [(8, 8), (16, 12), (21, 17), (32, 32)][seed % 4]
        tris = TS.random_triangles(8 + 2 * seed, w, h, 100 + seed)
        mesh = TS.mesh_from_pixels(tris, seed=seed)
        cam = TS.exact_camera(w, h)
        for cull in (False, True):
            opt = RendererCpuOption(diffuse_color=DiffuseColor.kVertex, backface_culling=cull)
            out = render(mesh, cam, opt)
            fid, depth, amb = brute_force(mesh, cam, cull)
            ok = (out["face_id"] == fid) | amb
            if not ok.all():
                return False, f"brute force seed {seed} cull {cull}: face_id differs at " \
                              f"{np.argwhere(~ok)[:3].tolist()}"
            sel = (fid >= 0) & ~amb
            if not np.allclose(out["depth"][sel], depth[sel], rtol=1e-5):
                return False, f"brute force seed {seed}: depth differs"
    return True, ""


def pin_ties_lowest_id():
    """Duplicated faces produce bitwise-equal depths: the lower id must win
    (strict less in face order, rasterizer.cc:229)."""
    tris = TS.random_triangles(10, 16, 16, 9, zs=(1.0,))
    tris = np.concatenate([tris, tris], 0)
    mesh = TS.mesh_from_pixels(tris, seed=2)
    out = render(mesh, TS.exact_camera(16, 16),
                 RendererCpuOption(diffuse_color=DiffuseColor.kVertex, backface_culling=False))
    f = out["face_id"]
    if (f >= 10).any():
        return False, "a duplicated face with the higher id won a depth tie"
    if not (f >= 0).any():
        return False, "nothing covered"
    return True, ""


def pin_face_zero_not_shaded():
    """Face 0 sets depth and face_id but is never shaded (`if (fid > 0)`,
    rasterizer.cc:254)."""
    tris = np.array([[(1, 1, 1), (14, 2, 1), (2, 14, 1)],
                     [(8, 8, 2), (15, 9, 2), (9, 15, 2)]], np.float64)[:, ::-1, :]
    mesh = TS.mesh_from_pixels(tris, seed=4)
    out = render(mesh, TS.exact_camera(16, 16),
                 RendererCpuOption(diffuse_color=DiffuseColor.kVertex))
    z = out["face_id"] == 0
    o = out["face_id"] == 1
    if not z.any() or not o.any():
        return False, "scene does not cover both faces"
    if out["mask"][z].any() or out["color"][z].any() or out["normal"][z].any():
        return False, "face 0 was shaded"
    if not (out["depth"][z] > 0).all():
        return False, "face 0 lost its depth"
    if not (out["mask"][o] == 255).all():
        return False, "face 1 not shaded"
    return True, ""


def _tilted_quad(texture=None, uvs=None, colors=None):
    """A quad on the plane z = 2 + 0.8 x (strongly tilted), seen by a 64x48
    camera at the origin; vertex colours linear in world x and y."""
    P = np.array([[-1.0, -0.8], [1.0, -0.8], [1.0, 0.8], [-1.0, 0.8]])
    verts = np.array([(x, y, 2.0 + 0.8 * x) for x, y in P], np.float32)
    faces = np.array([[0, 2, 1], [0, 3, 2]], np.int32)
    m = Mesh(verts, faces)
    m.CalcFaceNormal()
    m.set_normals(np.repeat(m.face_normals[:1], 4, axis=0))
    m.set_normal_indices(faces)
    if colors is None:
        colors = np.stack([(verts[:, 0] + 1) * 100 + 20, (verts[:, 1] + 0.8) * 120 + 10,
                           np.full(4, 200.0)], 1)
    m.set_vertex_colors(colors.astype(np.float32))
    if uvs is not None:
        m.set_uv(uvs)
        m.set_uv_indices(faces)
    if texture is not None:
        m.set_single_material(texture)
    cam = PinholeCamera(64, 48, np.eye(4), (31.5, 23.5), (40.0, 40.0))
    return m, cam


def _plane_hits(cam):
    """World points where each pixel's ray meets z = 2 + 0.8 x (float64)."""
    fx, cx, cy = 40.0, 31.5, 23.5
    ys, xs = np.mgrid[0:cam.height, 0:cam.width].astype(np.float64)
    dx, dy = (xs - cx) / fx, (ys - cy) / fx
    tz = 2.0 / (1.0 - 0.8 * dx)          # z = t, x = dx t
    return dx * tz, dy * tz, tz


def pin_perspective_correct():
    m, cam = _tilted_quad()
    out = render(m, cam, RendererCpuOption(diffuse_color=DiffuseColor.kVertex))
    X, Y, Z = _plane_hits(cam)
    sel = (out["face_id"] > 0) & (np.abs(X) < 0.97) & (np.abs(Y) < 0.77)
    sel |= (out["face_id"] == 0) & False
    cov = out["face_id"] >= 0
    if cov.sum() < 500:
        return False, "quad not covered"
    if not np.allclose(out["depth"][cov & (np.abs(X) < 0.97) & (np.abs(Y) < 0.77)],
                       Z[cov & (np.abs(X) < 0.97) & (np.abs(Y) < 0.77)], rtol=2e-5):
        return False, "depth is not the ray-plane intersection depth"
    exp_r = (X + 1) * 100 + 20
    exp_g = (Y + 0.8) * 120 + 10
    dr = np.abs(out["color"][..., 0][sel] - np.floor(exp_r[sel]))
    dg = np.abs(out["color"][..., 1][sel] - np.floor(exp_g[sel]))
    if dr.max() > 1 or dg.max() > 1:
        return False, f"vertex colour not perspective-correct (max err {dr.max()}, {dg.max()})"
    return True, ""


def pin_texture_sampling():
    """Texture with tex[r, c] = (c*2, r*3, 77): bilinear gives the linear
    function of the texel position, NN the value at the rounded position."""
    tw, th = 101, 71
    cc, rr = np.meshgrid(np.arange(tw), np.arange(th))
    tex = np.stack([cc * 2, rr * 3, np.full_like(cc, 77)], -1).astype(np.uint8)
    uvs = np.array([[0.05, 0.1], [0.95, 0.1], [0.95, 0.9], [0.05, 0.9]], np.float32)
    m, cam = _tilted_quad(texture=tex, uvs=uvs)
    X, Y, _ = _plane_hits(cam)
    # uv is linear in world x, y on the quad
    u = 0.05 + (X + 1) / 2 * 0.9
    v = 0.1 + (Y + 0.8) / 1.6 * 0.8
    fx_, fy_ = u * (tw - 1), (1 - v) * (th - 1)
    for interp in (ColorInterpolation.kBilinear, ColorInterpolation.kNn):
        out = render(m, cam, RendererCpuOption(diffuse_color=DiffuseColor.kTexture, interp=interp))
        sel = (out["face_id"] > 0) & (np.abs(X) < 0.97) & (np.abs(Y) < 0.77)
        if sel.sum() < 300:
            return False, "texture quad not covered"
        if interp == ColorInterpolation.kBilinear:
            er = np.abs(out["color"][..., 0][sel] - np.floor(fx_[sel] * 2))
            eg = np.abs(out["color"][..., 1][sel] - np.floor(fy_[sel] * 3))
            if er.max() > 1 or eg.max() > 1:
                return False, f"bilinear texture wrong ({er.max()}, {eg.max()})"
        else:
            # skip samples within 1e-3 texel of a .5 boundary
            safe = sel & (np.abs(fx_ - np.floor(fx_) - 0.5) > 1e-3) & \
                (np.abs(fy_ - np.floor(fy_) - 0.5) > 1e-3)
            er = out["color"][..., 0][safe] != np.round(fx_[safe]).astype(int) * 2
            eg = out["color"][..., 1][safe] != np.round(fy_[safe]).astype(int) * 3
            if er.any() or eg.any():
                return False, "nearest-neighbour texture is not rounded"
    return True, ""


def _sphere_scene(w=160, h=120, f=125.0, d=3.0, rotated=False, nseg=48):
    m = scenes.make_uv_sphere(nseg, nseg)
    n = m.vertices / np.linalg.norm(m.vertices, axis=1, keepdims=True)
    m.set_normals(n.astype(np.float32))
    m.set_normal_indices(m.vertex_indices)
    m.CalcFaceNormal()
    m.set_vertex_colors(np.full((len(m.vertices), 3), 200.0, np.float32))
    if rotated:
        pos = np.array([1.3, -0.9, -2.5])
        pos = pos / np.linalg.norm(pos) * d
        c2w = look_at_c2w(pos, (0, 0, 0), (0.2, 1.0, 0.1))
    else:
        c2w = np.eye(4)
        c2w[2, 3] = -d
    cam = PinholeCamera(w, h, c2w, (w * 0.5 - 0.5, h * 0.5 - 0.5), (f, f))
    return m, cam


def pin_sphere_closed_forms():
    m, cam = _sphere_scene()
    out = render(m, cam, RendererCpuOption(diffuse_color=DiffuseColor.kVertex))
    cov = out["face_id"] >= 0
    rad = 125.0 / math.sqrt(9.0 - 1.0)
    area = cov.sum()
    if abs(area / (math.pi * rad * rad) - 1) > 0.02:
        return False, f"silhouette area {area} vs {math.pi * rad * rad:.1f}"
    ys, xs = np.nonzero(cov)
    if abs(xs.mean() - 79.5) > 0.5 or abs(ys.mean() - 59.5) > 0.5:
        return False, f"silhouette centroid ({xs.mean():.2f},{ys.mean():.2f}) not at the principal point"
    dc = out["depth"][59:61, 79:81]
    if not (np.abs(dc - 2.0) < 2e-3).all():
        return False, f"near-point depth {dc.ravel()} != d - R"
    # (1-u-v)*200 + u*200 + v*200 truncates to 200 or, after rounding, 199
    col = out["color"][cov & (out["face_id"] > 0)]
    if not np.isin(col, (199, 200)).all() or (col == 200).mean() < 0.5:
        return False, "constant vertex colour is not reproduced"
    return True, ""


def _sphere_hits(cam):
    """Camera-space hit points and normals of each pixel's ray with the unit
    sphere (float64)."""
    a = cam.to_abi()
    R = np.array(a.w2c_R[:], np.float64).reshape(3, 3)
    t = np.array(a.w2c_t[:], np.float64)
    ccam = t  # sphere centre (world origin) in camera coordinates
    ys, xs = np.mgrid[0:cam.height, 0:cam.width].astype(np.float64)
    d = np.stack([(xs - a.cx) / a.fx, (ys - a.cy) / a.fy, np.ones_like(xs)], -1)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    b = d @ ccam
    disc = b * b - (ccam @ ccam - 1.0)
    hit = disc > 0
    s = b - np.sqrt(np.maximum(disc, 0))
    P = d * s[..., None]
    N = P - ccam
    return hit, P, N, d, R


def pin_normals_rotated_camera():
    m, cam = _sphere_scene(rotated=True, nseg=96)
    out = render(m, cam, RendererCpuOption(diffuse_color=DiffuseColor.kVertex,
                                           shading_normal=ShadingNormal.kVertex))
    hit, P, N, d, R = _sphere_hits(cam)
    sel = (out["face_id"] > 0) & hit
    # stay away from the silhouette
    sel &= (-(d * N).sum(-1)) > 0.5
    if sel.sum() < 1000:
        return False, "too few pixels"
    n_out = out["normal"][sel].astype(np.float64)
    n_out /= np.linalg.norm(n_out, axis=-1, keepdims=True)
    cosang = (n_out * N[sel]).sum(-1)
    if cosang.min() < 0.995:
        return False, f"camera-space normal deviates from the analytic normal (min cos {cosang.min():.4f})"
    return True, ""


def pin_lambertian():
    m, cam = _sphere_scene(nseg=96)
    for sn in (ShadingNormal.kFace, ShadingNormal.kVertex):
        out = render(m, cam, RendererCpuOption(diffuse_color=DiffuseColor.kVertex,
                                               shading_normal=sn,
                                               diffuse_shading=DiffuseShading.kLambertian))
        hit, P, N, d, R = _sphere_hits(cam)
        sel = (out["face_id"] > 0) & hit & ((-(d * N).sum(-1)) > 0.3)
        if sn == ShadingNormal.kFace:
            # the winning face's normal, rotated to camera space (float64)
            Nf = m.face_normals[out["face_id"].clip(0)].astype(np.float64) @ R.T
            k = -(d * Nf).sum(-1)
        else:
            k = -(d * N).sum(-1)  # analytic sphere normal (camera space)
        exp = np.floor(200 * k)
        err = np.abs(out["color"][..., 0][sel].astype(float) - exp[sel])
        tol = 1 if sn == ShadingNormal.kFace else 2
        if err.max() > tol:
            return False, f"Lambertian shading off by {err.max()} ({sn.name})"
    return True, ""


def pin_oren_nayar():
    """Oren-Nayar with light = view ray (pixel_shader.h:392-422), sigma 0.3,
    against the formula evaluated in float64 on the same normals."""
    m, cam = _sphere_scene(nseg=64)
    out = render(m, cam, RendererCpuOption(diffuse_color=DiffuseColor.kVertex,
                                           shading_normal=ShadingNormal.kFace,
                                           diffuse_shading=DiffuseShading.kOrenNayar))
    s2 = 0.3 * 0.3
    A = 1 - 0.5 * s2 / (s2 + 0.33)
    B = 0.45 * s2 / (s2 + 0.09)
    a = cam.to_abi()
    C = np.array(a.c2w_R[:], np.float64).reshape(3, 3)
    ys, xs = np.mgrid[0:cam.height, 0:cam.width].astype(np.float64)
    d = np.stack([(xs - a.cx) / a.fx, (ys - a.cy) / a.fy, np.ones_like(xs)], -1)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    r = d @ C.T
    f = out["face_id"]
    sel = f > 0
    n = m.face_normals[f[sel]].astype(np.float64)
    rr = r[sel]
    dl = -(rr * n).sum(-1)
    th = np.arccos(np.clip(dl, -1, 1))
    bl = -n * dl[:, None] - rr
    phi = np.maximum(0, (bl * bl).sum(-1))
    coeff = np.maximum(0, dl) * (A + B * phi * np.sin(th) * np.tan(th))
    exp = np.floor(coeff * 200)
    err = np.abs(out["color"][..., 0][sel] - exp)
    if err.max() > 1:
        return False, f"Oren-Nayar off by {err.max()}"
    return True, ""


def pin_depth_u16():
    d = np.array([2.9, 0.0, 65535.9, 65536.5, -1.5, 1234.999], np.float32)
    exp = np.array([2, 0, 65535, 0, 65535, 1234], np.uint16)
    got = O.depth_to_u16(d)
    if not np.array_equal(got, exp):
        return False, f"ConvertTo to uint16: {got} != {exp}"
    return True, ""


def pin_gbuffer():
    """G-buffer planes on the tilted quad against ray-plane hits: world and
    camera positions, the uv of an affine uv map, barycentrics of the hit
    point in its face; normal output = w2c_R * normal_wld in float32; geo_id
    0 exactly where the mask is set."""
    uvs = None
    m, cam = _tilted_quad()
    uvs = np.stack([(m.vertices[:, 0] + 1.0) / 2.0, (m.vertices[:, 1] + 0.8) / 1.6], 1)
    m.set_uv(uvs.astype(np.float32))
    m.set_uv_indices(m.vertex_indices)
    # a second pose: the same world, camera rotated about y and moved, so that
    # pos_cam differs from pos_wld by a non-symmetric rotation
    th = 0.3
    Ry = np.array([[math.cos(th), 0, math.sin(th)], [0, 1, 0], [-math.sin(th), 0, math.cos(th)]])
    for pose in ("identity", "rotated"):
        c2w = np.eye(4)
        if pose == "rotated":
            c2w[:3, :3] = Ry
            c2w[:3, 3] = (-0.6, 0.1, 0.2)
        cam2 = PinholeCamera(64, 48, c2w, (31.5, 23.5), (40.0, 40.0))
        opt = RendererCpuOption(diffuse_color=DiffuseColor.kVertex)
        rc, out, _ = O.render_gbuffer(m, O.camera_from_abi(cam2.to_abi()), opt,
                                      c2w_t=cam2.c2w[:3, 3])
        if rc != 0:
            return False, f"render_gbuffer rc {rc}"
        a = cam2.to_abi()
        R = np.array(a.w2c_R[:], np.float64).reshape(3, 3)
        t = np.array(a.w2c_t[:], np.float64)
        C = R.T
        o = -R.T @ t                      # camera centre
        ys, xs = np.mgrid[0:48, 0:64].astype(np.float64)
        d = np.stack([(xs - 31.5) / 40.0, (ys - 23.5) / 40.0, np.ones_like(xs)], -1) @ C.T
        # plane z = 2 + 0.8 x  <=>  n . X = 2 with n = (-0.8, 0, 1)
        n = np.array([-0.8, 0.0, 1.0])
        s_ = (2.0 - o @ n) / (d @ n)
        hit = o + d * s_[..., None]
        cov = out["mask"] == 255
        inner = cov & (np.abs(hit[..., 0]) < 0.97) & (np.abs(hit[..., 1]) < 0.77)
        if inner.sum() < 300:
            return False, f"{pose}: quad not covered"
        if not np.allclose(out["pos_wld"][inner], hit[inner], rtol=2e-5, atol=2e-5):
            return False, f"{pose}: pos_wld is not the ray-plane hit"
        pc = hit @ R.T + t
        if not np.allclose(out["pos_cam"][inner], pc[inner], rtol=2e-5, atol=2e-5):
            return False, f"{pose}: pos_cam != w2c_R pos_wld + w2c_t"
        uv_exp = np.stack([(hit[..., 0] + 1) / 2, (hit[..., 1] + 0.8) / 1.6], -1)
        if not np.allclose(out["uv"][inner][:, :2], uv_exp[inner], atol=2e-5):
            return False, f"{pose}: uv is not the affine map at the hit"
        if not np.allclose(out["uv"][inner][:, 2], 1 - uv_exp[inner].sum(-1), atol=4e-5):
            return False, f"{pose}: uv third component"
        V = m.vertices.astype(np.float64)
        fid = out["face_id"]
        for f in (0, 1):
            sel = inner & (fid == f)
            if f == 0 or sel.sum() == 0:
                continue              # face 0 is never shaded
            P0, P1, P2 = V[m.vertex_indices[f]]
            # hit = P0 + b1 (P1 - P0) + b2 (P2 - P0), least squares on the plane
            M = np.stack([P1 - P0, P2 - P0], 1)
            sol = np.linalg.lstsq(M, (hit[sel] - P0).T, rcond=None)[0]
            got = out["bary"][sel]
            if not (np.allclose(got[:, 0], sol[0], atol=2e-5) and
                    np.allclose(got[:, 1], sol[1], atol=2e-5) and
                    np.allclose(got[:, 2], 1 - sol[0] - sol[1], atol=4e-5)):
                return False, f"{pose}: barycentrics are not those of the hit point"
        # normal output is w2c_R (float32, Eigen order) times normal_wld
        nw = out["normal_wld"].astype(np.float32)
        Rf = np.array(a.w2c_R[:], np.float32).reshape(3, 3)
        rot = np.stack([Rf[r, 0] * nw[..., 0] + (Rf[r, 1] * nw[..., 1] + Rf[r, 2] * nw[..., 2])
                        for r in range(3)], -1)
        if not np.array_equal(rot[cov], out["normal"][cov]):
            return False, f"{pose}: normal != w2c_R normal_wld"
        if not (np.all(out["geo_id"][cov] == 0) and np.all(out["geo_id"][~cov] == -1)):
            return False, f"{pose}: geo_id"
        if np.any(out["pos_wld"][~cov] != 0) or np.any(out["bary"][~cov] != 0):
            return False, f"{pose}: uncovered pixels must be zero"
    return True, ""


# ---------------------------------------------------------------------------
# raytracer (oracle_raytrace)


def raytrace(mesh, cam, opt):
    rc, out, _ = O.raytrace(mesh, O.camera_from_abi(cam.to_abi()), opt, c2w_t=cam.c2w[:3, 3])
    assert rc == 0, rc
    return out


def pin_raytrace_brute_force(seeds=range(8)):
    """Ray casting sees the same nearest face as the float64 projective brute
    force (a pixel ray meets a face iff the projected pixel centre lies in the
    projected face), away from edges and depth ties; depth is the camera z."""
    for seed in seeds:
        w, h = [(8, 8), (16, 12), (21, 17), (32, 32)][seed % 4]
        tris = TS.random_triangles(8 + 2 * seed, w, h, 300 + seed)
        mesh = TS.mesh_from_pixels(tris, seed=seed)
        cam = TS.exact_camera(w, h)
        for cull in (False, True):
            opt = RendererCpuOption(diffuse_color=DiffuseColor.kVertex, backface_culling=cull)
            out = raytrace(mesh, cam, opt)
            fid, depth, amb = brute_force(mesh, cam, cull, all_edges_ambiguous=True)
            ok = (out["face_id"] == fid) | amb
            if not ok.all():
                return False, f"raytrace seed {seed} cull {cull}: face_id differs at " \
                              f"{np.argwhere(~ok)[:3].tolist()}"
            sel = (fid >= 0) & ~amb
            if not cull and sel.sum() < 10:
                return False, f"raytrace seed {seed}: scene not covered"
            if not np.allclose(out["depth"][sel], depth[sel], rtol=1e-5):
                return False, f"raytrace seed {seed}: depth differs"
            if not np.array_equal(out["mask"] == 255, out["face_id"] >= 0):
                return False, f"raytrace seed {seed}: mask != covered"
    return True, ""


def pin_raytrace_plane():
    """Tilted quad from an identity and a rotated, translated camera: face
    from the side of the diagonal, depth = camera z of the ray-plane hit times
    depth_scale, vertex colour with the raytracer's barycentric convention
    (u = weight of corner 0, v = weight of corner 1 after the standard
    conversion, colour = (1-u-v) c0 + u c1 + v c2), face 0 shaded."""
    m, _ = _tilted_quad()
    V = m.vertices.astype(np.float64)
    C = m.vertex_colors.astype(np.float64)
    th = 0.3
    Ry = np.array([[math.cos(th), 0, math.sin(th)], [0, 1, 0], [-math.sin(th), 0, math.cos(th)]])
    for pose in ("identity", "rotated"):
        c2w = np.eye(4)
        if pose == "rotated":
            c2w[:3, :3] = Ry
            c2w[:3, 3] = (-0.6, 0.1, 0.2)
        cam = PinholeCamera(64, 48, c2w, (31.5, 23.5), (40.0, 40.0))
        out = raytrace(m, cam, RendererCpuOption(diffuse_color=DiffuseColor.kVertex,
                                                 depth_scale=2.5))
        a = cam.to_abi()
        R = np.array(a.w2c_R[:], np.float64).reshape(3, 3)
        t = np.array(a.w2c_t[:], np.float64)
        o = -R.T @ t
        ys, xs = np.mgrid[0:48, 0:64].astype(np.float64)
        d = np.stack([(xs - 31.5) / 40.0, (ys - 23.5) / 40.0, np.ones_like(xs)], -1) @ R
        n = np.array([-0.8, 0.0, 1.0])
        hit = o + d * ((2.0 - o @ n) / (d @ n))[..., None]
        X, Y = hit[..., 0], hit[..., 1]
        inner = (np.abs(X) < 0.97) & (np.abs(Y) < 0.77) & (np.abs(Y - 0.8 * X) > 2e-3)
        if inner.sum() < 300:
            return False, f"{pose}: quad not covered"
        exp_f = np.where(Y < 0.8 * X, 0, 1)
        if not np.array_equal(out["face_id"][inner], exp_f[inner]):
            return False, f"{pose}: face id is not the side of the diagonal"
        if not (out["mask"][inner] == 255).all():
            return False, f"{pose}: mask"
        zc = (hit @ R.T + t)[..., 2] * 2.5
        if not np.allclose(out["depth"][inner], zc[inner], rtol=2e-5):
            return False, f"{pose}: depth is not depth_scale * camera z of the hit"
        for f in (0, 1):
            sel = inner & (out["face_id"] == f)
            i0, i1, i2 = m.vertex_indices[f]
            M = np.stack([V[i1] - V[i0], V[i2] - V[i0]], 1)
            l1, l2 = np.linalg.lstsq(M, (hit[sel] - V[i0]).T, rcond=None)[0]
            l0 = 1 - l1 - l2
            exp = l2[:, None] * C[i0] + l0[:, None] * C[i1] + l1[:, None] * C[i2]
            err = np.abs(out["color"][sel].astype(np.float64) - np.floor(exp))
            if err.max() > 1:
                return False, f"{pose}: face {f} colour is not (1-u-v)c0 + u c1 + v c2 " \
                              f"with the raytracer's (u, v) (max err {err.max()})"
        nrm = m.normals[0].astype(np.float64)
        if not np.allclose(out["normal"][inner], R @ nrm, atol=1e-5):
            return False, f"{pose}: normal is not w2c_R n"
    return True, ""


def pin_raytrace_sphere():
    """Outside a uv sphere (rotated camera): silhouette area of the closed
    form, hits on the unit sphere, normal = w2c_R face normal. Inside it
    (camera at the centre): every face is a back face, so culling leaves
    nothing and without culling every ray meets the sphere at distance ~1."""
    m, cam = _sphere_scene(rotated=True, nseg=96)
    out = raytrace(m, cam, RendererCpuOption(diffuse_color=DiffuseColor.kVertex,
                                             shading_normal=ShadingNormal.kFace))
    hit, P, N, d, R = _sphere_hits(cam)
    cov = out["face_id"] >= 0
    rad = 125.0 / math.sqrt(9.0 - 1.0)
    if abs(cov.sum() / (math.pi * rad * rad) - 1) > 0.02:
        return False, f"raytraced silhouette area {cov.sum()}"
    sel = cov & hit
    # away from the silhouette (ray-normal cosine > 0.3) the 96-segment sphere
    # is within 2.2e-3 of the unit sphere along the ray
    cosang = -(d * (N / np.linalg.norm(N, axis=-1, keepdims=True))).sum(-1)
    core = sel & (cosang > 0.3)
    s_ = out["depth"][core] / d[core][:, 2]      # distance along the unit ray
    dist = np.linalg.norm(P[core], axis=-1)
    if core.sum() < 1000 or np.abs(s_ - dist).max() > 4e-3:
        return False, "raytraced depth is not on the sphere"
    fnc = m.face_normals[out["face_id"][sel]].astype(np.float64) @ R.T
    if not np.allclose(out["normal"][sel], fnc, atol=1e-5):
        return False, "raytraced normal is not w2c_R face_normal"
    # inside
    c2w = np.eye(4)
    inside = PinholeCamera(48, 40, c2w, (23.5, 19.5), (30.0, 30.0))
    o1 = raytrace(m, inside, RendererCpuOption(diffuse_color=DiffuseColor.kVertex))
    if (o1["mask"] != 0).any() or (o1["face_id"] != -1).any():
        return False, "inside the sphere every face is culled"
    o2 = raytrace(m, inside, RendererCpuOption(diffuse_color=DiffuseColor.kVertex,
                                               backface_culling=False))
    if not (o2["mask"] == 255).all():
        return False, "inside the sphere every ray hits"
    ys, xs = np.mgrid[0:40, 0:48].astype(np.float64)
    dz = 1.0 / np.sqrt(((xs - 23.5) / 30) ** 2 + ((ys - 19.5) / 30) ** 2 + 1)
    s2 = o2["depth"] / dz
    if s2.min() < 0.999 or s2.max() > 1.0001:
        return False, f"inside distances {s2.min()}..{s2.max()} not ~1"
    return True, ""





# ---------------------------------------------------------------------------
# visibility test (oracle_visibility)


def visibility_scene(seed=0):
    """Unit uv sphere (analytic vertex normals) with a square occluder at
    z = -1.6 covering x in [-0.3, 0.9], y in [-0.35, 0.25]; camera at
    (0, 0, -3) looking down +z; a random colour image; mask all 255; depth
    from the oracle rasterizer (sphere and occluder)."""
    sph = scenes.make_uv_sphere(40, 40)
    n = sph.vertices / np.linalg.norm(sph.vertices, axis=1, keepdims=True)
    sv = sph.vertices.astype(np.float32)
    q = np.array([[-0.3, -0.35, -1.6], [0.9, -0.35, -1.6], [0.9, 0.25, -1.6], [-0.3, 0.25, -1.6]],
                 np.float32)
    verts = np.concatenate([sv, q], 0)
    faces = np.concatenate([sph.vertex_indices, len(sv) + np.array([[0, 2, 1], [0, 3, 2]])], 0)
    m = Mesh(verts, faces.astype(np.int32))
    m.CalcFaceNormal()
    normals = np.concatenate([n, np.tile([0.0, 0.0, -1.0], (4, 1))], 0).astype(np.float32)
    m.set_normals(normals)
    m.set_normal_indices(m.vertex_indices)
    m.set_vertex_colors(np.full((len(verts), 3), 128.0, np.float32))
    c2w = np.eye(4)
    c2w[2, 3] = -3.0
    cam = PinholeCamera(96, 80, c2w, (47.5, 39.5), (70.0, 70.0))
    rng = np.random.default_rng(seed)
    color = rng.integers(0, 256, (80, 96, 3), dtype=np.uint8)
    out = render(m, cam, RendererCpuOption(diffuse_color=DiffuseColor.kVertex,
                                           backface_culling=False))
    return m, normals, cam, color, out["depth"], np.full((80, 96), 255, np.uint8)


def vis_options(**kw):
    o = O.OracleVisOptions()
    o.backface_culling_angle_rad_th = float(np.float32(np.float32(90.0) * np.float32(math.pi) /
                                                       np.float32(180.0)))
    o.use_mask, o.use_depth, o.interp = 1, 1, 1
    o.max_viewing_angle = o.backface_culling_angle_rad_th
    o.max_distance_from_camera_to_vertex = float(np.finfo(np.float32).max)
    o.max_depth_difference = float(np.finfo(np.float32).max)
    o.eps = 0.1
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def visibility(m, normals, cam, color, depth, mask, opt):
    a = cam.to_abi()
    oc = O.camera_from_abi(a)
    lin = cam.w2c[:3, :3].astype(np.float32)
    for i in range(9):
        oc.w2c_R[i] = float(lin.reshape(-1)[i])
    return O.visibility(m, oc, cam.c2w[:3, 3], color, depth, mask, opt, m.vertices, m.vertices,
                        normals)


def pin_visibility():
    """Projection, viewing angle, distance, occlusion by the square, colour
    sampling (NN = rounded pixel, bilinear = numpy), intensity weights."""
    m, normals, cam, color, depth, mask = visibility_scene()
    nv = len(m.vertices) - 4
    V = m.vertices[:nv].astype(np.float64)
    C = np.array([0.0, 0.0, -3.0])
    d = V - C
    dist = np.linalg.norm(d, axis=1)
    ray = d / dist[:, None]
    cosang = -(ray * normals[:nv].astype(np.float64)).sum(1)
    # occluded: the segment camera -> vertex crosses z = -1.6 inside the square
    s = (-1.6 - C[2]) / d[:, 2]
    xy = C[:2] + s[:, None] * d[:, :2]
    occ = (s > 0) & (s < 1) & (xy[:, 0] > -0.3) & (xy[:, 0] < 0.9) & (xy[:, 1] > -0.35) & \
        (xy[:, 1] < 0.25)
    near_edge = (np.abs(xy[:, 0] + 0.3) < 0.02) | (np.abs(xy[:, 0] - 0.9) < 0.02) | \
        (np.abs(xy[:, 1] + 0.35) < 0.02) | (np.abs(xy[:, 1] - 0.25) < 0.02)
    for interp in (0, 1):
        r = visibility(m, normals, cam, color, depth, mask, vis_options(interp=interp))[:nv]
        vis = r["visible"] == 1
        clear_front = (cosang > 0.2) & ~occ & ~near_edge
        if not vis[clear_front].all():
            return False, "a clearly visible vertex was rejected"
        if vis[occ & ~near_edge].any():
            return False, "a vertex behind the occluder was reported visible"
        if vis[cosang < -0.05].any():
            return False, "a back-facing vertex was reported visible"
        fx = 70.0 * d[:, 0] / d[:, 2] + 47.5
        fy = 70.0 * d[:, 1] / d[:, 2] + 39.5
        if not (np.allclose(r["px"][vis], fx[vis], atol=1e-3) and
                np.allclose(r["py"][vis], fy[vis], atol=1e-3)):
            return False, "projected position"
        if not np.allclose(np.cos(r["angle"][vis]), cosang[vis], atol=2e-5):
            return False, "viewing angle is not acos(-n . ray)"
        # a ray through a mesh vertex can slip between the faces around it
        # (grazing rays then hit nothing): distance FLT_MAX, the vertex is
        # kept as its own hit (visibility_tester.cc:478-485)
        hit = vis & (r["dist"] < 1e30)
        if hit.sum() < 0.8 * vis.sum() or \
                not np.allclose(r["dist"][hit], dist[hit], rtol=2e-3):
            return False, "distance is not the ray length to the surface"
        px, py = r["px"][vis].astype(np.float64), r["py"][vis].astype(np.float64)
        col = np.stack([r["c0"], r["c1"], r["c2"]], -1)[vis]
        if interp == 0:
            exp = color[np.floor(py + 0.5).astype(int), np.floor(px + 0.5).astype(int)]
            if not np.array_equal(col, exp.astype(np.float32)):
                return False, "nearest colour is not the rounded pixel"
        else:
            x0, y0 = np.floor(px).astype(int), np.floor(py).astype(int)
            x1, y1 = np.minimum(x0 + 1, 95), np.minimum(y0 + 1, 79)
            lu, lv = (px - x0)[:, None], (py - y0)[:, None]
            exp = (1 - lu) * (1 - lv) * color[y0, x0] + lu * (1 - lv) * color[y0, x1] + \
                (1 - lu) * lv * color[y1, x0] + lu * lv * color[y1, x1]
            if np.abs(col - np.floor(exp)).max() > 1:
                return False, "bilinear colour"
        inten = 0.2989 * col[:, 0] + 0.5870 * col[:, 1] + 0.1140 * col[:, 2]
        if not np.allclose(r["intensity"][vis], inten, atol=1e-3):
            return False, "intensity weights"
    return True, ""





# ---------------------------------------------------------------------------
# texture mapping (oracle_texture_map)


def _tex_face(kf, src, dst):
    f = O.OracleTexFace()
    f.keyframe = kf
    for i in range(6):
        f.src[i] = float(src[i])
        f.dst[i] = float(dst[i])
    return f


def pin_texture_map():
    """One face copying a linear-gradient image (bilinear reproduces a linear
    function) into a 128^2 texture: coverage = the target triangle, values =
    the gradient at the barycentric map; overlapping faces: the later one
    wins; PaddingSimple = numpy 3x3 average of the unpadded texture."""
    H, W = 48, 64
    yy, xx = np.mgrid[0:H, 0:W]
    img = np.stack([xx + 2 * yy, 3 * xx, np.full_like(xx, 100)], -1).astype(np.uint8)
    src = np.array([5, 5, 50, 8, 20, 40], np.float32)
    dst = np.array([10, 10, 100, 20, 40, 90], np.float32)
    faces = (O.OracleTexFace * 1)(_tex_face(0, src, dst))
    tex, mask = O.texture_map([img], faces, 128, 128, 0)
    ty, tx = np.mgrid[0:128, 0:128].astype(np.float64)
    d = dst.astype(np.float64).reshape(3, 2)

    def e(a, b, px, py):
        return (px - a[0]) * (b[1] - a[1]) - (py - a[1]) * (b[0] - a[0])
    A = e(d[0], d[1], d[2][0], d[2][1])
    w = [e(d[1], d[2], tx, ty) / A, e(d[2], d[0], tx, ty) / A, e(d[0], d[1], tx, ty) / A]
    inside = (w[0] > 1e-4) & (w[1] > 1e-4) & (w[2] > 1e-4)
    outside = (w[0] < -1e-4) | (w[1] < -1e-4) | (w[2] < -1e-4)
    if not (mask[inside] == 255).all() or (mask[outside] == 255).any():
        return False, "texture coverage is not the target triangle"
    s = src.astype(np.float64).reshape(3, 2)
    sx = w[0] * s[0, 0] + w[1] * s[1, 0] + w[2] * s[2, 0]
    sy = w[0] * s[0, 1] + w[1] * s[1, 1] + w[2] * s[2, 1]
    er = np.abs(tex[..., 0].astype(float) - np.floor(sx + 2 * sy))[inside]
    eg = np.abs(tex[..., 1].astype(float) - np.floor(3 * sx))[inside]
    # a constant 100 can truncate to 99: the four float weights need not sum to 1
    if er.max() > 1 or eg.max() > 1 or not np.isin(tex[..., 2][inside], (99, 100)).all():
        return False, f"texel values are not the bilinear samples (err {er.max()}, {eg.max()})"
    # overlap: the later face wins
    black = np.full((H, W, 3), 10, np.uint8)
    white = np.full((H, W, 3), 200, np.uint8)
    faces2 = (O.OracleTexFace * 2)(_tex_face(0, src, dst), _tex_face(1, src, dst))
    t2, m2 = O.texture_map([black, white], faces2, 128, 128, 0)
    if not np.isin(t2[m2 == 255], (199, 200)).all():
        return False, "overlapping faces: the later face must win"
    # padding
    tp, mp = O.texture_map([img], faces, 128, 128, 3)
    exp_t, exp_m = tex.copy(), mask.copy()
    valid = mask == 255
    for j in range(1, 127):
        for i in range(1, 127):
            if valid[j, i]:
                continue
            win = valid[j - 1:j + 2, i - 1:i + 2]
            if not win.any():
                continue
            avg = tex[j - 1:j + 2, i - 1:i + 2][win].astype(np.float32).sum(0) / np.float32(win.sum())
            exp_t[j, i] = np.floor(avg).astype(np.uint8)
            exp_m[j, i] = 255
    if not (np.array_equal(tp, exp_t) and np.array_equal(mp, exp_m)):
        return False, "PaddingSimple"
    return True, ""



# ---------------------------------------------------------------------------
# differentiable model (oracle_diff_forward_f64 / oracle_diff_backward_f64)


def quad_f64():
    """The tilted quad's corners in float64, exactly on z = 2 + 0.8 x."""
    P = np.array([[-1.0, -0.8], [1.0, -0.8], [1.0, 0.8], [-1.0, 0.8]])
    return np.array([(x, y, 2.0 + 0.8 * x) for x, y in P], np.float64)


def lin_colors(x, y):
    return np.stack([30.0 + 50.0 * x + 20.0 * y, 100.0 - 40.0 * x + 60.0 * y,
                     7.0 + 3.0 * x - 5.0 * y], -1)


QUAD_UVS = np.array([[0.05, 0.1], [0.95, 0.1], [0.95, 0.9], [0.05, 0.9]], np.float32)


def lin_texture(tw=101, th=71):
    """tex[r, c] = (2c + 0.5r + 1, 3r - c + 100, 5): bilinear sampling of it is
    exactly the same affine function of the continuous texel position."""
    cc, rr = np.meshgrid(np.arange(tw, dtype=np.float64), np.arange(th, dtype=np.float64))
    return np.stack([2 * cc + 0.5 * rr + 1, 3 * rr - cc + 100, np.full_like(cc, 5.0)], -1)


def quad_diff_expected(cam, tw, th):
    """Closed forms on the tilted quad for colours _lin_colors of the corners
    and the texture lin_texture(tw, th) with uvs QUAD_UVS: the colour and the
    texture value at each pixel ray's hit with the plane."""
    X, Y, _ = _plane_hits(cam)
    # uv corners as stored (float32): u runs from a to b over x = -1..1, v
    # from c to d over y = -0.8..0.8
    a, b = (float(np.float32(t)) for t in (0.05, 0.95))
    c, d = (float(np.float32(t)) for t in (0.1, 0.9))
    u = a + (X + 1) / 2 * (b - a)
    v = c + (Y + 0.8) / 1.6 * (d - c)
    fxp, fyp = u * (tw - 1), (1 - v) * (th - 1)
    exp_c = lin_colors(X, Y)
    exp_t = np.stack([2 * fxp + 0.5 * fyp + 1, 3 * fyp - fxp + 100, np.full_like(fxp, 5.0)], -1)
    return exp_c, exp_t


def _quad_diff_setup():
    m, cam = _tilted_quad(uvs=QUAD_UVS)
    oc = O.camera_from_abi(cam.to_abi())
    fid = render(m, cam, RendererCpuOption())["face_id"]
    X, Y, _ = _plane_hits(cam)
    inner = (fid > 0) & (np.abs(X) < 0.97) & (np.abs(Y) < 0.77)
    return m, cam, oc, fid, X, Y, inner


def pin_diff_forward():
    """The float64 image model against the ray-plane hit of every pixel ray:
    colours linear in world x, y give that linear function at the hit
    (perspective-correct weights, u = w1 and v = w2 on the right corners);
    a texture linear in (column, row) with uv linear in world x, y gives the
    affine function of the texel position (1 - v) * (rows - 1) (row flip,
    bilinear weights); with both, their product."""
    m, cam, oc, fid, X, Y, inner = _quad_diff_setup()
    if inner.sum() < 500:
        return False, "quad not covered"
    V = quad_f64()
    C = lin_colors(V[:, 0], V[:, 1])
    tex = lin_texture()
    th, tw = tex.shape[:2]
    exp_c, exp_t = quad_diff_expected(cam, tw, th)
    for cols, tx, exp, name in ((C, None, exp_c, "colours"), (None, tex, exp_t, "texture"),
                               (C, tex, exp_c * exp_t, "colours x texture")):
        img = O.diff_forward_f64(oc, m, V, cols, tx, fid)
        err = np.abs(img[inner] - exp[inner]) / np.maximum(np.abs(exp[inner]), 1.0)
        if not err.max() <= 1e-9:
            return False, f"diff forward ({name}) off the ray-plane hit by {err.max():.3g}"
        if np.any(img[fid <= 0] != 0):
            return False, f"diff forward ({name}) wrote an uncovered or face-0 pixel"
    return True, ""


def pin_diff_forward_matches_u8_shader():
    """With the u8 rasterizer's inputs (vertex colours 0..255, a u8 texture as
    floats) and the rasterizer's face ids, floor(diff image) equals the
    reference colourizers' truncated u8 output within 1 on every shaded pixel:
    the float model interpolates the same corners with the same weights."""
    rng = np.random.default_rng(11)
    m, cam = _sphere_scene(rotated=True, nseg=24)
    cols = (rng.random((len(m.vertices), 3)) * 255.0).astype(np.float32)
    m.set_vertex_colors(cols)
    out = render(m, cam, RendererCpuOption(diffuse_color=DiffuseColor.kVertex))
    fid = out["face_id"]
    sh = fid > 0
    img = O.diff_forward_f64(O.camera_from_abi(cam.to_abi()), m, m.vertices, cols, None, fid)
    d = np.abs(np.floor(img[sh]) - out["color"][sh].astype(np.float64))
    if sh.sum() < 1000 or d.max() > 1:
        return False, f"vertex-colour model differs from the u8 shader by {d.max()}"
    texu8 = rng.integers(0, 256, (71, 101, 3), dtype=np.uint8)
    mq, camq = _tilted_quad(texture=texu8, uvs=QUAD_UVS)
    outq = render(mq, camq, RendererCpuOption(diffuse_color=DiffuseColor.kTexture,
                                              interp=ColorInterpolation.kBilinear))
    fq = outq["face_id"]
    sq = fq > 0
    imq = O.diff_forward_f64(O.camera_from_abi(camq.to_abi()), mq, mq.vertices, None,
                             texu8.astype(np.float64), fq)
    dq = np.abs(np.floor(imq[sq]) - outq["color"][sq].astype(np.float64))
    if sq.sum() < 500 or dq.max() > 1:
        return False, f"texture model differs from the u8 bilinear shader by {dq.max()}"
    return True, ""


def pin_diff_backward():
    """oracle_diff_backward_f64 against central differences of the
    pinned float64 forward for L = sum(G * image) on the tilted quad with
    colours and texture: every colour, a spread of texels and every vertex
    coordinate, within 1e-5 relative (h = 1e-5: truncation ~h^2, rounding
    ~1e-16 |L| / h)."""
    m, cam, oc, fid, X, Y, inner = _quad_diff_setup()
    V = quad_f64()
    C = lin_colors(V[:, 0], V[:, 1]) / 100.0
    tex = lin_texture() / 100.0
    rng = np.random.default_rng(5)
    G = rng.normal(size=(cam.height, cam.width, 3)) * inner[..., None]
    gv, gc, gt = O.diff_backward_f64(oc, m, V, C, tex, fid, G)

    def L(vv, cc, tt):
        return float((G * O.diff_forward_f64(oc, m, vv, cc, tt, fid)).sum())

    h = 1e-5
    worst = 0.0
    params = [("v", V, gv, range(V.size)), ("c", C, gc, range(C.size)),
              ("t", tex, gt, rng.choice(np.nonzero(gt.reshape(-1))[0], 24, replace=False))]
    for name, base, g, idx in params:
        for j in idx:
            p, q = base.copy().reshape(-1), base.copy().reshape(-1)
            p[j] += h
            q[j] -= h
            args_p, args_q = [V, C, tex], [V, C, tex]
            k = "vct".index(name)
            args_p[k], args_q[k] = p.reshape(base.shape), q.reshape(base.shape)
            fd = (L(*args_p) - L(*args_q)) / (2 * h)
            an = g.reshape(-1)[j]
            rel = abs(an - fd) / max(abs(fd), 1e-3)
            worst = max(worst, rel)
            if rel > 1e-5:
                return False, f"diff backward d/d{name}[{j}] = {an:.6g}, central {fd:.6g}"
    return True, f"max rel {worst:.2g}"


ALL_PINS = [pin_brute_force, pin_ties_lowest_id, pin_face_zero_not_shaded,
            pin_perspective_correct, pin_texture_sampling, pin_sphere_closed_forms,
            pin_normals_rotated_camera, pin_lambertian, pin_oren_nayar, pin_depth_u16,
            pin_gbuffer, pin_raytrace_brute_force, pin_raytrace_plane, pin_raytrace_sphere,
            pin_visibility, pin_texture_map, pin_diff_forward,
            pin_diff_forward_matches_u8_shader, pin_diff_backward]
