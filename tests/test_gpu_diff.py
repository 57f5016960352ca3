"""GPU: differentiable rendering against the float64 CPU reference and
central finite differences.

Scene: perturbed geodesic icosphere (frequency 71, 100,820 faces), 512x512
camera (fov 45 deg, distance 3), vertex colours and a 1024x1024 float texture
as leaf tensors; loss = squared error against a target rendered with seeded
N(0, 0.1) colour noise, over interior pixels (the pixel and its 4 neighbours
are covered). Tolerances: forward within 1e-5 relative of float64; gradients
within 1e-4 relative (norm-wise) of the float64 backward, and per parameter
within 1e-4 of central differences of the float64 loss: 128 seeded colours and
128 texels (h = 1e-3, exact for their linear dependence) and 64 positions
(Richardson extrapolation of h = 2e-6 and 1e-6, chosen by a conditioning rule
on the geometry), in the colour-only and the textured model; the float64
backward itself within 1e-6 of the same differences."""
import numpy as np
import pytest

import oracle as O
from ugu_amd import RendererCpuOption, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c4():
    import torch
    from ugu_amd.diff import DiffRenderer
    v, f = scenes.geodesic_unit(71)
    mesh = scenes.perturbed_sphere(v, f, texture=None)
    rng = np.random.default_rng(4)
    colors = rng.random((len(mesh.vertices), 3), dtype=np.float32)
    tex = scenes.checker_noise_texture(1024, 5, dtype=np.float32, block=64)
    cam = scenes.ring_cameras(1, 3.0, 11, 12, res_choices=[(512, 512)])[0]
    dr = DiffRenderer(mesh)
    dev = torch.device("cuda")
    P = dict(v=torch.tensor(mesh.vertices, device=dev), c=torch.tensor(colors, device=dev),
             t=torch.tensor(tex, device=dev))
    return mesh, cam, dr, P, colors, tex


def _interior(fid):
    ok = fid > 0
    inner = ok.copy()
    inner[1:, :] &= ok[:-1, :]
    inner[:-1, :] &= ok[1:, :]
    inner[:, 1:] &= ok[:, :-1]
    inner[:, :-1] &= ok[:, 1:]
    inner[0, :] = inner[-1, :] = False
    inner[:, 0] = inner[:, -1] = False
    return inner


def test_visibility_matches_reference_face_id(gpu_available, c4):
    mesh, cam, dr, P, colors, tex = c4
    imgs, fids = dr.render([cam], P["v"], P["c"], P["t"], return_face_ids=True)
    rc, ref, _ = O.render(mesh, O.camera_from_abi(cam.to_abi()), RendererCpuOption(),
                          c2w_t=cam.c2w[:3, 3])
    assert rc == 0
    fid = fids[0].cpu().numpy()
    print("[diff] covered", int((fid >= 0).sum()), flush=True)
    assert np.array_equal(fid, ref["face_id"])


def test_forward_matches_float64(gpu_available, c4):
    mesh, cam, dr, P, colors, tex = c4
    oc = O.camera_from_abi(cam.to_abi())
    for use_c, use_t in ((True, True), (True, False), (False, True)):
        imgs, fids = dr.render([cam], P["v"], P["c"] if use_c else None,
                               P["t"] if use_t else None, return_face_ids=True)
        fid = fids[0].cpu().numpy()
        ref = O.diff_forward_f64(oc, mesh, mesh.vertices, colors if use_c else None,
                                 tex if use_t else None, fid)
        got = imgs[0].cpu().numpy().astype(np.float64)
        err = np.abs(got - ref).max()
        print(f"[diff] forward colours={use_c} texture={use_t}: max abs err {err:.3g}", flush=True)
        assert err <= 1e-5 * max(1.0, np.abs(ref).max()), f"forward err {err}"


def test_forward_closed_forms(gpu_available):
    """The GPU image model against closed forms that do not come from the
    oracle: on a quad tilted in depth, colours linear in world x, y give that
    linear function at each pixel ray's hit with the plane; a texture linear
    in (column, row) with uv linear in world x, y gives the affine function of
    the texel position (1 - v) * (rows - 1); with both, their product.
    Tolerance 1e-5 relative to the product of the factors' magnitudes (each
    at least 1): the inputs are float32."""
    import torch
    import oracle_pins as OP
    from ugu_amd.diff import DiffRenderer
    m, cam = OP._tilted_quad(uvs=OP.QUAD_UVS)
    tex = OP.lin_texture()
    th, tw = tex.shape[:2]
    exp_c, exp_t = OP.quad_diff_expected(cam, tw, th)
    X, Y, _ = OP._plane_hits(cam)
    dev = torch.device("cuda")
    dr = DiffRenderer(m)
    V = torch.tensor(m.vertices, device=dev)
    C = torch.tensor(OP.lin_colors(m.vertices[:, 0].astype(np.float64),
                                   m.vertices[:, 1].astype(np.float64)).astype(np.float32),
                     device=dev)
    T = torch.tensor(tex.astype(np.float32), device=dev)
    for c, t, exp, name in ((C, None, exp_c, "colours"), (None, T, exp_t, "texture"),
                            (C, T, exp_c * exp_t, "colours x texture")):
        imgs, fids = dr.render([cam], V, c, t, return_face_ids=True)
        fid = fids[0].cpu().numpy()
        img = imgs[0].cpu().numpy().astype(np.float64)
        inner = (fid > 0) & (np.abs(X) < 0.97) & (np.abs(Y) < 0.77)
        assert inner.sum() > 500
        # scale of each factor (at least 1): the float32 inputs (corners,
        # colours) carry absolute errors relative to that scale, and a factor
        # near zero multiplies the other factor's scale
        sc = np.maximum(np.abs(exp_c[inner]), 1.0) if c is not None else 1.0
        st = np.maximum(np.abs(exp_t[inner]), 1.0) if t is not None else 1.0
        err = np.abs(img[inner] - exp[inner]) / (sc * st)
        print(f"[diff] closed form {name}: max rel err {err.max():.3g}", flush=True)
        assert err.max() <= 1e-5, name
        assert not img[fid <= 0].any()


def _loss_setup(c4, use_tex=True):
    import torch
    mesh, cam, dr, P, colors, tex = c4
    rng = np.random.default_rng(7)
    target_c = P["c"] + torch.tensor(rng.normal(0, 0.1, colors.shape).astype(np.float32),
                                     device=P["c"].device)
    with torch.no_grad():
        target, fids = dr.render([cam], P["v"], target_c, P["t"] if use_tex else None,
                                 return_face_ids=True)
    fid = fids[0].cpu().numpy()
    mask = torch.tensor(_interior(fid), device=P["v"].device)[..., None].float()
    return target[0], mask, fid


def test_backward_matches_float64_and_finite_differences(gpu_available, c4):
    import torch
    mesh, cam, dr, P, colors, tex = c4
    oc = O.camera_from_abi(cam.to_abi())
    for use_tex in (True, False):
        target, mask, fid = _loss_setup(c4, use_tex)
        v = P["v"].clone().requires_grad_(True)
        c = P["c"].clone().requires_grad_(True)
        t = P["t"].clone().requires_grad_(True) if use_tex else None
        img = dr.render([cam], v, c, t)[0]
        loss = (((img - target) ** 2) * mask).sum()
        loss.backward()
        # float64 reference gradient with the same dL/dimage
        G = (2.0 * (img.detach() - target) * mask).cpu().numpy().astype(np.float64)
        texd = tex.astype(np.float64) if use_tex else None
        gv, gc, gt = O.diff_backward_f64(oc, mesh, mesh.vertices, colors, texd, fid, G)
        pairs = [("verts", v.grad, gv), ("colors", c.grad, gc)]
        if use_tex:
            pairs.append(("texture", t.grad, gt))
        for name, g_gpu, g_ref in pairs:
            g_gpu = g_gpu.cpu().numpy().astype(np.float64)
            rel = np.linalg.norm(g_gpu - g_ref) / np.linalg.norm(g_ref)
            print(f"[diff] tex={use_tex} {name}: |g|={np.linalg.norm(g_ref):.4g} "
                  f"rel err vs float64 {rel:.3g}", flush=True)
            assert rel < 1e-4, f"{name} gradient rel err {rel}"

        # central differences of the float64 fixed-visibility loss
        Tn = target.cpu().numpy().astype(np.float64)
        Mn = mask.cpu().numpy().astype(np.float64)
        inner = Mn[..., 0] > 0
        base = {"verts": mesh.vertices.astype(np.float64),
                "colors": colors.astype(np.float64), "texture": texd}

        def loss64(name, j, step):
            args = dict(base)
            x = base[name].copy().reshape(-1)
            x[j] += step
            args[name] = x.reshape(base[name].shape)
            im = O.diff_forward_f64(oc, mesh, args["verts"], args["colors"], args["texture"], fid)
            return (((im - Tn) ** 2) * Mn).sum()

        def central(name, j, step):
            return (loss64(name, j, step) - loss64(name, j, -step)) / (2 * step)

        # the float64 backward for the loss the differences take: its dL/dimage
        # comes from the float64 image (the GPU's from its float32 image)
        im64 = O.diff_forward_f64(oc, mesh, base["verts"], base["colors"], texd, fid)
        G64 = 2.0 * (im64 - Tn) * Mn
        gv64, gc64, gt64 = O.diff_backward_f64(oc, mesh, mesh.vertices, colors, texd, fid, G64)
        g_gpu_all = {"verts": v.grad.cpu().numpy(), "colors": c.grad.cpu().numpy()}
        g_ref_all = {"verts": gv64, "colors": gc64}
        groups = ["colors", "verts"]
        if use_tex:
            g_gpu_all["texture"] = t.grad.cpu().numpy()
            g_ref_all["texture"] = gt64
            groups.append("texture")
        rng = np.random.default_rng(256)
        for name in groups:
            flat_ref = g_ref_all[name].reshape(-1)
            # samples: parameters whose gradient is not negligible
            cand = np.nonzero(np.abs(flat_ref) > 1e-2 * np.abs(flat_ref).max())[0]
            cand = rng.permutation(cand)
            if name == "verts":
                idx, examined = _conditioned_positions(mesh, cam, fid, inner, cand, H_POS,
                                                       tex.shape[:2] if use_tex else None, 64)
                # the conditioning rule excludes at most a fifth of the candidates
                assert len(idx) == 64 and len(idx) >= 0.8 * examined, (len(idx), examined)
                step = H_POS
            else:
                # linear in colours and texels: central differences are exact
                idx, examined, step = cand[:128], min(len(cand), 128), 1e-3
                assert len(idx) == 128
            if name == "verts":
                # Richardson extrapolation of central differences at h and
                # h/2 (error O(h^4)): the bilinear texture term curves
                # strongly on the scale of a texel
                fd = np.array([(4.0 * central(name, j, step / 2) - central(name, j, step)) / 3.0
                               for j in idx])
            else:
                fd = np.array([central(name, j, step) for j in idx])
            gg = g_gpu_all[name].reshape(-1)[idx].astype(np.float64)
            g64 = flat_ref[idx]
            rel = np.abs(gg - fd) / np.abs(fd)
            rel64 = np.abs(g64 - fd) / np.abs(fd)
            print(f"[diff] tex={use_tex} FD {name} (h={step:g}, {len(idx)} params of {examined} "
                  f"examined): GPU max rel {rel.max():.3g} median {np.median(rel):.3g}; "
                  f"float64 backward max rel {rel64.max():.3g}", flush=True)
            assert rel.max() <= 1e-4, (name, idx[np.argmax(rel)], rel.max())
            assert rel64.max() <= 1e-6, (name, idx[np.argmax(rel64)], rel64.max())


# Step of the position central differences. With it, the screen positions move
# by about 1e-3 pixel, far inside the conditioning limits below, and the
# rounding of the float64 loss (about 1e-16 of 1e3, divided by the step) stays
# near 1e-7 of the gradients compared.
H_POS = 2e-6


def _project(cam, P):
    """Screen x, y and camera depth of world points (float64), as the image
    model projects them: Q = R P + t, x = fx Qx / Qz + cx, y = fy Qy / Qz + cy."""
    R = np.asarray(cam.to_abi().w2c_R, np.float64).reshape(3, 3)
    t = np.asarray(cam.to_abi().w2c_t, np.float64)
    a = cam.to_abi()
    Q = P @ R.T + t
    return (a.fx * Q[..., 0] / Q[..., 2] + a.cx, a.fy * Q[..., 1] / Q[..., 2] + a.cy, Q[..., 2])


def _edge(ax, ay, bx, by, cx, cy):
    return (cx - ax) * (by - ay) - (cy - ay) * (bx - ax)


def _conditioned_positions(mesh, cam, fid, inner, cand, h, tex_hw, n):
    """Position parameters (vertex * 3 + axis) taken from `cand` in order,
    keeping those for which the image model is smooth over [-h, h]; returns
    (kept[:n], number examined). Rule, evaluated on the geometry only, before
    any difference of the loss is taken:
      1. every face incident to the vertex that owns an interior pixel keeps
         its screen area within a relative 1e-3 at both ends of the step (far
         from the area's zero, the pole of the perspective-correct weights);
      2. with a texture, every interior pixel of those faces keeps its
         bilinear cell (floor of the texel coordinates (tw - 1) u and
         (th - 1)(1 - v)) at both ends of the step: bilinear sampling has
         kinks at texel boundaries."""
    F = np.asarray(mesh.vertex_indices, np.int64)
    V = np.asarray(mesh.vertices, np.float64)
    ys, xs = np.nonzero(inner)
    pf = fid[ys, xs].astype(np.int64)
    faces_of = {}
    for fidx, tri in enumerate(F):
        for k in tri:
            faces_of.setdefault(int(k), []).append(fidx)
    order = np.argsort(pf, kind="stable")
    pf_s = pf[order]
    if tex_hw is not None:
        UV = np.asarray(mesh.uv, np.float64)
        UI = np.asarray(mesh.uv_indices, np.int64)
        th, tw = tex_hw

    def state(Vm, faces, px, py, pface):
        X, Y, Z = _project(cam, Vm[F[faces]])  # (m, 3)
        area = _edge(X[:, 0], Y[:, 0], X[:, 1], Y[:, 1], X[:, 2], Y[:, 2])
        cells = None
        if tex_hw is not None and len(px):
            Xp, Yp, Zp = _project(cam, Vm[F[pface]])
            A = _edge(Xp[:, 0], Yp[:, 0], Xp[:, 1], Yp[:, 1], Xp[:, 2], Yp[:, 2])
            e0 = _edge(Xp[:, 1], Yp[:, 1], Xp[:, 2], Yp[:, 2], px, py)
            e1 = _edge(Xp[:, 2], Yp[:, 2], Xp[:, 0], Yp[:, 0], px, py)
            e2 = _edge(Xp[:, 0], Yp[:, 0], Xp[:, 1], Yp[:, 1], px, py)
            q = np.stack([e0 / (A * Zp[:, 0]), e1 / (A * Zp[:, 1]), e2 / (A * Zp[:, 2])], 1)
            b = q / q.sum(1, keepdims=True)
            uvc = UV[UI[pface]]  # (m, 3, 2)
            iu = (b * uvc[..., 0]).sum(1)
            iv = (b * uvc[..., 1]).sum(1)
            cells = np.stack([np.floor(iu * (tw - 1)), np.floor((1.0 - iv) * (th - 1))], 1)
        return area, cells

    kept, examined = [], 0
    for j in cand:
        if len(kept) == n:
            break
        vtx, axis = int(j) // 3, int(j) % 3
        faces = np.array([f for f in faces_of.get(vtx, []) if
                          np.searchsorted(pf_s, f, "right") > np.searchsorted(pf_s, f, "left")],
                         np.int64)
        if len(faces) == 0:
            continue
        examined += 1
        sel = np.concatenate([order[np.searchsorted(pf_s, f, "left"):
                                    np.searchsorted(pf_s, f, "right")] for f in faces])
        px, py, pface = xs[sel].astype(np.float64), ys[sel].astype(np.float64), pf[sel]
        a0, c0 = state(V, faces, px, py, pface)
        ok = True
        for sgn in (-1.0, 1.0):
            Vm = V.copy()
            Vm[vtx, axis] += sgn * h
            a1, c1 = state(Vm, faces, px, py, pface)
            ok = ok and np.all(np.abs(a1 - a0) <= 1e-3 * np.abs(a0))
            if c0 is not None:
                ok = ok and np.array_equal(c0, c1)
        if ok:
            kept.append(int(j))
    return np.array(kept, np.int64), examined


def test_batched_views_gradients_add_up(gpu_available, c4):
    """Gradients of a multi-view batch equal the sum of per-view gradients."""
    import torch
    mesh, cam, dr, P, colors, tex = c4
    cams = scenes.ring_cameras(4, 3.0, 21, 22, res_choices=[(256, 256), (320, 240)])
    c = P["c"].clone().requires_grad_(True)
    imgs = dr.render(cams, P["v"], c, P["t"])
    sum(im.square().sum() for im in imgs).backward()
    g_batch = c.grad.clone()
    g_sum = torch.zeros_like(g_batch)
    for cm in cams:
        c2 = P["c"].clone().requires_grad_(True)
        dr.render([cm], P["v"], c2, P["t"])[0].square().sum().backward()
        g_sum += c2.grad
    rel = (g_batch - g_sum).norm() / g_sum.norm()
    assert rel < 1e-5
