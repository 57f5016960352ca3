#!/usr/bin/env python3
"""Benchmark of the gfx950 rasterizer (one JSON line on stdout).

Headline workload (default): the 1,310,720-face perturbed icosphere (subdivision
level 8, 2048x2048 RGB8 texture, bilinear, vertex normals, back-face culling)
rendered from 64 ring cameras of mixed resolution (320x240 .. 1920x1080) in
one batched pipeline (one hipGraph replay per step), all five outputs written
to device memory. A step = one forward render of the 64 views; inputs are
resident in HBM before timing starts.

metric  = Mtri*views/s = faces * views / step time / 1e6
extras  = Mpix/s, the per-view GPU loop (64 separate Render calls) and its
          ratio, the single 1080p view, the 32.5k-face UV sphere, per-kernel
          HIP-event timings (roofline), the raytracer on the same views, the
          differentiable legs and the CPU oracle on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--quick]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def alloc_outputs(torch, cams, which=("color", "depth", "normal", "mask", "face_id")):
    bufs = []
    for c in cams:
        h, w = c.height, c.width
        b = {}
        if "color" in which:
            b["color"] = torch.empty((h, w, 3), dtype=torch.uint8, device="cuda")
        if "depth" in which:
            b["depth"] = torch.empty((h, w), dtype=torch.float32, device="cuda")
        if "normal" in which:
            b["normal"] = torch.empty((h, w, 3), dtype=torch.float32, device="cuda")
        if "mask" in which:
            b["mask"] = torch.empty((h, w), dtype=torch.uint8, device="cuda")
        if "face_id" in which:
            b["face_id"] = torch.empty((h, w), dtype=torch.int32, device="cuda")
        bufs.append(b)
    return bufs


def algorithmic_bytes(mesh, cams, texture_touched=0):
    """Inputs once per launch + 24 B per output pixel (colour 3, depth 4,
    normal 12, mask 1, face_id 4), for texture colouring with vertex normals."""
    V, F = len(mesh.vertices), len(mesh.vertex_indices)
    b_in = 12 * V + 12 * F + 12 * F          # positions, faces, face normals
    b_in += 12 * len(mesh.normals) + 12 * F  # vertex normals + normal indices
    b_in += 8 * len(mesh.uv) + 12 * F + 4 * F  # uv, uv indices, material ids
    b_in += texture_touched
    px = sum(c.width * c.height for c in cams)
    return b_in, 24 * px


PROFILE_SUMMARY = os.path.join(ROOT, "profiles", "raster_64view_icosphere8.md")


def measured_traffic(tiling):
    """HBM bytes per k_raster launch of this workload: FETCH_SIZE x 2 +
    WRITE_SIZE x 1 (factors from a 1 GiB copy in the same profile) of the
    committed rocprofv3 summary, if it was taken with the same tiling."""
    tag = f"{tiling['tile_w']}x{tiling['tile_h']}:{'p' if tiling['persistent'] else 'g'}"
    try:
        with open(PROFILE_SUMMARY) as f:
            for line in f:
                parts = line.split()
                if (len(parts) == 4 and parts[0] == "traffic_per_launch_bytes"
                        and parts[1] == "k_raster" and parts[2] == tag):
                    return int(parts[3])
    except OSError:
        pass
    return None


def time_calls(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def cpu_baseline(mesh, cams, opt, budget_s):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    done, t_total = 0, 0.0
    for c in cams:
        # wall time of the whole Render (ray table + vertex + face + resolve loops)
        t0 = time.perf_counter()
        rc, _, _ = O.render(mesh, O.camera_from_abi(c.to_abi()), opt, c2w_t=c.c2w[:3, 3])
        t_total += time.perf_counter() - t0
        assert rc == 0
        done += 1
        if t_total > budget_s:
            break
    return done, t_total, O.num_threads()


def backward_roofline(dr, step, mesh, F, tex_texels, colors, positions):
    """k_bwd_items of one profiled fit step: algorithmic bytes = 28 B per item
    (the 16-B (face, view, pixel) item and the pixel's 12-B dL/dimage) + the
    inputs once (positions 12 V, faces 12 F; colours 12 V; uv 8 U, uv indices
    12 F and the float texture 12 T) + the gradients once (12 V per vertex
    quantity, 12 T); updates = per-item gradient contributions before the
    wavefront and LDS pre-reduction (9 colour, 9 position, 12 texel values)."""
    dr.enable_profiling(True)
    step()
    tm = dr.timings()
    dr.enable_profiling(False)
    n, ms = tm["bwd_items"], tm["bwd_items_ms"]
    V, U = len(mesh.vertices), len(mesh.uv)
    b = 28 * n + 12 * V + 12 * F
    per_item = 0
    if colors:
        b += 24 * V
        per_item += 9
    if positions:
        b += 12 * V
        per_item += 9
    if tex_texels:
        b += 8 * U + 12 * F + 24 * tex_texels
        per_item += 12
    gbs = b / (ms * 1e-3) / 1e9
    return {"kernel": "k_bwd_items", "bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS, "algorithmic_bytes": b, "items": n,
            "kernel_ms": ms, "sort_ms": tm["bwd_sort_ms"],
            "updates_per_s": n * per_item / (ms * 1e-3)}


def diff_benchmarks(torch, np, scenes, mesh, cams, args):
    """Forward+backward legs of the differentiable renderer (float colours and
    float texture as leaf tensors, L2 image loss)."""
    from ugu_amd.diff import DiffRenderer
    out = {}
    dev = torch.device("cuda")
    # the L2 loss as one reduction (and one gradient kernel) and Adam as one
    # fused update: the torch side of a step stays a few passes over the
    # images and the texture
    l2 = lambda a, b: torch.nn.functional.mse_loss(a, b, reduction="sum")

    def fit_step(dr, cam_list, v, c, t, targets, opt=None):
        imgs = dr.render(cam_list, v, c, t)
        loss = sum(l2(im, tg) for im, tg in zip(imgs, targets))
        if opt is not None:
            opt.zero_grad(set_to_none=True)
        loss.backward()
        if opt is not None:
            opt.step()
        return loss

    # -- 100k-face perturbed icosphere (geodesic frequency 71), one 512x512 view
    log("diff: 100k faces, 512x512, fwd+bwd")
    gv, gf = scenes.geodesic_unit(71)
    m4 = scenes.perturbed_sphere(gv, gf, texture=None)
    rng = np.random.default_rng(4)
    cols = torch.tensor(rng.random((len(m4.vertices), 3), dtype=np.float32), device=dev)
    tex = torch.tensor(scenes.checker_noise_texture(1024, 5, dtype=np.float32), device=dev)
    cam4 = scenes.ring_cameras(1, 3.0, 11, 12, res_choices=[(512, 512)])[0].to_abi()
    dr4 = DiffRenderer(m4)
    v4 = torch.tensor(m4.vertices, device=dev)
    with torch.no_grad():
        noise = torch.tensor(rng.normal(0, 0.1, cols.shape).astype(np.float32), device=dev)
        tgt4 = dr4.render([cam4], v4, cols + noise, tex)
    pv, pc, pt = (x.clone().requires_grad_(True) for x in (v4, cols, tex))
    t4 = time_calls(torch, lambda: fit_step(dr4, [cam4], pv, pc, pt, tgt4), args.steps, 3)
    out["diff_100k_512"] = {"iters_per_s": 1.0 / t4, "ms": t4 * 1e3, "faces": len(m4.vertex_indices)}
    del dr4

    # -- 1.31M-face mesh, the 64 ring views: batched vs per-view loop
    log("diff: 64 views fwd+bwd, batched vs per-view")
    dr = DiffRenderer(mesh)
    vv = torch.tensor(mesh.vertices, device=dev)
    cc = torch.tensor(np.random.default_rng(6).random((len(mesh.vertices), 3), dtype=np.float32),
                      device=dev)
    tt = torch.tensor(scenes.checker_noise_texture(2048, 7, dtype=np.float32), device=dev)
    abi = [c.to_abi() for c in cams]
    with torch.no_grad():
        tg_flat, _ = dr.render_flat(abi, vv, cc * 0.9, tt)
    tg, off = [], 0
    for c in cams:
        n = c.width * c.height * 3
        tg.append(tg_flat[off:off + n].view(c.height, c.width, 3))
        off += n
    pv, pc, pt = (x.clone().requires_grad_(True) for x in (vv, cc, tt))

    def batched_step():
        flat, _ = dr.render_flat(abi, pv, pc, pt)
        l2(flat, tg_flat).backward()
    tb = time_calls(torch, batched_step, max(3, args.steps // 2), 2)

    def per_view():
        for a, g in zip(abi, tg):
            fit_step(dr, [a], pv, pc, pt, [g])
    tl = time_calls(torch, per_view, 3, 1)
    out["diff_64view"] = {"batched_iters_per_s": 1.0 / tb, "per_view_iters_per_s": 1.0 / tl,
                          "batched_vs_loop": tl / tb, "batched_ms": tb * 1e3, "loop_ms": tl * 1e3,
                          "roofline": backward_roofline(dr, batched_step, mesh, len(mesh.vertex_indices),
                                                        tt.shape[0] * tt.shape[1], True, True)}

    # -- inverse texturing: 32 views (16 at 512x512, 16 at 1024x768), GT
    #    2048x2048 float texture, learned texture from 0.5 grey, Adam lr 1e-2
    log("diff: inverse texturing, 32 views, 200 Adam iterations")
    vc = scenes.inverse_texturing_cameras()
    a5 = [c.to_abi() for c in vc]
    gt_tex = torch.tensor(scenes.checker_noise_texture(2048, 8, dtype=np.float32), device=dev)
    with torch.no_grad():
        tg5, _ = dr.render_flat(a5, vv, None, gt_tex)
    learn = torch.full_like(gt_tex, 0.5).requires_grad_(True)
    opt = torch.optim.Adam([learn], lr=1e-2, fused=True)
    seen = None
    n_iter = 200
    for it in range(n_iter + 3):
        if it == 3:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
        flat, _ = dr.render_flat(a5, vv, None, learn)
        opt.zero_grad(set_to_none=True)
        l2(flat, tg5).backward()
        opt.step()
        if seen is None:
            # texels that receive gradient from at least one view
            seen = (learn.grad.abs().sum(-1) > 0) if learn.grad is not None else None
    torch.cuda.synchronize()
    t5 = (time.perf_counter() - t0) / n_iter
    with torch.no_grad():
        m = seen if seen is not None else torch.ones(gt_tex.shape[:2], dtype=torch.bool, device=dev)
        mse = ((learn.clamp(0, 1) - gt_tex) ** 2).mean(-1)[m].mean().item()

    def tex_step():
        flat, _ = dr.render_flat(a5, vv, None, learn)
        l2(flat, tg5).backward()
    roof5 = backward_roofline(dr, tex_step, mesh, len(mesh.vertex_indices),
                              gt_tex.shape[0] * gt_tex.shape[1], False, False)
    out["inverse_texturing_32view"] = {"iters_per_s": 1.0 / t5, "ms_per_iter": t5 * 1e3,
                                       "iterations": n_iter,
                                       "psnr_db": 10.0 * math.log10(1.0 / max(mse, 1e-12)),
                                       "texels_seen": int(m.sum().item()),
                                       "roofline": roof5}
    del dr
    return out


def raytrace_benchmarks(torch, scenes, mesh, cams, cam1080, opt, args):
    """Raytracer legs (ugu_hip_raytrace): the 64 ring views in one launch and
    the single 1080p view of the same mesh, all five outputs in device
    memory; plus the kernel time from HIP events and the host BVH build."""
    from ugu_amd import Raytracer
    rt = Raytracer(opt)
    rt.set_mesh(mesh)
    t0 = time.perf_counter()
    assert rt.PrepareMesh()
    prep = time.perf_counter() - t0
    abi = [c.to_abi() for c in cams]
    bufs = alloc_outputs(torch, cams)
    px = sum(c.width * c.height for c in cams)
    t64 = time_calls(torch, lambda: rt.render_device(abi, bufs), max(3, args.steps // 2), 1)
    rt.enable_profiling(True)
    rt.render_device(abi, bufs)
    k64 = rt.timings()["raster_ms"]
    rt.enable_profiling(False)
    c1 = [cam1080.to_abi()]
    b1 = alloc_outputs(torch, [cam1080])
    t1 = time_calls(torch, lambda: rt.render_device(c1, b1), args.steps, args.warmup)
    F = len(mesh.vertex_indices)
    return {"raytrace": {
        "views64_ms": t64 * 1e3, "views64_mrays_per_s": px / t64 / 1e6,
        "views64_mtri_views_per_s": F * len(cams) / t64 / 1e6,
        "views64_kernel_ms": k64,
        "single_1080p_ms": t1 * 1e3, "single_1080p_mrays_per_s": 1920 * 1080 / t1 / 1e6,
        "prepare_mesh_and_bvh_s": prep}}


def visibility_benchmarks(torch, np, mesh, cam1080, opt):
    """Visibility test (ugu_hip_visibility_test): every vertex of the mesh
    against the 1080p keyframe (colour = the rasterizer's render, depth and
    mask from it), vertex-wise and face-wise items; host arrays in and out,
    so the time includes the PCIe copies."""
    from ugu_amd import Rasterizer, visibility as V
    r = Rasterizer(opt)
    r.set_mesh(mesh)
    assert r.PrepareMesh()
    out = r.render_batch([cam1080])[0]
    kf = V.Keyframe(0, out["color"], out["depth"], out["mask"], cam1080)
    t = V.VisibilityTester()
    t.set_data(mesh.vertices, mesh.normals, mesh.vertex_indices)
    t0 = time.perf_counter()
    assert t.PrepareData()
    prep = time.perf_counter() - t0
    t.set_keyframe(kf)
    res = {"prepare_s": prep}
    for mode, items in (("vertex", (t.vertices_, t.vertices_, t.normals_)),
                        ("facewise", t.facewise_items())):
        t.test_items(*items)  # warm-up
        t0 = time.perf_counter()
        reps = 5
        for _ in range(reps):
            rec, _ = t.test_items(*items)
        dt = (time.perf_counter() - t0) / reps
        res[mode] = {"items": len(items[0]), "ms": dt * 1e3, "mitems_per_s": len(items[0]) / dt / 1e6,
                     "visible": int(rec["visible"].sum())}
    return {"visibility_1080p": res}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline only")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--tile", default="auto",
                    help="auto | 8x8 | 16x16 | 32x8, optionally with :p (persistent) or :g (grid)")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiable legs")
    ap.add_argument("--diff-only", action="store_true",
                    help="run only the differentiable legs (prints their JSON)")
    ap.add_argument("--raytrace-only", action="store_true",
                    help="run only the raytracer legs (prints their JSON)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run a 1 GiB device-to-device copy first (counter calibration)")
    ap.add_argument("--shading", default="texture", choices=("texture", "vertex-normals",
                                                             "face-normals"),
                    help="options of the headline leg, for attributing k_raster's reads under "
                         "a counter profile: texture (the workload), vertex-normals "
                         "(DiffuseColor kNone: no texels), face-normals (kNone and "
                         "ShadingNormal kFace: no per-face shading records either)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the views "
              "would shard across GPUs with no exchange; one GPU's share is what is "
              "measured)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from ugu_amd import (ColorInterpolation, DiffuseColor, Rasterizer, RendererCpuOption,
                         ShadingNormal, scenes)

    assert torch.cuda.is_available(), "bench needs a GPU"
    torch.cuda.init()
    if args.calib_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    log("building scenes")
    mesh, cam2 = scenes.scene_icosphere8()
    cams = scenes.multiview_cameras(64)
    F = len(mesh.vertex_indices)
    if args.diff_only:
        print(json.dumps(diff_benchmarks(torch, np, scenes, mesh, cams, args)))
        return
    opt = RendererCpuOption(diffuse_color=DiffuseColor.kTexture,
                            interp=ColorInterpolation.kBilinear)
    if args.raytrace_only:
        res = raytrace_benchmarks(torch, scenes, mesh, cams, cam2, opt, args)
        res.update(visibility_benchmarks(torch, np, mesh, cam2, opt))
        print(json.dumps(res))
        return
    head_opt = opt
    if args.shading != "texture":
        head_opt = RendererCpuOption(
            diffuse_color=DiffuseColor.kNone, interp=ColorInterpolation.kBilinear,
            shading_normal=(ShadingNormal.kFace if args.shading == "face-normals"
                            else ShadingNormal.kVertex))
    r = Rasterizer(head_opt)
    r.set_mesh(mesh)
    assert r.PrepareMesh()
    abi_cams = [c.to_abi() for c in cams]
    bufs = alloc_outputs(torch, cams)

    if args.tile == "auto":
        log("autotuning tile size / bin capacity")
        tiling = r.autotune(cams, repeats=3)
    else:
        shape, _, mode = args.tile.partition(":")
        tw, th = (int(x) for x in shape.split("x"))
        r.set_tiling(tw, th, 1024, mode != "g")
        r.render_device(abi_cams, bufs)  # sizes the bins
        tiling = r.get_tiling()
    log(f"tiling {tiling}")

    # ---- headline: batched 64-view forward (graph replay), called through a
    # binding of the output tensors (arguments marshalled once, as a C/C++
    # caller of ugu_hip_render_batch passes them); render_device, which
    # rebuilds the ctypes arguments every call, is timed beside it
    log("timing batched 64-view forward")
    binding = r.bind_device(abi_cams, bufs)
    t_batch = time_calls(torch, lambda: r.render_bound(binding), args.steps, args.warmup)
    t_unbound = time_calls(torch, lambda: r.render_device(abi_cams, bufs, graph=True),
                           max(3, args.steps // 2), 1)
    px_total = sum(c.width * c.height for c in cams)
    value = F * len(cams) / t_batch / 1e6
    mpix = px_total / t_batch / 1e6

    # ---- per-kernel timings (HIP events on the library's stream, eager launches)
    r.enable_profiling(True)
    kt = {"clear_ms": 0.0, "setup_bin_ms": 0.0, "raster_ms": 0.0}
    n_prof = 5
    for _ in range(n_prof):
        r.render_device(abi_cams, bufs, graph=False)
        t = r.timings()
        for k in kt:
            kt[k] += t[k] / n_prof
    r.enable_profiling(False)
    b_in, b_out = algorithmic_bytes(mesh, cams)
    traffic = measured_traffic(tiling) if args.shading == "texture" else None
    dom = max(("setup_bin_ms", "raster_ms"), key=lambda k: kt[k])
    raster_gbs = (b_in + b_out) / (kt["raster_ms"] * 1e-3) / 1e9
    pipeline_ms = kt["clear_ms"] + kt["setup_bin_ms"] + kt["raster_ms"]
    log(f"kernels {kt} dominant {dom}")

    extra = {}
    if not args.quick:
        # ---- per-view GPU loop: 64 separate Render calls (device outputs)
        log("timing per-view loop")
        t_loop = time_calls(torch, lambda: r.render_device(abi_cams, bufs, batched=False),
                            max(3, args.steps // 4), 1)
        extra["per_view_loop_ms"] = t_loop * 1e3
        extra["batched_vs_loop"] = t_loop / t_batch
        # ---- single 1080p view of the same mesh (config: large single view)
        r2 = Rasterizer(opt)
        r2.set_mesh(mesh)
        assert r2.PrepareMesh()
        c2 = [cam2.to_abi()]
        b2 = alloc_outputs(torch, [cam2])
        r2.autotune([cam2], repeats=3)
        t2 = time_calls(torch, lambda: r2.render_device(c2, b2), args.steps, args.warmup)
        extra["single_1080p"] = {"ms": t2 * 1e3, "mtri_per_s": F / t2 / 1e6,
                                 "mpix_per_s": 1920 * 1080 / t2 / 1e6,
                                 "tiling": r2.get_tiling()}
        del r2
        # ---- UV sphere 128x128, 640x480, vertex colours
        m1, cam1 = scenes.scene_uv_sphere_128()
        r1 = Rasterizer(RendererCpuOption(diffuse_color=DiffuseColor.kVertex))
        r1.set_mesh(m1)
        assert r1.PrepareMesh()
        b1 = alloc_outputs(torch, [cam1])
        c1 = [cam1.to_abi()]
        r1.autotune([cam1], repeats=3)
        t1 = time_calls(torch, lambda: r1.render_device(c1, b1), args.steps, args.warmup)
        extra["uv_sphere_640x480"] = {"ms": t1 * 1e3,
                                      "mtri_per_s": len(m1.vertex_indices) / t1 / 1e6,
                                      "mpix_per_s": 640 * 480 / t1 / 1e6}
        del r1

    if not args.quick:
        log("raytracer legs")
        extra.update(raytrace_benchmarks(torch, scenes, mesh, cams, cam2, opt, args))
        log("visibility leg")
        extra.update(visibility_benchmarks(torch, np, mesh, cam2, opt))
    if not args.quick and not args.no_diff:
        extra.update(diff_benchmarks(torch, np, scenes, mesh, cams, args))

    cpu = None
    if not args.no_cpu:
        log("CPU oracle baseline (bounded sample)")
        done, t_cpu, threads = cpu_baseline(mesh, cams, opt, args.cpu_budget)
        cpu_val = F * done / t_cpu / 1e6
        cpu = {"value": cpu_val, "unit": "Mtri*views/s", "cores": threads, "kind": "port",
               "sample": f"first {done} of the 64 views, C++ restatement of Render() "
                         f"(single-threaded vertex/face/resolve loops as the reference; its "
                         f"ray table on {threads} OpenMP threads), {t_cpu:.1f} s",
               "host_cpus": os.cpu_count()}
        extra["gpu_vs_cpu"] = value / cpu_val

    line = {
        "metric": "Mtri*views/s (forward, 64 views batched)",
        "value": value,
        "unit": "Mtri*views/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_batch * 1e3,
        "render_device_ms": t_unbound * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "icosphere-L8 1,310,720 faces, 64 ring views 320x240..1920x1080 "
                               "(56.5 Mpx), texture 2048^2 RGB8 bilinear, vertex normals, "
                               "culling, 5 outputs",
                   "faces": F, "views": len(cams), "pixels": px_total, "tiling": tiling,
                   "shading": args.shading},
        "mpix_per_s": mpix,
        "roofline": {"bound": "hbm", "kernel": "k_raster",
                     "achieved": raster_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": raster_gbs / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": b_in + b_out,
                     "kernel_ms": kt["raster_ms"]},
        "kernels_ms": kt,
        "pipeline_gbs": (b_in + b_out) / (pipeline_ms * 1e-3) / 1e9,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
