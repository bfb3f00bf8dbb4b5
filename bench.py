#!/usr/bin/env python3
"""Camera-path benchmark on one MI355X: the full-scale forward frame.

Workload (default): the seeded 2,013,432-triangle scene (dome + 100 cylinders + 100 boxes, 256 materials,
16 quad lights), three KB4 cameras 1920x1080 at 64 spp, paths of depth 5 with NEE, then per camera the fused
sensor epilogue (RGGB, Poisson shot + 2 DN Gaussian read noise, 12 bit) and the ISP to 8-bit RGB and YUV 4:2:0
(bilinear demosaic, bilateral NR, sRGB). One step = one frame: every camera traced (the wavefront form by
default), then the three sensor and ISP passes. Inputs are resident in HBM
before the timed region; the timed region is exactly `--steps` frames between two device synchronisations.

Prints one JSON line. `value` is Mrays/s (camera + extension + shadow rays counted on the device). `roofline`
is the dominant kernel (the extension-ray trace) against the roof that binds it, VALU issue: its VALU lane
operations per launch (counters committed under profiles/ for the same shape) over the launch duration measured
here; beside it the HBM fraction and the algorithmic figures of the reference-order BVH2 traversal of the same
seeded paths. After the headline frame the default run adds the rest of the metric on short runs of their own:
the adjoint / forward time ratio (3 KB4 cameras 1280x960, 32 spp), the batched / per-camera speed-up (24 uneven
cameras) and BF16 + refinement against FP32 with its error against FP64 (dense radiosity, N = 32,768, 48 RHS).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "veridien-twin_amd"))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s and spp·Mpix/s per frame on 1 MI355X; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0
# VALU lane operations per second: MI355X's FP32 vector 157.3 TFLOP/s counts an FMA as 2 and a packed FP32 op as 2
# lanes; the library is built without packed FP32, so 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz = 39.3 T lane-op/s
VALU_PEAK_T = 157.3 / 4
# sensorParams' ISP defaults (the reference's 'none' noise-reduction string enables the bilateral)
ISP = dict(demosaic=0, nr=1, nr_sigma_px=0.4, nr_range_scale=2.0, tone=0, chroma420=1)


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(sc, cam, budget_s=12.0, nthreads=16, budget_1t=4.0):
    """The oracle (C restatement of gpu/bvh.js traversal + the shading pieces, FP64) on a bounded sample of the
    same workload: 32x32 tiles of camera 0 around the image centre at the camera's spp, until ~budget_s on
    `nthreads` threads; then the same tiles on one thread for ~budget_1t."""
    import oracle

    t0 = time.time()
    ob = oracle.build_bvh(sc["pos"], sc["idx"])
    import numpy as np
    assert np.array_equal(ob["nodes"].view(np.uint32), sc["nodes"].view(np.uint32))
    log("oracle BVH built in %.1f s (identical to the uploaded one)" % (time.time() - t0))
    w, h = cam["width"], cam["height"]
    tiles = []
    cx, cy = (w // 2) // 32 * 32, (h // 2) // 32 * 32
    for r in range(0, 20):
        for dy in range(-r, r + 1):
            for dx in range(-r, r + 1):
                if max(abs(dx), abs(dy)) != r:
                    continue
                x0, y0 = cx + 32 * dx, cy + 32 * dy
                if 0 <= x0 <= w - 32 and 0 <= y0 <= h - 32:
                    tiles.append((x0, y0))
    rays = 0
    vis = np.zeros(5, np.uint64)
    px = 0
    t1 = time.time()
    for (x0, y0) in tiles:
        _, r, v = oracle.render_tile(sc, cam, x0, y0, 32, 32, 0, cam["spp"], nthreads=nthreads)
        rays += r
        vis += v
        px += 32 * 32
        if time.time() - t1 > budget_s:
            break
    dt = time.time() - t1
    ext = max(int(vis[2]), 1)
    # one thread, same tiles from the centre
    rays1, t2 = 0, time.time()
    for (x0, y0) in tiles:
        _, r, _ = oracle.render_tile(sc, cam, x0, y0, 32, 32, 0, cam["spp"], nthreads=1)
        rays1 += r
        if time.time() - t2 > budget_1t:
            break
    dt1 = time.time() - t2
    model = cpu_model()
    return dict(value=rays / dt / 1e6, unit="Mrays/s", cores=nthreads, kind="port",
                sample="%d pixels (%d 32x32 tiles at the centre of camera 0) x %d spp, depth %d: %d rays in %.1f s on "
                       "%d threads of %s" % (px, px // 1024, cam["spp"], cam["max_depth"], rays, dt, nthreads, model),
                value_1thread=round(rays1 / dt1 / 1e6, 3), host_cpus=os.cpu_count(), cpu_model=model,
                rays=rays, node_visits_per_ray=float(vis[0]) / max(rays, 1), tri_tests_per_ray=float(vis[1]) / max(rays, 1),
                ext_share=ext / max(rays, 1), node_visits_per_ext_ray=float(vis[3]) / ext, tri_tests_per_ext_ray=float(vis[4]) / ext)


def bench_c3(args, ctx=None, sc=None, steps=None, quiet=False):
    """Batching: the full-scale scene with 24 uneven cameras, one persistent launch vs one launch per camera. With
    ctx (the full-scale scene uploaded) and sc, runs on that context and returns the result instead of printing."""
    import numpy as np
    from vtwin import lib, scenes

    own = ctx is None
    steps = steps or args.steps
    sc, cams = scenes.c3(scene=sc)
    if own:
        ctx = lib.Context(0)
        ctx.upload_scene(sc)
    for k, c in enumerate(cams):
        ctx.set_camera(k, c)
    slots = list(range(len(cams)))
    tuned = ctx.autotune(slots)        # untimed: launch shapes of both forms for this frame
    log("autotune: %s" % tuned)
    res = {}
    for name, mode, graph in (("per_camera", lib.VT_RENDER_PER_CAMERA, False), ("batched", lib.VT_RENDER_BATCHED, True),
                              ("wavefront", lib.VT_RENDER_WAVEFRONT, False)):
        ctx.render(slots, mode=mode, graph=graph)
        ctx.sync()
        t1 = time.perf_counter()
        rays = 0.0
        for _ in range(steps):
            ctx.render(slots, mode=mode, graph=graph)
            st = ctx.stats()
            rays += st["rays"]
        t2 = time.perf_counter()
        res[name] = dict(ms_per_frame=1e3 * (t2 - t1) / steps, mrays=rays / (t2 - t1) / 1e6, launches=st["launches"])
        log("%s: %.1f ms/frame, %.1f Mrays/s" % (name, res[name]["ms_per_frame"], res[name]["mrays"]))
    # each camera alone (its own launch): its time and rays. The GPU's throughput when one launch fills it is the
    # best of these; no batching can beat all rays / that throughput, which bounds the batched speed-up
    alone = []
    for k in slots:
        ctx.render([k], mode=lib.VT_RENDER_PER_CAMERA)
        ctx.sync()
        ctx.render([k], mode=lib.VT_RENDER_PER_CAMERA)
        st = ctx.stats()
        alone.append((st["ms_render"], st["rays"]))
    r_full = max(r / (m * 1e-3) for m, r in alone if m > 0)
    t_sum = sum(m for m, _ in alone)
    t_floor = sum(r for _, r in alone) / r_full * 1e3
    # identical images from the three forms
    imgs = []
    for mode in (lib.VT_RENDER_BATCHED, lib.VT_RENDER_PER_CAMERA, lib.VT_RENDER_WAVEFRONT):
        ctx.render(slots, mode=mode)
        imgs.append(ctx.readback(5, lib.VT_KIND_LINEAR, cams[5]["width"], cams[5]["height"]))
    spp_px = sum(c["spp"] * c["width"] * c["height"] for c in cams)
    out = {"metric": METRIC, "config": {"workload": "batching: 24 KB4 cameras 320x240..2048x1536, spp in {1,4,16,64}, depth 5, "
                                                     "2,013,432-tri scene"},
           "value": round(res["wavefront"]["mrays"], 2), "unit": "Mrays/s", "steps": steps,
           "batched_ms_per_frame": round(res["batched"]["ms_per_frame"], 2),
           "wavefront_ms_per_frame": round(res["wavefront"]["ms_per_frame"], 2),
           "per_camera_ms_per_frame": round(res["per_camera"]["ms_per_frame"], 2),
           "speedup_batched_over_per_camera": round(res["per_camera"]["ms_per_frame"] / res["batched"]["ms_per_frame"], 3),
           "speedup_wavefront_over_per_camera": round(res["per_camera"]["ms_per_frame"] / res["wavefront"]["ms_per_frame"], 3),
           "spp_mpix_per_s": round(spp_px / (res["wavefront"]["ms_per_frame"] * 1e-3) / 1e6, 2),
           "images_identical": bool(np.array_equal(imgs[0], imgs[1]) and np.array_equal(imgs[0], imgs[2])),
           "autotune": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in tuned.items()},
           "cameras": [[c["width"], c["height"], c["spp"]] for c in cams],
           "alone_ms": [round(m, 3) for m, _ in alone], "alone_mrays_s": [round(r / (m * 1e-3) / 1e6, 1) for m, r in alone],
           "sum_alone_ms": round(t_sum, 2), "floor_ms_at_best_throughput": round(t_floor, 2),
           "batching_speedup_bound": round(t_sum / t_floor, 3)}
    if not own:
        return out
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_c4(args, quiet=False):
    """Dense node radiosity: 32,768 dome nodes, 48 RHS, Jacobi to 1e-6; FP32 vs BF16 MFMA + FP32 refinement vs FP64."""
    import numpy as np
    from vtwin import lib, scenes

    nodes, sc = scenes.c4_nodes()
    n = nodes.shape[0]
    e = scenes.c4_emission(n)
    ctx = lib.Context(0)
    ctx.upload_scene(sc)
    t0 = time.perf_counter()
    ctx.dense_formfactors(nodes, 0.9, e)
    ctx.sync()
    t_ff = time.perf_counter() - t0
    log("form factors %d x %d in %.2f s" % (n, n, t_ff))
    out = {"metric": METRIC, "config": {"workload": "dense node radiosity: N=32768 dome nodes (visibility through the table box), "
                                                     "3 channels x 16 emitter configs = 48 RHS, Jacobi to 1e-6"},
           "form_factor_s": round(t_ff, 3), "form_factor_rays": n * n}
    sols = {}
    for name, method, inner in (("fp64", lib.VT_DENSE_FP64, 0), ("fp32", lib.VT_DENSE_FP32, 0),
                                ("bf16_refine", lib.VT_DENSE_BF16_REFINE, args.inner)):
        ctx.dense_solve(n, e.shape[1], method, tol=1e-6, max_sweeps=600, inner_sweeps=inner, want_x=False)  # warm
        x, r = ctx.dense_solve(n, e.shape[1], method, tol=1e-6, max_sweeps=600, inner_sweeps=inner)
        sols[name] = x
        out[name] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
        log("%s: %s" % (name, out[name]))
    ref = sols["fp64"].astype(np.float64)
    for name in ("fp32", "bf16_refine"):
        out[name]["rel_err_vs_fp64"] = float(np.linalg.norm(sols[name] - ref) / np.linalg.norm(ref))
    out["speedup_bf16_over_fp32"] = round(out["fp32"]["ms"] / out["bf16_refine"]["ms"], 3)
    out["hbm_read_ceiling_GBs"] = round(ctx.read_bandwidth(2 << 30), 1)   # streaming reads of 2 GiB, the BF16 F's size
    sweep_bytes_fp32 = 4.0 * n * n + 2 * 4 * 48 * n
    out["fp32_sweep_GBs"] = round(sweep_bytes_fp32 / (out["fp32"]["ms_per_sweep"] * 1e-3) / 1e9, 1)
    out["value"] = out["speedup_bf16_over_fp32"]
    out["unit"] = "x (BF16+refinement vs FP32 solve time)"
    ctx.close()
    if quiet:
        return out
    print(json.dumps(out), flush=True)


def bench_c5(args, ctx=None, sc=None, steps=None, warmup=None):
    """Differentiation: targets rendered from the true materials, lights and cameras, parameters perturbed (albedo
    +-20 %, emitters +-30 %, KB4 k1..k4 +-0.02, pose +-1 degree / 5 mm); one forward frame and one path-replay backward pass of the loss 1/2 ||I - I_target||^2 over
    three 1280x960 cameras at 32 spp, with the gradients of materials, lights and the cameras' intrinsics and
    pose; reports the backward / forward time ratio."""
    import numpy as np
    from vtwin import lib, scenes

    own = ctx is None
    steps = steps or args.steps
    warmup = args.warmup if warmup is None else warmup
    sc, cams = scenes.c5(scene=sc)
    if own:
        ctx = lib.Context(0)
        ctx.upload_scene(sc)
    for k, c in enumerate(cams):
        ctx.set_camera(k, c)
    slots = list(range(len(cams)))
    n_mat, n_light = sc["materials"].size // 16, sc["lights"].size // 4
    ctx.render(slots)
    targets = [ctx.readback(k, lib.VT_KIND_LINEAR, c["width"], c["height"]) for k, c in enumerate(cams)]
    r = scenes.Rng(4711)
    mats = sc["materials"].reshape(-1, 16).copy()
    mats[:256, 0:3] *= (1 + r.uniform(-0.2, 0.2, 256 * 3).reshape(256, 3)).astype(np.float32)
    lights = sc["lights"].reshape(-1, 4).copy()
    lights[:, 0:3] *= (1 + r.uniform(-0.3, 0.3, n_light)[:, None]).astype(np.float32)
    ctx.update_materials(mats)
    ctx.update_lights(lights)
    # cameras: KB4 k1..k4 +-0.02, orientation +-1 degree about a random axis, centre +-5 mm
    for k, c in enumerate(cams):
        c2 = dict(c)
        lens = np.asarray(c["lens"], np.float32).copy()
        lens[4:8] += r.uniform(-0.02, 0.02, 4).astype(np.float32)
        axis = r.uniform(-1, 1, 3)
        Rm = np.asarray(c["R"], np.float64).reshape(3, 3).T
        Rm = scenes.rot_axis(axis / np.linalg.norm(axis), np.radians(r.uniform(-1, 1))) @ Rm
        c2["R"] = Rm.T.reshape(-1).astype(np.float32)
        c2["pos"] = (np.asarray(c["pos"], np.float64) + r.uniform(-0.005, 0.005, 3)).astype(np.float32)
        c2["lens"] = lens
        ctx.set_camera(k, c2)
    fwd, bwd, rays_f, rays_b = [], [], [], []
    for it in range(warmup + steps):
        ctx.render(slots, mode=lib.VT_RENDER_WAVEFRONT)
        st = ctx.stats()
        imgs = [ctx.readback(k, lib.VT_KIND_LINEAR, c["width"], c["height"]) for k, c in enumerate(cams)]
        adj = [im - t for im, t in zip(imgs, targets)]
        gm, gl, gc = ctx.backward(slots, adj, n_mat, n_light, cameras=True)
        sb = ctx.stats()
        if it >= warmup:
            fwd.append(st["ms_render"])
            bwd.append(sb["ms_backward"])
            rays_f.append(st["rays"])
            rays_b.append(sb["rays_backward"])
        log("step %d: forward %.1f ms, backward %.1f ms" % (it, st["ms_render"], sb["ms_backward"]))
    loss = 0.5 * sum(float(np.sum((im[:, :, :3] - t[:, :, :3]).astype(np.float64) ** 2)) for im, t in zip(imgs, targets))
    # the gradient points downhill: a small step along -g lowers the loss
    step = 0.05 * float(np.max(np.abs(mats[:256, 0:3]))) / float(np.max(np.abs(gm[:256])) + 1e-30)
    mats2 = mats.copy()
    mats2[:256, 0:3] -= (step * gm[:256]).astype(np.float32)
    ctx.update_materials(mats2)
    ctx.render(slots)
    imgs2 = [ctx.readback(k, lib.VT_KIND_LINEAR, c["width"], c["height"]) for k, c in enumerate(cams)]
    loss2 = 0.5 * sum(float(np.sum((im[:, :, :3] - t[:, :, :3]).astype(np.float64) ** 2)) for im, t in zip(imgs2, targets))
    out = {"metric": METRIC, "config": {"workload": "differentiation: 2,013,432-tri scene, 3 KB4 cameras 1280x960, 32 spp, "
                                                     "depth 5; gradients w.r.t. 256 material colours (rgb), 16 emitters (rgb) "
                                                     "and 3 x (8 KB4 intrinsics + 6 pose)"},
           "value": round(float(np.mean(bwd)) / float(np.mean(fwd)), 3), "unit": "x (adjoint / forward GPU time)",
           "forward_mode": "wavefront", "forward_ms": round(float(np.mean(fwd)), 2), "backward_ms": round(float(np.mean(bwd)), 2),
           "forward_mrays_s": round(float(np.mean(rays_f)) / (float(np.mean(fwd)) * 1e-3) / 1e6, 1),
           "backward_mrays_s": round(float(np.mean(rays_b)) / (float(np.mean(bwd)) * 1e-3) / 1e6, 1),
           "loss_perturbed": loss, "loss_after_gradient_step": loss2, "steps": steps,
           "parameters": int(3 * 256 + 3 * n_light + gc.size),
           "camera_gradients": [[round(float(v), 6) for v in row] for row in gc]}
    # leave the scene as uploaded
    ctx.update_materials(sc["materials"])
    ctx.update_lights(sc["lights"])
    if not own:
        return out
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_copy(args):
    """Counter calibration: one elementwise kernel reading 1 GiB and writing 1 GiB (y = x * 1), three dispatches.
    Run under rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE); the counter's KiB per dispatch over 1,048,576 KiB is
    the factor that scales the trace kernel's counters to bytes."""
    import torch

    n = 256 << 20
    x = torch.ones(n, device="cuda", dtype=torch.float32)
    y = torch.empty_like(x)
    for _ in range(3):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"config": {"workload": "copy calibration: y = x * 1 over 2^28 float32 (1 GiB read, 1 GiB written)"},
                      "bytes_read_per_dispatch": 4 * n, "bytes_written_per_dispatch": 4 * n,
                      "check": float(y[-1].item())}), flush=True)


def bench_sensor(args):
    """The sensor chain alone on one 1920x1080 camera (rendered once, untimed): the raw stage with glare (the
    twin's physics settings: glare 0.1 on the 1/16 level, black 168, 12 bit) and the ISP (edge-aware demosaic,
    bilateral NR, sRGB + S-curve, BT.601 4:2:0), `--steps` passes between two synchronisations. Algorithmic bytes
    per pixel: 22 (raw: 16 B radiance + 4 B RI in, 2 B DN out) + 64.5 (ISP: DN -> green -> RGB -> NR -> RGBA +
    YUV)."""
    import numpy as np
    from vtwin import lib, scenes

    sc, cams = scenes.c1()
    cam = dict(cams[0], width=1920, height=1080, spp=4)
    cam["lens"] = np.array([900.0, 900.0, 959.5, 539.5] + list(cams[0]["lens"][4:]), np.float32)
    sens = scenes.default_sensor(cfa=0, frame=0, seed=1006)
    sens.update(black=168.0, glare_fraction=0.1, glare_sigma=0.25 * float(np.hypot(1920, 1080)) / 16.0,
                glare_level=16)
    isp = dict(demosaic=1, nr=1, nr_sigma_px=0.4, nr_range_scale=2.0, tone=3, s_curve_c=1.39, black_offset=0.014,
               chroma420=1)
    ctx = lib.Context(0)
    ctx.upload_scene(sc)
    ctx.set_camera(0, cam)
    ctx.render([0])
    for _ in range(max(1, args.warmup)):
        ctx.sensor(0, sens)
        ctx.isp(0, isp)
    ctx.sync()
    ms_s, ms_i = [], []
    t1 = time.perf_counter()
    for _ in range(args.steps):
        ctx.sensor(0, sens)
        ctx.isp(0, isp)
        st = ctx.stats()
        ms_s.append(st["ms_sensor"])
        ms_i.append(st["ms_isp"])
    dt = time.perf_counter() - t1
    ctx.close()
    px = 1920 * 1080
    s_ms, i_ms = float(np.mean(ms_s)), float(np.mean(ms_i))
    print(json.dumps({"metric": METRIC, "config": {"workload": "sensor chain: one 1920x1080 camera, raw stage with "
                                                               "glare + ISP (edge-aware, NR, S-curve, 4:2:0)"},
                      "steps": args.steps, "ms_per_pass_wall": round(1e3 * dt / args.steps, 4),
                      "raw_stage_ms": round(s_ms, 4), "isp_ms": round(i_ms, 4),
                      "raw_stage_GBps": round(22 * px / (s_ms * 1e-3) / 1e9, 1),
                      "isp_GBps": round(64.5 * px / (i_ms * 1e-3) / 1e9, 1), "peak_GBps": HBM_PEAK_GBS}), flush=True)


def bench_twin(args):
    """The real twin scene (twin_assets/: the reference's scene.json + meshes + physics.json): 993,618 triangles
    after the 3 mm node subdivision, 14 posed bodies at the default state, the scope's light outlet with its
    radiance profile; the three 1280x720 KB4 tracking cameras at 64 spp, depth 5, NEE, in one frame."""
    import numpy as np
    from vtwin import lib, twin

    root = os.path.join(ROOT, "twin_assets")
    if not os.path.exists(os.path.join(root, "scene.json")):
        print("twin_assets/ missing: python tests/golden/fetch_twin_assets.py", file=sys.stderr)
        sys.exit(2)
    t0 = time.time()
    scene, meshes, physics = twin.load(root)
    prep = twin.prepare(scene, meshes, physics)
    t_prep = time.time() - t0
    cams = twin.cameras(scene, spp=64, depth=5, seed=11, prep=prep)
    poses = twin.poses_from_state(scene, twin.default_state(scene))
    ctx = lib.Context(0)
    t0 = time.time()
    ctx.upload_scene(prep)
    ctx.upload_profiles(prep)
    ctx.set_instances(twin.instances(prep, poses))
    for k, c in enumerate(cams):
        ctx.set_camera(k, c)
    ctx.sync()
    t_up = time.time() - t0
    slots = list(range(len(cams)))
    args.mode = args.mode or "auto"
    mode = {"auto": lib.VT_RENDER_AUTO, "wavefront": lib.VT_RENDER_WAVEFRONT, "batched": lib.VT_RENDER_BATCHED,
            "per-camera": lib.VT_RENDER_PER_CAMERA}[args.mode]
    tuned = ctx.autotune(slots) if args.mode == "auto" else None
    for _ in range(args.warmup):
        ctx.render(slots, mode=mode)
        ctx.sync()
    rays = []
    t1 = time.perf_counter()
    for i in range(args.steps):
        ctx.render(slots, mode=mode)
        st = ctx.stats()
        rays.append(st["rays"])
        log("twin step %d: %.1f ms, %.3g rays" % (i, st["ms_render"], st["rays"]))
    ctx.sync()
    dt = time.perf_counter() - t1
    assert st["stack_overflows"] == 0
    ctx.close()
    spp_px = sum(c["spp"] * c["width"] * c["height"] for c in cams) * args.steps
    print(json.dumps({"metric": METRIC, "config": {"workload": "real twin scene: %d triangles (%d static after the 3 mm "
                                                               "node subdivision), %d posed bodies, scope outlet with "
                                                               "its profile; 3 KB4 cameras 1280x720, 64 spp, depth 5, NEE"
                                                               % (prep["tris"].size // 16, prep["stats"]["static_tris"],
                                                                  len(prep["bodies"])), "mode": args.mode},
                      "value": round(float(np.sum(rays)) / dt / 1e6, 2), "unit": "Mrays/s",
                      "ms_per_frame": round(1e3 * dt / args.steps, 2), "spp_mpix_per_s": round(spp_px / dt / 1e6, 2),
                      "rays_per_frame": float(np.mean(rays)), "prepare_s": round(t_prep, 2),
                      "upload_s": round(t_up, 2),
                      "autotune": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in tuned.items()}
                      if tuned else None}), flush=True)


TWIN_PRESETS = {
    # the reference's operating points (gpu/renderer.js:31-67) and its published whole-frame times on an M4 Air
    "live": dict(spp=1, nl=1, refine=0, nl_refine=1, refine_shadows=False, every=2, iters=1, published_ms=45),
    "balanced": dict(spp=1, nl=2, refine=2, nl_refine=1, refine_shadows=False, every=2, iters=1, published_ms=84),
    "judge": dict(spp=1, nl=4, refine=6, nl_refine=1, refine_shadows=True, every=1, iters=4,
                  published_ms="190-245"),
}


def bench_twin_preset(args, preset):
    """A reference operating point on the real twin scene, its published whole-frame workload (3 x 1280x720,
    tools moving, scope still): per frame the instruments move; every `every`-th frame the node cache takes its
    direct light and `iters` Jacobi steps; the three cameras run the reference estimator (shadeHit over the cache,
    shadow culling) with the preset's sub-samples, light samples and refinement pass; each camera's sensor and ISP
    run with the twin's physics settings. Times the whole frame between two synchronisations."""
    import numpy as np
    from vtwin import lib, twin

    P = TWIN_PRESETS[preset]
    root = os.path.join(ROOT, "twin_assets")
    if not os.path.exists(os.path.join(root, "scene.json")):
        print("twin_assets/ missing: python tests/golden/fetch_twin_assets.py", file=sys.stderr)
        sys.exit(2)
    scene, meshes, physics = twin.load(root)
    prep = twin.prepare(scene, meshes, physics)
    state = twin.default_state(scene)
    cams = twin.cameras(scene, spp=P["spp"], depth=1, seed=0, prep=prep)
    # the reference's defaults: shadow culling, the fixed cameras' first-hit and still-light occlusion caches
    flags = lib.VT_OPT_REFERENCE | lib.VT_OPT_CULL_SHADOWS | (lib.VT_OPT_REFINE_SHADOWS if P["refine_shadows"] else 0)
    if not args.no_caches:
        flags |= lib.VT_OPT_HIT_CACHE | lib.VT_OPT_SHADOW_CACHE
    for c in cams:
        c.update(flags=flags, light_samples=P["nl"], refine=P["refine"], refine_light_samples=P["nl_refine"])
    ctx = lib.Context(0)
    ctx.upload_scene(prep)
    ctx.upload_profiles(prep)
    ctx.set_instances(twin.instances(prep, twin.poses_from_state(scene, state)))
    t0 = time.time()
    info = ctx.radiosity_build()
    interior = ctx.interior_bvh(True)          # the reference builds it with the node cache
    n_mat = np.asarray(prep["materials"]).size // 16
    twin.rho_eff(prep, ctx.radiosity_tables(6, (n_mat, 4), np.float32)[:, :3])
    ctx.radiosity_update(direct=True, iters=96)
    ctx.sync()
    t_build = time.time() - t0
    for k, c in enumerate(cams):
        ctx.set_camera(k, c)
    sens = scenes_default_sensor()
    sens.update(black=168.0, glare_fraction=0.1, glare_sigma=0.25 * float(np.hypot(1280, 720)) / 16.0, glare_level=16,
                exposure=0.0157)
    isp = dict(demosaic=1, nr=1, nr_sigma_px=0.4, nr_range_scale=2.0, tone=3, s_curve_c=1.39, black_offset=0.014,
               chroma420=1)
    slots = list(range(len(cams)))

    solved = [twin.instances(prep, twin.poses_from_state(scene, state))]
    kinds = {}

    def moved_mask(inst):
        """bodies whose pose record differs from the last node update's (_movedSinceSolve)"""
        m = 0
        for b, (x, y) in enumerate(zip(inst, solved[0])):
            same = (x is None and y is None) or (x is not None and y is not None and all(
                np.array_equal(np.asarray(x[k]), np.asarray(y[k])) for k in ("q0", "q1", "p0", "p1") if k in x))
            if not same:
                m |= 1 << b
        return m

    def frame(i):
        st = dict(state, instruments=[dict(q, yaw=q["yaw"] + 0.002 * i, jaw=q["jaw"] + 0.01 * (i % 5))
                                      for q in state["instruments"]])
        inst = twin.instances(prep, twin.poses_from_state(scene, st))
        ctx.set_instances(inst)
        if i % P["every"] == 0:
            # the reference's policy: only tools moved, so the node pass relights the nodes near them
            kind = ctx.radiosity_update_moved(moved_mask(inst), iters=P["iters"])
            kinds[kind] = kinds.get(kind, 0) + 1
            solved[0] = inst
        ctx.render(slots, mode=lib.VT_RENDER_BATCHED)
        for k in slots:
            ctx.sensor(k, dict(sens, frame=i, cam_index=k))
            ctx.isp(k, isp)

    for i in range(max(2, args.warmup)):
        frame(i)
    ctx.sync()
    n = max(args.steps, 20)
    t1 = time.perf_counter()
    for i in range(n):
        frame(i)
    ctx.sync()
    dt = time.perf_counter() - t1
    st = ctx.stats()
    ctx.close()
    px = sum(c["width"] * c["height"] for c in cams)
    print(json.dumps({"metric": METRIC, "config": {"workload": "reference '%s' preset on the real twin scene: 3 KB4 "
                                                               "cameras 1280x720, reference estimator, spp %d, %d light "
                                                               "samples, refine %d x %d light samples, shadow culling%s, interior tree, "
                                                               "tools moving, node radiosity (incremental direct + %d Jacobi steps) "
                                                               "every %s frame, sensor + ISP"
                                                   % (preset, P["spp"], P["nl"], P["refine"], P["nl_refine"],
                                                      "" if args.no_caches else ", fixed-camera hit and occlusion caches",
                                                      P["iters"], "2nd" if P["every"] == 2 else "")},
                      "frames": n, "ms_per_frame": round(1e3 * dt / n, 3), "fps": round(n / dt, 1),
                      "refined_pixel_fraction": round((st["samples"] - px * P["spp"]) / max(1, P["refine"]) / px, 4)
                      if P["refine"] else 0.0,
                      "nodes": info["nodes"], "active_nodes": info["active"], "node_build_s": round(t_build, 2),
                      "node_updates": kinds,
                      "interior_static_tris": interior,
                      "last_render_ms": round(st["ms_render"], 3), "last_isp_ms": round(st["ms_isp"], 3),
                      "reference_published_ms_m4_air": P["published_ms"]}), flush=True)


def scenes_default_sensor():
    from vtwin import scenes
    return scenes.default_sensor(cfa=0, frame=0, seed=1006)


def profile_figures(name):
    """Per-launch counter figures of the dominant kernel, read from its committed rocprofv3 summary (lines
    `key: value`): traffic_bytes_per_launch (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel),
    valu_lane_ops_per_launch (SQ_THREAD_CYCLES_VALU), valu_busy and valu_lane_use (rocprof's VALUBusy and
    VALUUtilization), l2_hit. Missing keys are absent."""
    path = os.path.join(ROOT, "profiles", name)
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        k, sep, v = line.partition(":")
        if sep and k.strip() in ("traffic_bytes_per_launch", "valu_lane_ops_per_launch", "valu_busy", "valu_lane_use",
                                 "l2_hit", "shape"):
            out[k.strip()] = v.strip() if k.strip() == "shape" else float(v)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", choices=["auto", "wavefront", "batched", "per-camera"], default=None,
                    help="wavefront (the default of the headline frame): per-vertex kernels with sorted rays, at the shape the committed "
                         "counter profile was taken at (the default batch, 8-entry LDS stack, refill 16); "
                         "auto (the default of --config twin): the autotuner times both forms on a probe of the "
                         "frame (untimed) and the faster runs; batched: one persistent launch")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline (profiling runs)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the adjoint, batching and dense-radiosity legs of the default run (profiling runs)")
    ap.add_argument("--no-caches", action="store_true",
                    help="twin presets: trace every first hit and static shadow ray (no fixed-camera caches)")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--stack", type=int, default=0, help="LDS stack entries per lane (0: the mode's default)")
    ap.add_argument("--batch", type=int, default=0, help="wavefront: paths per batch (0: the library's default, 134,217,728 paths within an eighth of the "
                         "device's memory)")
    ap.add_argument("--shadow-sort", type=int, default=0, help="wavefront: 1 = shadow rays traced in key order, "
                                                               "2 = in slot order, 0 = the default")
    ap.add_argument("--queues", type=int, default=0, help="ray queues of the trace launches: 1, or 8 (one per XCD); "
                                                          "0: the mode's default")
    ap.add_argument("--config", choices=["c2", "c3", "c4", "c5", "copy", "sensor", "twin", "twin-live", "twin-balanced",
                                         "twin-judge"], default="c2",
                    help="c2: the headline full-scale frame (default); c3: batching; c4: dense radiosity; c5: adjoint; "
                         "copy: counter calibration kernel; sensor: the sensor chain alone; twin: the real twin scene; "
                         "twin-live / twin-balanced / twin-judge: the reference's presets on the twin scene")
    ap.add_argument("--inner", type=int, default=0, help="BF16 sweeps per refinement step (c4; 0: adaptive)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one GPU per box); run --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.config == "c3":
        return bench_c3(args)
    if args.config == "c4":
        return bench_c4(args)
    if args.config == "c5":
        return bench_c5(args)
    if args.config == "copy":
        return bench_copy(args)
    if args.config == "sensor":
        return bench_sensor(args)
    if args.config == "twin":
        return bench_twin(args)
    if args.config.startswith("twin-"):
        return bench_twin_preset(args, args.config[5:])
    import numpy as np
    from vtwin import lib, scenes

    t0 = time.time()
    sc, cams = scenes.c2()
    log("scene: %d triangles, %d BVH2 nodes, built in %.1f s" % (sc["tris"].size // 16, sc["nodes"].size // 16, time.time() - t0))
    ctx = lib.Context(0)
    ctx.upload_scene(sc)
    for k, c in enumerate(cams):
        ctx.set_camera(k, c)
    slots = list(range(len(cams)))
    sensor = [scenes.default_sensor(cfa=0, frame=0, seed=1006) for _ in cams]
    args.mode = args.mode or "wavefront"
    mode = {"auto": lib.VT_RENDER_AUTO, "wavefront": lib.VT_RENDER_WAVEFRONT, "batched": lib.VT_RENDER_BATCHED,
            "per-camera": lib.VT_RENDER_PER_CAMERA}[args.mode]
    tuned = None
    if args.mode == "auto":
        t0 = time.time()
        tuned = ctx.autotune(slots)
        log("autotune (%.1f s): %s" % (time.time() - t0, tuned))

    def frame():
        ctx.render(slots, mode=mode, block=args.block, chunk=args.chunk, stack_lds=args.stack, batch=args.batch,
                   queues=args.queues, shadow_sort=args.shadow_sort)
        for k in slots:
            ctx.sensor(k, sensor[k])
            ctx.isp(k, ISP)

    for i in range(args.warmup):
        frame()
        ctx.sync()
        log("warmup %d done" % i)
    ms_kernel, rays_step, launches, ext_ms, ext_rays, ext_launches, isp_ms = [], [], 0, [], [], 0, []
    ctx.sync()
    t1 = time.perf_counter()
    for i in range(args.steps):
        frame()
        st = ctx.stats()           # syncs; its event times are this frame's
        ms_kernel.append(st["ms_kernel_avg"])
        rays_step.append(st["rays"])
        launches = st["launches"]
        ext_ms.append(st["ms_ext_avg"])
        ext_rays.append(st["rays_extension"])
        ext_launches = st["ext_launches"]
        isp_ms.append(st["ms_isp"])
        log("step %d: %.1f ms trace, %.3g rays" % (i, st["ms_render"], st["rays"]))
    ctx.sync()
    t2 = time.perf_counter()
    assert st["stack_overflows"] == 0
    dt = t2 - t1
    total_rays = float(np.sum(rays_step))
    spp_px = sum(c["spp"] * c["width"] * c["height"] for c in cams) * args.steps
    value = total_rays / dt / 1e6
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * dt / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded mulberry32 scene)",
        "config": {"workload": "full-scale forward: 2,013,432-tri seeded dome scene, 256 materials, 16 lights; "
                               "3 KB4 cameras 1920x1080, 64 spp, depth 5, NEE; RGGB sensor, shot + 2 DN read noise, 12 bit; "
                               "ISP: bilinear demosaic, bilateral NR, sRGB, BT.601 4:2:0",
                   "cameras": len(cams), "resolution": [cams[0]["width"], cams[0]["height"]], "spp": 64, "depth": 5,
                   "mode": args.mode,
                   "form": "wavefront" if ext_launches else ("per-camera" if args.mode == "per-camera" else "batched")},
        "spp_mpix_per_s": round(spp_px / dt / 1e6, 2),
        "rays_per_frame": total_rays / args.steps,
        "trace_ms_per_launch": round(float(np.mean(ms_kernel)), 3),
        "trace_launches_per_frame": launches,
        "isp_ms_per_camera": round(float(np.mean(isp_ms)), 4),
        "extension_trace_launches_per_frame": ext_launches,
        "autotune": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in tuned.items()} if tuned else None,
    }
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline(sc, cams[0], budget_s=args.cpu_budget)
        out["cpu_baseline"] = {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample", "value_1thread",
                                                   "host_cpus", "cpu_model")}
        if ext_launches:
            # the dominant kernel of the wavefront frame: the extension-ray trace (camera and bounce rays)
            vn, vt = cpu["node_visits_per_ext_ray"], cpu["tri_tests_per_ext_ray"]
            rays_launch = float(np.mean(ext_rays)) / ext_launches
            ms_launch = float(np.mean(ext_ms))
            kernel, prof = "k_wf_trace_dyn<false, 8> (camera and extension rays)", "k_wf_trace_c2_full_scale_forward.md"
        else:
            vn, vt = cpu["node_visits_per_ray"], cpu["tri_tests_per_ray"]
            rays_launch = float(np.mean(rays_step)) / max(launches, 1)
            ms_launch = float(np.mean(ms_kernel))
            kernel, prof = "k_render<256>", "k_render_c2_full_scale_forward.md"
        out["roofline"] = roofline(ctx, kernel, prof, vn, vt, rays_launch, ms_launch)
    if not args.no_extras:
        extras(args, ctx, sc, out)
    print(json.dumps(out), flush=True)
    ctx.close()


def roofline(ctx, kernel, prof, vn, vt, rays_launch, ms_launch):
    """The dominant kernel against the roof that binds it: VALU issue. achieved = the VALU lane operations one
    launch executes (SQ_THREAD_CYCLES_VALU per launch, from the committed counter summary of the same shape) over
    the launch's duration measured here with HIP events; the same counters give rocprof's VALUBusy x
    VALUUtilization. Beside it: HBM (FETCH_SIZE + WRITE_SIZE per launch over the same duration, against 8 TB/s),
    and the algorithmic figures of the reference-order BVH2 traversal of the same seeded paths (the oracle's
    visit counts): B_ray = 32 + 16 + 64 V_node + 48 V_tri bytes and 2 x 20 ops per node visit + 30 per triangle
    test."""
    f = profile_figures(prof)
    t = ms_launch * 1e-3
    b_ray = 32 + 16 + 64 * vn + 48 * vt
    ops_ray = 2 * 20 * vn + 30 * vt
    r = {"bound": "valu", "peak": round(VALU_PEAK_T, 2), "unit": "Tlane-op/s", "kernel": kernel,
         "rays_per_launch": round(rays_launch), "ms_per_launch": round(ms_launch, 3), "profile": "profiles/" + prof}
    if "valu_lane_ops_per_launch" in f:
        ach = f["valu_lane_ops_per_launch"] / t / 1e12
        r.update(achieved=round(ach, 3), frac=round(ach / VALU_PEAK_T, 4))
    else:
        r.update(achieved=None, frac=None)
    if "valu_busy" in f and "valu_lane_use" in f:
        r.update(valu_busy=f["valu_busy"], valu_lane_use=f["valu_lane_use"],
                 counter_frac=round(f["valu_busy"] * f["valu_lane_use"], 4))
    traffic = f.get("traffic_bytes_per_launch")
    r["traffic"] = traffic
    if traffic:
        r.update(hbm_GBs=round(traffic / t / 1e9, 1), hbm_frac=round(traffic / t / 1e9 / HBM_PEAK_GBS, 4))
    if "l2_hit" in f:
        r["l2_hit"] = f["l2_hit"]
    if "shape" in f:
        r["profile_shape"] = f["shape"]
    alg_gbs = b_ray * rays_launch / t / 1e9
    alg_ops = ops_ray * rays_launch / t / 1e12
    r.update(achieved_algorithmic_GBs=round(alg_gbs, 1), bytes_per_ray=round(b_ray, 1),
             achieved_algorithmic_Tops=round(alg_ops, 3), algorithmic_ops_frac=round(alg_ops / VALU_PEAK_T, 4),
             ops_per_ray=round(ops_ray, 1), node_visits_per_ray=round(vn, 2), tri_tests_per_ray=round(vt, 2))
    mall = ctx.copy_bandwidth(64 << 20)   # 2 x 64 MiB: resident in the 256 MB Infinity Cache
    r.update(mall_copy_GBs=round(mall, 1))
    return r


def extras(args, ctx, sc, out):
    """The rest of the metric on the same box and scene, each on a short run (untimed by the frame clock above):
    the adjoint / forward time ratio (differentiation configuration), the batched / per-camera speed-up (24
    uneven cameras) and BF16 + refinement against FP32 with its error against FP64 (dense radiosity)."""
    from vtwin import lib
    import numpy as np
    if out["config"]["form"] == "wavefront":
        # the same frame as one hipGraph (captured on its first use, the bin sort between vertices): its time per
        # replay and whether its image equals the eager frame's (camera 0, the last timed frame above)
        t0 = time.time()
        w0, h0 = out["config"]["resolution"]
        slots = list(range(out["config"]["cameras"]))
        eager = ctx.readback(0, lib.VT_KIND_LINEAR, w0, h0)
        ms = []
        for i in range(3):
            ctx.render(slots, mode=lib.VT_RENDER_WAVEFRONT, graph=True)
            st = ctx.stats()
            if i:
                ms.append(st["ms_render"])
        same = bool(np.array_equal(ctx.readback(0, lib.VT_KIND_LINEAR, w0, h0), eager))
        out["wavefront_graph"] = {"ms_per_frame": round(float(np.mean(ms)), 2), "graph_used": st["graph_used"],
                                  "graph_builds": st["graph_builds"], "image_equals_eager": same,
                                  "mrays_s": round(st["rays"] / (np.mean(ms) * 1e-3) / 1e6, 1)}
        log("wavefront graph: %.1f ms per frame, image equals eager %s (%.1f s)" % (np.mean(ms), same, time.time() - t0))
    t0 = time.time()
    c5 = bench_c5(args, ctx=ctx, sc=sc, steps=2, warmup=1)
    out["adjoint_forward_ratio"] = c5["value"]
    out["c5"] = {k: c5[k] for k in ("forward_ms", "backward_ms", "forward_mrays_s", "backward_mrays_s", "parameters",
                                    "loss_perturbed", "loss_after_gradient_step")}
    log("c5: adjoint / forward %.3f (%.1f s)" % (c5["value"], time.time() - t0))
    t0 = time.time()
    c3 = bench_c3(args, ctx=ctx, sc=sc, steps=2)
    out["batched_over_per_camera"] = c3["speedup_batched_over_per_camera"]
    out["c3"] = {k: c3[k] for k in ("batched_ms_per_frame", "wavefront_ms_per_frame", "per_camera_ms_per_frame",
                                    "speedup_batched_over_per_camera", "speedup_wavefront_over_per_camera",
                                    "spp_mpix_per_s", "images_identical", "sum_alone_ms",
                                    "floor_ms_at_best_throughput", "batching_speedup_bound")}
    log("c3: batched / per-camera %.3f (%.1f s)" % (c3["speedup_batched_over_per_camera"], time.time() - t0))
    t0 = time.time()
    c4 = bench_c4(args, quiet=True)
    out["bf16_over_fp32_solve"] = c4["speedup_bf16_over_fp32"]
    out["c4"] = {"fp32_ms": c4["fp32"]["ms"], "bf16_refine_ms": c4["bf16_refine"]["ms"],
                 "fp32_ms_per_sweep": c4["fp32"]["ms_per_sweep"], "bf16_ms_per_sweep": c4["bf16_refine"]["ms_per_sweep"],
                 "bf16_rel_err_vs_fp64": c4["bf16_refine"]["rel_err_vs_fp64"],
                 "fp32_rel_err_vs_fp64": c4["fp32"]["rel_err_vs_fp64"], "fp32_sweep_GBs": c4["fp32_sweep_GBs"],
                 "hbm_read_ceiling_GBs": c4["hbm_read_ceiling_GBs"]}
    log("c4: bf16 / fp32 %.3f (%.1f s)" % (c4["speedup_bf16_over_fp32"], time.time() - t0))


if __name__ == "__main__":
    main()
