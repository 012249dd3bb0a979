"""Benchmark of the MI355X ray-tracing hot path.

Headline (`value`): closest-hit Mrays/s of the incoherent diffuse path tracer
on the 2,097,152-triangle fBm-displaced octahedral sphere (amplitude 0.1),
1920x1080, 16 spp, depth 8 — every ray segment (camera rays and cosine
bounces) is one closest-hit query of the radix-tree LBVH.  One step = one full
frame; the BVH build and scene generation are outside the timed region.

Also reported (`secondary`): LBVH build time of that mesh, primary-ray
closest-hit Mrays/s (65,536-triangle soup, 1024^2 primary rays), the batched
build of 256 meshes, volumetric path samples/s (256^3 fBm grid, 512^2, 64 spp,
depth 32) and the differentiable forward + adjoint cost against the plain
forward (128^3 grid, 256^2, 16 spp).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--fast]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cam_of(N, pos, look, fov=45.0):
    return N.Camera.make(pos, look, (0.0, 1.0, 0.0), fov)


def timed(fn, steps, warmup, sync):
    for _ in range(warmup):
        fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    sync()
    return (time.perf_counter() - t0) / steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--fast", action="store_true", help="headline only (no secondary workloads)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--graph", action="store_true", help="replay each frame from a captured hipGraph")
    ap.add_argument("--shape", default="",
                    help="fixed launch shape block,blocks_per_cu,stack_mode,ray_fetch for every frame "
                         "(skips the autotuner; profiling runs use the shape it picked)")
    ap.add_argument("--profile-pass", action="store_true",
                    help="one calibration copy + one headline frame, no JSON (for counter runs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the hot path has no multi-GPU exchange to time "
              "(rays and paths would shard by image tiles with a replicated BVH). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from fledge_amd import _native as N
    from fledge_amd import api

    dev = torch.device("cuda", 0)
    ctx = api.Context(0)
    sync = torch.cuda.synchronize

    # ------------------------------------------------------ headline workload
    W, H, SPP, DEPTH = 1920, 1080, 16, 8
    log("generating the 2M-triangle displaced octasphere ...")
    v, ind = api.gen_octasphere(9, 0.1, seed=7, base_freq=2.0)
    ntris = ind.size // 3
    dv, di = torch.from_numpy(v).to(dev), torch.from_numpy(ind).to(dev)
    sync()
    t0 = time.perf_counter()
    bvh = api.LBVH.build(ctx, dv, di)
    sync()
    build_ms_first = (time.perf_counter() - t0) * 1e3
    build_ms = timed(lambda: api.LBVH.build(ctx, dv, di).close(), 3, 1, sync) * 1e3
    cam2 = cam_of(N, (0.0, 0.0, -3.0), (0.0, 0.0, 0.0))

    shape = {}
    if args.shape:
        b, bpc, sm, rf = (int(x) for x in args.shape.split(","))
        shape = {"trace_block": b, "blocks_per_cu": bpc, "stack_mode": sm, "ray_fetch": rf}
    info = bvh.pathtrace(cam2, W, H, SPP, DEPTH, **shape)  # also warms the queues
    segments = info["segments"]
    log("frame: %d paths, %d ray segments" % (info["paths"], segments))

    if args.profile_pass:
        # for hardware-counter runs: a plain device copy of known size (the
        # calibration of FETCH_SIZE / WRITE_SIZE), then one headline frame
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        sync()
        log("calibration copy: %d bytes read, %d bytes written" % (src.numel() * 4, dst.numel() * 4))
        del src, dst
        bvh.pathtrace(cam2, W, H, SPP, DEPTH, **shape)
        sync()
        log("profile pass done: one frame of %d ray segments" % segments)
        return

    # launch shape picked by the library's autotuner on this frame (untimed)
    tuning = None
    if shape:
        tuning = {"block_size": shape["trace_block"], "blocks_per_cu": shape["blocks_per_cu"],
                  "stack_mode": shape["stack_mode"], "ray_fetch": shape["ray_fetch"], "fixed": True}
    elif not args.no_tune:
        tuning = bvh.autotune(cam2, W, H, SPP, DEPTH)
        log("autotune: %s" % tuning)

    def frame(use_graph=args.graph):
        cfg = N.PtCfg()
        cfg.width, cfg.height, cfg.spp, cfg.max_depth = W, H, SPP, DEPTH
        cfg.use_graph = 1 if use_graph else 0
        if shape:
            cfg.trace_block, cfg.blocks_per_cu = shape["trace_block"], shape["blocks_per_cu"]
            cfg.stack_mode, cfg.ray_fetch = shape["stack_mode"], shape["ray_fetch"]
        ctx.check(ctx.lib.flg_pathtrace_diffuse(ctx.h, bvh.h, cam2, cfg, None, None, None, None))

    # timed region: K frames, HIP-event kernel timers recorded live
    for _ in range(args.warmup):
        frame()
    sync()
    ctx.profiling(True)
    ctx.reset_profiling()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame()
    sync()
    dt = (time.perf_counter() - t0) / args.steps
    trace_ms, trace_launches = ctx.kernel_time("pathtrace_bounce")
    total_ms, _ = ctx.kernel_time("pathtrace_total")
    if args.graph:  # a replayed graph has no per-launch timers: one timed eager frame
        ctx.reset_profiling()
        frame(use_graph=False)
        sync()
        trace_ms, trace_launches = ctx.kernel_time("pathtrace_bounce")
        total_ms, _ = ctx.kernel_time("pathtrace_total")
        trace_ms, trace_launches = trace_ms * args.steps, trace_launches * args.steps
        total_ms *= args.steps
    ctx.profiling(False)
    value = segments / dt / 1e6

    # algorithmic bytes per ray segment: 32 (ray) + 8 (hit) + 64 per node
    # visit + 36 per triangle test, visits counted by the instrumented kernel
    cnt = bvh.pathtrace(cam2, W, H, SPP, DEPTH, count_visits=True, **shape)
    vn = cnt["node_visits"] / segments
    vt = cnt["tri_tests"] / segments
    bytes_per_ray = 32 + 8 + 64 * vn + 36 * vt
    launches_per_frame = trace_launches / args.steps
    avg_launch_ms = trace_ms / trace_launches
    bytes_per_launch = bytes_per_ray * segments / launches_per_frame
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(),
                "kernel": ("k_pt_bounce (closest-hit traversal + diffuse bounce, one launch per bounce)"
                           if not tuning or tuning.get("ray_fetch", 1) != 2 else
                           "k_pt_bounce (first bounce) then k_pt_trace + k_pt_shade per bounce "
                           "(closest-hit traversal with per-wave ray refill, then the diffuse bounce); "
                           "a launch = one bounce"),
                "algorithmic_bytes_per_ray": round(bytes_per_ray, 2),
                "node_visits_per_ray": round(vn, 3), "tri_tests_per_ray": round(vt, 3),
                "avg_launch_ms": round(avg_launch_ms, 4), "launches_per_frame": launches_per_frame,
                "trace_kernel_share": round(trace_ms / max(total_ms, 1e-9), 4),
                "launch_shape": tuning, "graph": bool(args.graph)}

    secondary = {"lbvh_build_2M_ms": round(build_ms, 3), "lbvh_build_2M_first_ms": round(build_ms_first, 3),
                 "lbvh_build_2M_mtris_per_s": round(ntris / build_ms / 1e3, 2)}

    # -------------------------------------------------- secondary workloads
    if not args.fast:
        # 65,536-triangle soup, 1024^2 jittered primary rays
        v0, i0 = api.gen_soup(65536)
        b0 = api.LBVH.build(ctx, torch.from_numpy(v0).to(dev), torch.from_numpy(i0).to(dev))
        cam0 = cam_of(N, (0.5, 0.5, -1.5), (0.5, 0.5, 0.0))
        rays0 = api.camera_rays(ctx, cam0, 1024, 1024)
        hits0 = torch.empty((1024 * 1024, 2), dtype=torch.int32, device=dev)
        t = timed(lambda: b0.trace(rays0, hits0), 20, 5, sync)
        secondary["soup65k_primary_closest_hit_mrays_per_s"] = round(1024 * 1024 / t / 1e6, 2)
        secondary["soup65k_build_ms"] = round(timed(
            lambda: api.LBVH.build(ctx, torch.from_numpy(v0).to(dev), torch.from_numpy(i0).to(dev)).close(),
            5, 2, sync) * 1e3, 3)
        # batched build: 256 meshes, log-uniform 1K..256K triangles, one batch
        sizes = api.gen_batch_sizes(256)
        meshes = []
        for k, n in enumerate(sizes):
            vv, ii = api.gen_soup(int(n), initstate=42, initseq=1000 + k)
            meshes.append((torch.from_numpy(vv).to(dev), torch.from_numpy(ii).to(dev)))
        sync()

        def batch():
            for h in api.LBVH.build_batched(ctx, meshes):
                h.close()
        tb = timed(batch, 3, 1, sync)
        secondary["batched_build_256_meshes_ms"] = round(tb * 1e3, 3)
        secondary["batched_build_total_tris"] = int(sizes.sum())
        secondary["batched_build_mtris_per_s"] = round(int(sizes.sum()) / tb / 1e6, 2)
        # the scene accel over the same 256 meshes, each moved to a cell of a
        # 16 x 16 grid (pitch 1.25), against one LBVH over all their
        # triangles: 1024^2 camera rays from above
        placed, voff, allv, alli = [], 0, [], []
        for k, (vv, ii) in enumerate(meshes):
            off = torch.tensor([1.25 * (k % 16), 1.25 * (k // 16), 0.0], dtype=torch.float32, device=dev)
            pv = (vv.view(-1, 3) + off).reshape(-1)
            placed.append((pv, ii))
            allv.append(pv)
            alli.append(ii + voff)
            voff += pv.numel() // 3
        blas = api.LBVH.build_batched(ctx, placed)
        sacc = api.SceneAccel(ctx, blas)
        cam_s = N.Camera.make((10.0, 10.0, 30.0), (10.0, 10.0, 0.0), (0.0, 1.0, 0.0), 40.0)
        rays_s = api.camera_rays(ctx, cam_s, 1024, 1024)
        hits_s = torch.empty((1024 * 1024, 4), dtype=torch.int32, device=dev)
        ts_ = timed(lambda: sacc.trace(rays_s, hits_s), 10, 2, sync)
        secondary["scene_256_meshes_closest_hit_mrays_per_s"] = round(1024 * 1024 / ts_ / 1e6, 2)
        secondary["scene_256_meshes_hit_fraction"] = round(float((hits_s[:, 0] >= 0).float().mean()), 4)
        merged = api.LBVH.build(ctx, torch.cat(allv), torch.cat(alli))
        hits_m = torch.empty((1024 * 1024, 2), dtype=torch.int32, device=dev)
        tm_ = timed(lambda: merged.trace(rays_s, hits_m), 10, 2, sync)
        secondary["merged_single_lbvh_closest_hit_mrays_per_s"] = round(1024 * 1024 / tm_ / 1e6, 2)
        sacc.close()
        merged.close()
        for h in blas:
            h.close()
        del placed, allv, alli, meshes
        # volumetric path tracing: 256^3 fBm grid, 512x512, 64 spp, depth 32
        g1 = api.gen_fbm_grid(256, seed=1)
        vol = api.Volume(ctx, torch.from_numpy(g1).to(dev), [0, 0, 0, 1, 1, 1], 6.4, 1.6, 0.3)
        sc1 = api.vol_scene(cam0)
        out1 = torch.empty((512, 512, 3), dtype=torch.float32, device=dev)
        _, st1 = vol.render(sc1, 512, 512, 64, 32, out=out1, stats=True)
        t1 = timed(lambda: vol.render(sc1, 512, 512, 64, 32, out=out1), 3, 1, sync)
        paths1 = 512 * 512 * 64
        t1b = timed(lambda: vol.render(sc1, 512, 512, 64, 32, out=out1, schedule=1), 3, 1, sync)
        secondary["vol_ms_per_frame_lane_per_path"] = round(t1b * 1e3, 3)
        secondary["vol_path_samples_per_s"] = round(paths1 / t1, 1)
        secondary["vol_ms_per_frame"] = round(t1 * 1e3, 3)
        secondary["vol_lookups_per_path"] = round(st1["lookups"] / paths1, 3)
        lb = 32 * st1["lookups"] + 12 * 512 * 512
        secondary["vol_algorithmic_GBps"] = round(lb / t1 / 1e9, 2)
        if hasattr(ctx.lib, "flg_volpath_adjoint"):
            try:
                from fledge_amd import diff
                secondary.update(diff.bench_adjoint(ctx, timed, sync))
            except Exception as e:  # reported, not hidden
                secondary["adjoint_error"] = repr(e)

        # surfaces + media (VolPathIntegrator): the 8,388,608-triangle displaced
        # octahedral sphere (amplitude 0.05) as the interface around a 512^3
        # grid whose 8^3 bricks are active with probability 0.15 and hold fBm,
        # 1920x1080, 64 spp, depth 16, one frame
        vc, ic = api.gen_octasphere(10, 0.05, seed=7, base_freq=2.0)
        gc = api.gen_fbm_grid(512, seed=3)
        bricks = np.random.default_rng(4).random((64, 64, 64)) < 0.15
        gc *= np.repeat(np.repeat(np.repeat(bricks, 8, 0), 8, 1), 8, 2)
        bc = api.LBVH.build(ctx, torch.from_numpy(vc).to(dev), torch.from_numpy(ic).to(dev))
        gct = torch.from_numpy(gc).to(dev)
        volc = api.Volume(ctx, gct, [0, 0, 0, 1, 1, 1], 6.4, 1.6, 0.3)
        gmax = float(gc.max())
        del gc, vc, ic
        scc = api.vol_scene(cam2)
        outc = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
        _, stc = volc.render_mesh(bc, scc, W, H, 64, 16, out=outc, stats=True)
        tc = timed(lambda: volc.render_mesh(bc, scc, W, H, 64, 16, out=outc), 1, 0, sync)
        secondary["combined_8M_512cube_ms_per_frame"] = round(tc * 1e3, 2)
        secondary["combined_path_samples_per_s"] = round(W * H * 64 / tc, 1)
        secondary["combined_lookups_per_path"] = round(stc["lookups"] / (W * H * 64), 3)
        secondary["combined_config"] = ("VolPathIntegrator, 8,388,608-triangle interface mesh around a 512^3 grid "
                                        "(8^3 bricks, 15% active), 1920x1080, 64 spp, depth 16")
        volc.close()
        # the same scene through the differentiable estimator (sigma_t 8 per
        # unit density, albedo 0.8: the medium above), forward with the walk
        # tape, then the adjoint of an L2-like loss with respect to the
        # active bricks' voxels
        from fledge_amd import diff
        dvc = diff.DiffVolume(ctx, gct, g=0.3)
        layout = diff.BrickLayout(ctx, gct, 8)
        del gct
        tape = diff.new_tape(ctx, W, H, 64, 16)
        lpc = torch.empty(W * H * 64, dtype=torch.float32, device=dev)
        dlc = torch.full((H, W, 3), 1.0 / (W * H * 3), dtype=torch.float32, device=dev)
        gb = layout.zeros()
        gac = torch.zeros(1, dtype=torch.float32, device=dev)
        mu_c = 8.0 * gmax
        tf = timed(lambda: diff.mesh_forward(dvc, bc, scc, W, H, 64, 16, 8.0, 0.8, mu_c, out=outc, lpath=lpc,
                                             tape=tape), 1, 1, sync)
        ta = timed(lambda: diff.mesh_adjoint(dvc, bc, scc, W, H, 64, 16, 8.0, 0.8, mu_c, lpc, dlc, grad=gb,
                                             galpha=gac, tape=tape, bricks=layout), 1, 0, sync)
        secondary["combined_diff_forward_ms"] = round(tf * 1e3, 2)
        secondary["combined_adjoint_ms"] = round(ta * 1e3, 2)
        secondary["combined_active_bricks"] = layout.n_active
        secondary["combined_fwd_adj_over_forward"] = round((tf + ta) / tc, 3)
        del tape, lpc, gb
        bc.close()

    # ------------------------------------------------------ CPU baseline
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(v, ind, W, H, SPP, DEPTH, args.cpu_seconds)

    out = {
        "metric": "Mrays/s closest-hit (2M tris); vol path samples/s; fwd+adjoint ms/iter vs plain",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "incoherent diffuse closest-hit on a 2,097,152-triangle fBm-displaced "
                               "octahedral sphere (amplitude 0.1), 1920x1080, 16 spp, depth 8",
                   "triangles": ntris, "rays_per_step": segments, "paths_per_step": W * H * SPP,
                   "max_depth": DEPTH},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(out))


def measured_traffic():
    """HBM bytes per k_pt_bounce launch from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE runs of `bench.py --profile-pass`, each scaled
    by what the counter reads for a 1 GiB device copy); None if absent."""
    path = os.path.join(ROOT, "profiles", "pt_bounce_octasphere2M_traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_baseline(v, ind, W, H, SPP, DEPTH, seconds):
    """The oracle's LBVH + per-path DFS traversal (the restated fledge CPU path)
    on a bounded sample of the same frame: whole image rows spread over the
    frame, sized to about `seconds` of work on this host's cores."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc

    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    o = orc.Lbvh(v, ind)
    cam10 = [0.0, 0.0, -3.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 45.0]
    row = W * SPP
    # calibrate on one central row
    t0 = time.perf_counter()
    _, _, _, r = o.diffuse_render(cam10, W, H, SPP, DEPTH, (H // 2) * row, (H // 2 + 1) * row,
                                  record=False, nthreads=cores)
    per_row = time.perf_counter() - t0
    nrows = int(max(4, min(H, seconds / max(per_row, 1e-6))))
    rows = np.unique(np.linspace(0, H - 1, nrows).astype(int))
    t0 = time.perf_counter()
    rays = 0
    for y in rows:
        _, _, _, r = o.diffuse_render(cam10, W, H, SPP, DEPTH, int(y) * row, int(y + 1) * row,
                                      record=False, nthreads=cores)
        rays += r
    dt = time.perf_counter() - t0
    return {"value": round(rays / dt / 1e6, 3), "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": "%d of %d image rows (evenly spaced), %d paths, %d ray segments, %.1f s"
                      % (len(rows), H, len(rows) * row, rays, dt)}


if __name__ == "__main__":
    main()
