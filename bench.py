"""Benchmark of the MI355X ray-tracing hot path.

Headline (`value`): closest-hit Mrays/s of the incoherent diffuse path tracer
on the 2,097,152-triangle fBm-displaced octahedral sphere (amplitude 0.1),
1920x1080, 16 spp, depth 8 — every ray segment (camera rays and cosine
bounces) is one closest-hit query of the radix-tree LBVH.  One step = one full
frame; the BVH build and scene generation are outside the timed region.

The metric's other legs are reported under `legs`, each with its own value,
roofline (algorithmic bytes over the kernel's HIP-event time, against the
8 TB/s HBM peak) and CPU baseline (the CPU restatement of the reference's
algorithm timed on this host's cores over a bounded sample):
  vol_path       volumetric path samples/s (256^3 fBm grid, 512^2, 64 spp, depth 32)
  fwd_adjoint    differentiable forward + path-replay adjoint, ms per iteration
                 and its ratio to the plain forward (128^3 grid, 256^2, 16 spp),
                 at the uniform start state and at a non-uniform fBm state
  lbvh_build     LBVH build of the 2M-triangle mesh and of the 65,536-triangle soup
  batched_build  256 meshes (log-uniform 1K..256K triangles) in one build
  primary_rays   closest hit of 1024^2 jittered primary rays on the 65,536-triangle soup
  combined       the 8,388,608-triangle interface mesh around a 512^3 brick grid
                 (1920x1080, 64 spp, depth 16): forward, differentiable forward
                 and adjoint over the active bricks

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
CPU_SHARE = 16  # host cores of one GPU's share on the measurement machine
BUILD_BYTES_PER_TRI = 248  # 36 verts + 8 key/index + 4 passes x 16 + 16 gathers + 64 node + 60 refit


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


def roofline(achieved_gbs, kernel, bytes_per_unit, unit_name, **extra):
    r = {"bound": "hbm", "achieved": round(achieved_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(achieved_gbs / HBM_PEAK_GBS, 5), "traffic": None, "kernel": kernel,
         "algorithmic_bytes_per_" + unit_name: bytes_per_unit}
    r.update(extra)
    return r


def kernel_ms(ctx, names, calls):
    """Mean HIP-event time per call of the library timers `names` (summed)."""
    tot = 0.0
    for n in names:
        ms, _ = ctx.kernel_time(n)
        tot += ms
    return tot / max(calls, 1)


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(n, CPU_SHARE))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--fast", action="store_true", help="headline only (no other legs)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU work of the headline's baseline")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--graph", action="store_true", help="replay each frame from a captured hipGraph")
    ap.add_argument("--shape", default="",
                    help="fixed launch shape block,blocks_per_cu,stack_mode,ray_fetch[,ring_depth] for every "
                         "frame (skips the autotuner; profiling runs use the shape it picked)")
    ap.add_argument("--profile-pass", action="store_true",
                    help="one calibration copy + one headline frame, no JSON (for counter runs)")
    ap.add_argument("--profile-leg", default="", choices=["", "vol_path", "fwd_adjoint", "combined", "build"],
                    help="one calibration copy + one pass of that leg's workload (no headline, no JSON; for "
                         "counter runs)")
    ap.add_argument("--only", default="",
                    help="comma-separated legs to run besides the headline (vol_path, fwd_adjoint, "
                         "lbvh_build, batched_build, primary_rays, combined); default all")
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
    if args.profile_leg:
        profile_leg(args.profile_leg, ctx, N, api, torch, dev, sync)
        return
    want = set(x for x in args.only.split(",") if x) or {"vol_path", "fwd_adjoint", "lbvh_build",
                                                           "batched_build", "primary_rays", "combined"}
    if args.fast:
        want = set()

    # ------------------------------------------------------ headline workload
    W, H, SPP, DEPTH = 1920, 1080, 16, 8
    log("generating the 2M-triangle displaced octasphere ...")
    v, ind = api.gen_octasphere(9, 0.1, seed=7, base_freq=2.0)
    ntris = ind.size // 3
    dv, di = torch.from_numpy(v).to(dev), torch.from_numpy(ind).to(dev)
    sync()
    bvh = api.LBVH.build(ctx, dv, di)
    cam2 = cam_of(N, (0.0, 0.0, -3.0), (0.0, 0.0, 0.0))

    shape = {}
    if args.shape:
        vals = [int(x) for x in args.shape.split(",")]
        b, bpc, sm, rf = vals[:4]
        rd = vals[4] if len(vals) > 4 else 0
        shape = {"trace_block": b, "blocks_per_cu": bpc, "stack_mode": sm, "ray_fetch": rf, "ring_depth": rd}
    info = bvh.pathtrace(cam2, W, H, SPP, DEPTH, **shape)  # also warms the queues
    segments = info["segments"]
    paths = info["paths"]
    log("frame: %d paths, %d ray segments" % (paths, segments))

    if args.profile_pass:
        # for hardware-counter runs: a plain device copy of known size (the
        # calibration of FETCH_SIZE / WRITE_SIZE), then one headline frame
        calibration_copy(torch, dev, sync)
        bvh.pathtrace(cam2, W, H, SPP, DEPTH, **shape)
        sync()
        log("profile pass done: one frame of %d ray segments" % segments)
        return

    # launch shape picked by the library's autotuner on this frame (untimed)
    tuning = None
    if shape:
        tuning = {"block_size": shape["trace_block"], "blocks_per_cu": shape["blocks_per_cu"],
                  "stack_mode": shape["stack_mode"], "ray_fetch": shape["ray_fetch"],
                  "ring_depth": shape["ring_depth"], "fixed": True}
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
            cfg.ring_depth = shape["ring_depth"]
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
    if args.graph:  # a replayed graph has no per-launch timers: timed eager frames after it
        ctx.reset_profiling()
        for _ in range(args.steps):
            frame(use_graph=False)
        sync()
    trace_ms, trace_launches = ctx.kernel_time("pathtrace_bounce")
    total_ms, _ = ctx.kernel_time("pathtrace_total")
    first_ms, _ = ctx.kernel_time("pathtrace_first_bounce")
    later_ms, _ = ctx.kernel_time("pathtrace_later_bounces")
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
    traffic = measured_traffic()
    head_roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                 "kernel": {1: "k_pt_bounce (closest-hit traversal + diffuse bounce, one launch per bounce)",
                            2: "k_pt_bounce (first bounce) then k_pt_trace + k_pt_shade per bounce "
                               "(closest-hit traversal with per-wave ray refill, then the diffuse bounce); "
                               "a launch = one bounce",
                            3: "k_pt_bounce (first bounce) then k_pt_paths (every later bounce in one launch: "
                               "each lane follows its path to the end, refilling from the queue); a launch = "
                               "one of these two",
                            4: "k_pt_bounce (first bounce), k_pt_trace + k_pt_shade (second), then k_pt_paths "
                               "(every later bounce in one launch, each lane following its path to the end); a "
                               "launch = one of these three"}[(tuning or {}).get("ray_fetch", 1) or 1],
                 "algorithmic_bytes_per_ray": round(bytes_per_ray, 2),
                 "node_visits_per_ray": round(vn, 3), "tri_tests_per_ray": round(vt, 3),
                 "avg_launch_ms": round(avg_launch_ms, 4), "launches_per_frame": launches_per_frame,
                 "trace_kernel_share": round(trace_ms / max(total_ms, 1e-9), 4),
                 "launch_shape": tuning, "graph": bool(args.graph)}
    if traffic:
        # the nodes and triangles stay in the 256 MB MALL and the L2s: HBM
        # itself moves far fewer bytes than the traversal reads, and the
        # kernels are limited by VALU issue and lane divergence, not bandwidth
        meas = traffic / (avg_launch_ms * 1e-3) / 1e9
        head_roof.update({"hbm_measured_GBps": round(meas, 2), "hbm_measured_frac": round(meas / HBM_PEAK_GBS, 5),
                          "limiter": "VALU issue and lane divergence (per-lane traversal); the algorithmic "
                                     "bytes are served mostly from L2 and MALL"})
    incoherent = segments - paths  # every segment after a path's camera ray
    secondary = {
        "first_bounce_ms": round(first_ms / args.steps, 4),
        "incoherent_bounces_ms": round(later_ms / args.steps, 4),
        "incoherent_segments": incoherent,
        "incoherent_grays_per_s": round(incoherent / (later_ms / args.steps * 1e-3) / 1e9, 4),
        "camera_segments": paths,
        "camera_grays_per_s": round(paths / (first_ms / args.steps * 1e-3) / 1e9, 4),
    }
    legs = {}
    cpu_jobs = []  # CPU baselines run after all GPU work

    # ------------------------------------------------------------ LBVH build
    log("headline: %.1f Mrays/s" % value)
    if "lbvh_build" in want:
        log("leg: lbvh_build")
        legs["lbvh_build"] = bench_build(ctx, api, torch, dev, sync, v, ind, dv, di, cpu_jobs)
    if "batched_build" in want:
        log("leg: batched_build")
        legs["batched_build"] = bench_batched(ctx, api, torch, dev, sync, cpu_jobs)
    if "primary_rays" in want:
        log("leg: primary_rays")
        legs["primary_rays"] = bench_primary(ctx, N, api, torch, dev, sync, cpu_jobs, secondary)
    if "vol_path" in want:
        log("leg: vol_path")
        legs["vol_path"] = bench_vol(ctx, N, api, torch, dev, sync, cpu_jobs)
    if "fwd_adjoint" in want:
        log("leg: fwd_adjoint")
        legs["fwd_adjoint"] = bench_fwd_adjoint(ctx, sync, cpu_jobs)
    if "combined" in want:
        log("leg: combined")
        legs["combined"] = bench_combined(ctx, N, api, torch, np, dev, sync, cam2, W, H, cpu_jobs=cpu_jobs)

    # ------------------------------------------------------ CPU baselines
    cpu = None
    if not args.no_cpu:
        cores = host_cores()
        cpu = cpu_baseline(v, ind, W, H, SPP, DEPTH, args.cpu_seconds, cores)
        if cores > 1:
            c1 = cpu_baseline(v, ind, W, H, SPP, DEPTH, args.cpu_seconds / 2, cores - 1)
            cpu["value_nproc_minus_1"] = c1["value"]
            cpu["cores_nproc_minus_1"] = cores - 1
        for job in cpu_jobs:
            log("cpu baseline: %s" % job.__qualname__.split(".")[0])
            job(cores)

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
                   "triangles": ntris, "rays_per_step": segments, "paths_per_step": paths,
                   "max_depth": DEPTH},
        "roofline": head_roof,
        "cpu_baseline": cpu,
        "secondary": secondary,
        "legs": legs,
    }
    print(json.dumps(out))


def calibration_copy(torch, dev, sync):
    """A plain device copy of 1 GiB: what FETCH_SIZE and WRITE_SIZE read for
    it calibrates their counts of a kernel's HBM bytes."""
    src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    sync()
    log("calibration copy: %d bytes read, %d bytes written" % (src.numel() * 4, dst.numel() * 4))


def profile_leg(name, ctx, N, api, torch, dev, sync):
    """One pass of a leg's workload after the calibration copy, for counter
    runs: vol_path = one 256^3 frame (k_svol_regen); fwd_adjoint = one plain
    forward plus one differentiable iteration at the fBm state (k_svol_regen,
    k_diff_regen<false>, k_diff_regen<true>); combined = one frame of the
    combined scene, its differentiable forward and its adjoint (k_volpath_regen,
    k_vpdiff_regen); build = six builds of the 2M-triangle octasphere, then
    three batched builds of the 256 meshes."""
    calibration_copy(torch, dev, sync)
    if name == "vol_path":
        w = vol_workload(N, api, torch, dev, ctx)
        w["vol"].render(w["scene"], w["W"], w["H"], w["spp"], 32, out=w["out"])
        sync()
        w["vol"].close()
    elif name == "fwd_adjoint":
        from fledge_amd import diff
        diff.bench_adjoint(ctx, None, sync, state="fbm", profile=True)
    elif name == "combined":
        bench_combined(ctx, N, api, torch, __import__("numpy"), dev, sync, None, 1920, 1080, profile=True)
    elif name == "build":
        v, ind = api.gen_octasphere(9, 0.1, seed=7, base_freq=2.0)
        dv, di = torch.from_numpy(v).to(dev), torch.from_numpy(ind).to(dev)
        for _ in range(6):
            api.LBVH.build(ctx, dv, di, check_indices=False).close()
        sync()
        log("six builds of the 2,097,152-triangle octasphere done")
        meshes = []
        for k, n in enumerate(api.gen_batch_sizes(256)):
            vv, ii = api.gen_soup(int(n), initstate=42, initseq=1000 + k)
            meshes.append((torch.from_numpy(vv).to(dev), torch.from_numpy(ii).to(dev)))
        for _ in range(3):
            for h in api.LBVH.build_batched(ctx, meshes, check_indices=False):
                h.close()
        sync()
        log("three batched builds of the 256 meshes done")
    log("profile pass done: %s" % name)


# ---------------------------------------------------------------- legs ---
def bench_build(ctx, api, torch, dev, sync, v, ind, dv, di, cpu_jobs):
    """LBVH build of the 2M-triangle octasphere and the 65,536-triangle soup:
    wall time of the build call (output pooled per context after the first
    build) and the build kernels' HIP-event time."""
    names = ["lbvh_bound", "lbvh_morton", "lbvh_sort", "lbvh_emit", "lbvh_refit", "lbvh_soup"]
    res = {}
    v0, i0 = api.gen_soup(65536)
    dv0, di0 = torch.from_numpy(v0).to(dev), torch.from_numpy(i0).to(dev)
    for tag, (vv, ii, n) in {"octasphere_2M": (dv, di, ind.size // 3),
                             "soup_65536": (dv0, di0, i0.size // 3)}.items():
        def one():
            api.LBVH.build(ctx, vv, ii, check_indices=False).close()
        one()
        wall = timed(one, 10, 2, sync)
        ctx.profiling(True)
        ctx.reset_profiling()
        for _ in range(5):
            one()
        sync()
        kms = kernel_ms(ctx, names, 5)
        parts = {n: round(ctx.kernel_time(n)[0] / 5, 4) for n in names}
        ctx.profiling(False)
        res[tag] = {"ms": round(wall * 1e3, 4), "kernel_ms": round(kms, 4), "mtris_per_s": round(n / wall / 1e6, 2),
                    "stages_ms": parts, "triangles": n,
                    "roofline": roofline(BUILD_BYTES_PER_TRI * n / (kms * 1e-3) / 1e9,
                                         "the build's kernels (bound, Morton, sort, emit, refit + soup)",
                                         BUILD_BYTES_PER_TRI, "triangle")}
    out = {"metric": "LBVH build ms (2,097,152-triangle displaced octasphere)", "value": res["octasphere_2M"]["ms"],
           "unit": "ms", "higher_is_better": False, "roofline": res["octasphere_2M"]["roofline"],
           "builds": res, "cpu_baseline": None}

    def cpu_job(cores):
        out["cpu_baseline"] = cpu_build_baseline(v, ind)
    cpu_jobs.append(cpu_job)
    return out


def bench_batched(ctx, api, torch, dev, sync, cpu_jobs):
    sizes = api.gen_batch_sizes(256)
    meshes = []
    for k, n in enumerate(sizes):
        vv, ii = api.gen_soup(int(n), initstate=42, initseq=1000 + k)
        meshes.append((torch.from_numpy(vv).to(dev), torch.from_numpy(ii).to(dev)))
    sync()

    def batch():
        for h in api.LBVH.build_batched(ctx, meshes, check_indices=False):
            h.close()
    batch()
    tb = timed(batch, 5, 1, sync)
    names = ["lbvh_bound", "lbvh_morton", "lbvh_sort", "lbvh_emit", "lbvh_refit", "lbvh_soup"]
    ctx.profiling(True)
    ctx.reset_profiling()
    for _ in range(3):
        batch()
    sync()
    kms = kernel_ms(ctx, names, 3)
    parts = {n: round(ctx.kernel_time(n)[0] / 3, 4) for n in names}
    ctx.profiling(False)
    tot = int(sizes.sum())
    out = {"metric": "batched LBVH build of 256 meshes (log-uniform 1K..256K triangles) in one call, ms",
           "value": round(tb * 1e3, 3), "unit": "ms", "higher_is_better": False, "triangles": tot,
           "mtris_per_s": round(tot / tb / 1e6, 2), "kernel_ms": round(kms, 4), "stages_ms": parts,
           "roofline": roofline(BUILD_BYTES_PER_TRI * tot / (kms * 1e-3) / 1e9, "the batch's build kernels",
                                BUILD_BYTES_PER_TRI, "triangle"),
           "cpu_baseline": None}
    host = [(m[0].cpu().numpy(), m[1].cpu().numpy()) for m in meshes[:32]]

    def cpu_job(cores):
        import oracle_lib as orc
        t0 = time.perf_counter()
        n = 0
        for vv, ii in host:
            orc.Lbvh(vv, ii)
            n += ii.size // 3
        dtc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(dtc * 1e3 * tot / n, 1), "unit": "ms", "cores": 1, "kind": "port",
                               "sample": "the first 32 of the 256 meshes (%d triangles) built one after another "
                                         "by the oracle's serial build, %.2f s, scaled to the batch's %d "
                                         "triangles" % (n, dtc, tot)}
    cpu_jobs.append(cpu_job)
    # the scene accel over the same 256 meshes, each moved to a cell of a
    # 16 x 16 grid (pitch 1.25), against one LBVH over all their triangles:
    # 1024^2 camera rays from above
    from fledge_amd import _native as N
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
    out["scene_accel_256_meshes_mrays_per_s"] = round(1024 * 1024 / ts_ / 1e6, 2)
    out["scene_accel_hit_fraction"] = round(float((hits_s[:, 0] >= 0).float().mean()), 4)
    merged = api.LBVH.build(ctx, torch.cat(allv), torch.cat(alli))
    hits_m = torch.empty((1024 * 1024, 2), dtype=torch.int32, device=dev)
    tm_ = timed(lambda: merged.trace(rays_s, hits_m), 10, 2, sync)
    out["merged_single_lbvh_mrays_per_s"] = round(1024 * 1024 / tm_ / 1e6, 2)
    sacc.close()
    merged.close()
    for h in blas:
        h.close()
    return out


def bench_primary(ctx, N, api, torch, dev, sync, cpu_jobs, secondary):
    v0, i0 = api.gen_soup(65536)
    b0 = api.LBVH.build(ctx, torch.from_numpy(v0).to(dev), torch.from_numpy(i0).to(dev))
    cam0 = cam_of(N, (0.5, 0.5, -1.5), (0.5, 0.5, 0.0))
    rays0 = api.camera_rays(ctx, cam0, 1024, 1024)
    hits0 = torch.empty((1024 * 1024, 2), dtype=torch.int32, device=dev)
    t = timed(lambda: b0.trace(rays0, hits0), 20, 5, sync)
    ctx.profiling(True)
    ctx.reset_profiling()
    for _ in range(10):
        b0.trace(rays0, hits0)
    sync()
    kms = kernel_ms(ctx, ["trace_closest"], 10)
    ctx.profiling(False)
    c = b0.count(rays0).cpu().numpy()
    R = 1024 * 1024
    vn, vt = c[0] / R, c[1] / R
    bpr = 32 + 8 + 64 * vn + 36 * vt
    out = {"metric": "closest-hit Mrays/s, 1024^2 jittered primary rays on the 65,536-triangle soup",
           "value": round(R / t / 1e6, 2), "unit": "Mrays/s", "higher_is_better": True,
           "kernel_ms": round(kms, 4),
           "roofline": roofline(bpr * R / (kms * 1e-3) / 1e9, "k_trace", round(bpr, 2), "ray",
                                node_visits_per_ray=round(vn, 3), tri_tests_per_ray=round(vt, 3)),
           "cpu_baseline": None}
    rays_h = rays0.cpu().numpy()

    def cpu_job(cores):
        import oracle_lib as orc
        orc.NTHREADS = cores
        o = orc.Lbvh(v0, i0)
        sub = rays_h[::8]
        t0 = time.perf_counter()
        o.trace(sub)
        dtc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(len(sub) / dtc / 1e6, 3), "unit": "Mrays/s", "cores": cores,
                               "kind": "port",
                               "sample": "every 8th of the 1024^2 rays (%d), the oracle's per-ray DFS, %.2f s"
                                         % (len(sub), dtc)}
    cpu_jobs.append(cpu_job)
    b0.close()
    return out


def vol_workload(N, api, torch, dev, ctx):
    """The 256^3 fBm volume (seed 1, sigma_s 6.4, sigma_a 1.6, g 0.3), its
    camera and lights, 512x512 at 64 spp."""
    g1 = api.gen_fbm_grid(256, seed=1)
    vol = api.Volume(ctx, torch.from_numpy(g1).to(dev), [0, 0, 0, 1, 1, 1], 6.4, 1.6, 0.3)
    cam0 = cam_of(N, (0.5, 0.5, -1.5), (0.5, 0.5, 0.0))
    out = torch.empty((512, 512, 3), dtype=torch.float32, device=dev)
    return {"grid": g1, "vol": vol, "scene": api.vol_scene(cam0), "W": 512, "H": 512, "spp": 64, "out": out}


def bench_vol(ctx, N, api, torch, dev, sync, cpu_jobs):
    """256^3 fBm grid, 512x512, 64 spp, depth 32: path samples/s; roofline of
    the path kernel over 32 B per trilinear lookup + 12 B per pixel."""
    w = vol_workload(N, api, torch, dev, ctx)
    g1, vol, sc1, Wv, Hv, S1, out1 = w["grid"], w["vol"], w["scene"], w["W"], w["H"], w["spp"], w["out"]
    _, st1 = vol.render(sc1, Wv, Hv, S1, 32, out=out1, stats=True)
    t1 = timed(lambda: vol.render(sc1, Wv, Hv, S1, 32, out=out1), 5, 2, sync)
    t1b = timed(lambda: vol.render(sc1, Wv, Hv, S1, 32, out=out1, schedule=1), 3, 1, sync)
    t1c = timed(lambda: vol.render(sc1, Wv, Hv, S1, 32, out=out1, schedule=3, use_graph=True), 3, 1, sync)
    ctx.profiling(True)
    ctx.reset_profiling()
    for _ in range(5):
        vol.render(sc1, Wv, Hv, S1, 32, out=out1)
    sync()
    kms = kernel_ms(ctx, ["volpath_paths"], 5)
    ctx.profiling(False)
    P = Wv * Hv * S1
    L = st1["lookups"] / P
    lb = 32 * st1["lookups"] + 12 * Wv * Hv
    out = {"metric": "vol path samples/s (256^3 fBm grid, 512x512, 64 spp, depth 32)",
           "value": round(P / t1, 1), "unit": "path samples/s", "higher_is_better": True,
           "ms_per_frame": round(t1 * 1e3, 3), "ms_per_frame_lane_per_path": round(t1b * 1e3, 3),
           "ms_per_frame_wavefront_graph": round(t1c * 1e3, 3),
           "kernel_ms": round(kms, 4), "lookups_per_path": round(L, 4),
           "roofline": roofline(lb / (kms * 1e-3) / 1e9, "k_svol_regen (persistent lanes with path regeneration)",
                                round(32 * L + 12 / S1, 3), "path"),
           "cpu_baseline": None}

    def cpu_job(cores):
        import oracle_lib as orc
        import numpy as np
        lib = orc.lib()
        rgb = np.zeros((Hv, Wv, 3), np.float32)
        lk = np.zeros(1, np.uint64)
        pa = np.zeros(1, np.uint64)
        cam10 = np.array([0.5, 0.5, -1.5, 0.5, 0.5, 0.0, 0.0, 1.0, 0.0, 45.0], np.float32)
        stride = 1  # the whole frame
        t0 = time.perf_counter()
        lib.orc_volpath_render_strided(
            orc._p(g1), 256, 256, 256, orc._p(np.array([0, 0, 0, 1, 1, 1], np.float32)),
            orc._p(np.array([6.4, 1.6, 0.3], np.float32)), orc._p(orc.lights11([1, 2, 2, -2, 50.0, 1, 0.5])),
            orc._p(cam10), Wv, Hv, S1, 32, stride, orc._p(rgb), cores, orc._p(lk), orc._p(pa))
        dtc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(float(pa[0]) / dtc, 1), "unit": "path samples/s", "cores": cores,
                               "kind": "port",
                               "sample": "the whole frame (%d paths: 512 x 512 pixels, 64 spp each), the oracle's "
                                         "SVolIntegrator restatement, %.2f s" % (int(pa[0]), dtc)}
    cpu_jobs.append(cpu_job)
    vol.close()
    return out


def bench_fwd_adjoint(ctx, sync, cpu_jobs):
    from fledge_amd import diff
    states = {}
    for st in ("uniform", "fbm"):
        states[st] = diff.bench_adjoint(ctx, timed, sync, state=st)
    out = {"metric": "differentiable forward + adjoint ms/iter vs the plain forward (128^3, 256x256, 16 spp)",
           "unit": "ms/iter", "higher_is_better": False, "states": {}}
    for st, r in states.items():
        P = r["paths"]
        lf, la = r["lookups_forward"] / P, r["lookups_adjoint"] / P
        kms = r["kernel_forward_ms"] + r["kernel_adjoint_ms"]
        bytes_iter = P * (32 * lf + 64 * la)
        rec = {k: (round(x, 4) if isinstance(x, float) else x) for k, x in r.items()
               if not isinstance(x, dict)}
        rec.update({k: {kk: round(vv, 4) for kk, vv in x.items()} for k, x in r.items() if isinstance(x, dict)})
        rec["lookups_per_path_forward"] = round(lf, 4)
        rec["lookups_per_path_adjoint"] = round(la, 4)
        rec["roofline"] = roofline(bytes_iter / (kms * 1e-3) / 1e9,
                                   "k_diff_regen<false> + k_diff_regen<true> (forward, then the replay)",
                                   round(32 * lf + 64 * la, 3), "path",
                                   bytes_per_iteration=int(bytes_iter))
        out["states"][st] = rec
    out["value"] = out["states"]["fbm"]["fwd_adjoint_ms_per_iter"]
    out["ratio_vs_plain_forward"] = out["states"]["fbm"]["fwd_adjoint_vs_plain"]
    out["ratio_vs_plain_forward_uniform_state"] = out["states"]["uniform"]["fwd_adjoint_vs_plain"]
    out["roofline"] = out["states"]["fbm"]["roofline"]
    out["cpu_baseline"] = None

    def cpu_job(cores):
        out["cpu_baseline"] = cpu_adjoint_baseline(cores)
    cpu_jobs.append(cpu_job)
    return out


def bench_combined(ctx, N, api, torch, np, dev, sync, cam2, W, H, profile=False, cpu_jobs=None):
    """VolPathIntegrator: the 8,388,608-triangle displaced octahedral sphere
    (amplitude 0.05) as the interface around a 512^3 sparse-occupancy grid
    whose 8^3 bricks are active with probability 0.15 and hold fBm, stored as
    a brick pool (flg_volume_create_sparse); 1920x1080, 64 spp, depth 16;
    then the differentiable forward with the walk tape and the adjoint over
    the active bricks.  The same frame on the dense copy of the grid is timed
    beside it.  With `profile`: one pass of each sparse kernel, no timing."""
    from fledge_amd import diff
    if cam2 is None:
        cam2 = cam_of(N, (0.0, 0.0, -3.0), (0.0, 0.0, 0.0))
    SPPc, Dc = 64, 16
    vc, ic = api.gen_octasphere(10, 0.05, seed=7, base_freq=2.0)
    gc = api.gen_fbm_grid(512, seed=3)
    bricks = np.random.default_rng(4).random((64, 64, 64)) < 0.15
    gc *= np.repeat(np.repeat(np.repeat(bricks, 8, 0), 8, 1), 8, 2)
    bc = api.LBVH.build(ctx, torch.from_numpy(vc).to(dev), torch.from_numpy(ic).to(dev))
    gct = torch.from_numpy(gc).to(dev)
    origins, values = api.leaves_of(gct)
    gmax = float(gc.max())
    host_grid = gc if (cpu_jobs is not None and not profile) else None
    host_mesh = (vc, ic) if host_grid is not None else None
    del gc
    volc = api.Volume.sparse(ctx, origins, values, (512, 512, 512), [0, 0, 0, 1, 1, 1], 6.4, 1.6, 0.3)
    scc = api.vol_scene(cam2)
    outc = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
    _, stc = volc.render_mesh(bc, scc, W, H, SPPc, Dc, out=outc, stats=True)
    dvc = diff.DiffVolume.sparse(ctx, origins, values, (512, 512, 512), g=0.3)
    layout = diff.BrickLayout.of_volume(dvc.vol)
    tape = diff.new_tape(ctx, W, H, SPPc, Dc)
    lpc = torch.empty((W * H * SPPc, 3), dtype=torch.float32, device=dev)
    dlc = torch.full((H, W, 3), 1.0 / (W * H * 3), dtype=torch.float32, device=dev)
    gb = layout.zeros()
    gac = torch.zeros(1, dtype=torch.float32, device=dev)
    mu_c = 8.0 * gmax
    if profile:
        volc.render_mesh(bc, scc, W, H, SPPc, Dc, out=outc)
        diff.mesh_forward(dvc, bc, scc, W, H, SPPc, Dc, 8.0, 0.8, mu_c, out=outc, lpath=lpc, tape=tape)
        diff.mesh_adjoint(dvc, bc, scc, W, H, SPPc, Dc, 8.0, 0.8, mu_c, lpc, dlc, grad=gb, galpha=gac, tape=tape,
                          bricks=layout)
        sync()
        bc.close()
        return None
    tc = timed(lambda: volc.render_mesh(bc, scc, W, H, SPPc, Dc, out=outc), 2, 1, sync)
    ctx.profiling(True)
    ctx.reset_profiling()
    volc.render_mesh(bc, scc, W, H, SPPc, Dc, out=outc)
    sync()
    kms = kernel_ms(ctx, ["volpath_mesh_paths"], 1)
    ctx.profiling(False)
    tf = timed(lambda: diff.mesh_forward(dvc, bc, scc, W, H, SPPc, Dc, 8.0, 0.8, mu_c, out=outc, lpath=lpc,
                                         tape=tape), 1, 1, sync)
    ta = timed(lambda: diff.mesh_adjoint(dvc, bc, scc, W, H, SPPc, Dc, 8.0, 0.8, mu_c, lpc, dlc, grad=gb,
                                         galpha=gac, tape=tape, bricks=layout), 1, 0, sync)
    del tape, lpc, gb
    # the same frame on the dense grid (512 MiB plus its border)
    vold = api.Volume(ctx, gct, [0, 0, 0, 1, 1, 1], 6.4, 1.6, 0.3)
    outd = torch.empty_like(outc)
    td = timed(lambda: vold.render_mesh(bc, scc, W, H, SPPc, Dc, out=outd), 1, 1, sync)
    volc.render_mesh(bc, scc, W, H, SPPc, Dc, out=outc)  # (outc last held the differentiable forward)
    same = bool(torch.equal(outc.view(torch.int32), outd.view(torch.int32)))
    vold.close()
    del gct
    P = W * H * SPPc
    rays = stc["rays"]
    look = stc["lookups"]
    # algorithmic bytes: per trilinear lookup 32 B of voxels + the 4-byte
    # slot-map entry; per ray traced on the mesh 32 B in + 8 B out, 64 B per
    # node visit and 36 B per triangle test (counted by the same kernel's
    # counting build); 12 B per pixel written
    lb = 36 * look + 40 * rays + 64 * stc["node_visits"] + 36 * stc["tri_tests"] + 12 * W * H
    pool_bytes = origins.shape[0] * 4000 + 65 * 65 * 65 * 32
    out = {"metric": "combined scene: VolPathIntegrator path samples/s (8,388,608-triangle interface mesh around "
                     "a 512^3 sparse grid of 8^3 bricks, 15% active, in a brick pool; 1920x1080, 64 spp, depth 16)",
           "value": round(P / tc, 1), "unit": "path samples/s", "higher_is_better": True,
           "ms_per_frame": round(tc * 1e3, 2), "kernel_ms": round(kms, 3),
           "mrays_per_s": round(rays / tc / 1e6, 1), "rays_per_frame": rays,
           "lookups_per_path": round(look / P, 3),
           "dense_grid_ms_per_frame": round(td * 1e3, 2), "dense_and_sparse_images_bit_equal": same,
           "active_bricks": int(origins.shape[0]), "grid_bytes_sparse": pool_bytes,
           "grid_bytes_dense": 514 ** 3 * 4,
           "diff_forward_ms": round(tf * 1e3, 2), "adjoint_ms": round(ta * 1e3, 2),
           "fwd_adjoint_over_forward": round((tf + ta) / tc, 3),
           "roofline": roofline(lb / (kms * 1e-3) / 1e9,
                                "k_volpath_regen (persistent lanes over the interface mesh and the brick pool)",
                                round(lb / P, 3), "path",
                                node_visits_per_ray=round(stc["node_visits"] / rays, 3),
                                tri_tests_per_ray=round(stc["tri_tests"] / rays, 3)),
           "cpu_baseline": None}
    bc.close()
    if host_grid is not None:
        def cpu_job(cores):
            import oracle_lib as orc
            vv, ii = host_mesh
            m = orc.Lbvh(vv, ii)
            stride = 5  # 414,720 of 2,073,600 pixels, all 64 samples each
            cam10 = [0.0, 0.0, -3.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 45.0]
            t0 = time.perf_counter()
            _, _, pa = m.volpath_mesh_strided(host_grid, [6.4, 1.6, 0.3], [1, 2, 2, -2, 50.0, 1, 0.5], cam10, W, H,
                                              SPPc, Dc, stride, nthreads=cores)
            dtc = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": round(pa / dtc, 1), "unit": "path samples/s", "cores": cores,
                                   "kind": "port",
                                   "sample": "every %d-th pixel of the frame (%d paths, all 64 spp each) on the "
                                             "dense copy of the grid, the oracle's VolPathIntegrator restatement "
                                             "over its own LBVH, %.2f s" % (stride, pa, dtc)}
        cpu_jobs.append(cpu_job)
    return out


# ------------------------------------------------------- CPU baselines ---
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


def cpu_baseline(v, ind, W, H, SPP, DEPTH, seconds, cores):
    """The oracle's LBVH + per-path DFS traversal (the restated fledge CPU path)
    on a bounded sample of the same frame: whole image rows spread over the
    frame, sized to about `seconds` of work on `cores` host threads."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc

    o = orc.Lbvh(v, ind)
    cam10 = [0.0, 0.0, -3.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 45.0]
    row = W * SPP
    # calibrate on one central row
    t0 = time.perf_counter()
    o.diffuse_render(cam10, W, H, SPP, DEPTH, (H // 2) * row, (H // 2 + 1) * row, record=False, nthreads=cores)
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
            "sample": "%d of %d image rows (evenly spaced), %d paths, %d ray segments, %.1f s on %d threads "
                      "(the host share of one GPU is %d cores)"
                      % (len(rows), H, len(rows) * row, rays, dt, cores, CPU_SHARE)}


def cpu_build_baseline(v, ind):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    t0 = time.perf_counter()
    orc.Lbvh(v, ind)
    dtc = time.perf_counter() - t0
    return {"value": round(dtc * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
            "sample": "one build of the whole 2,097,152-triangle mesh by the oracle's serial restatement "
                      "(stable sort, Karras emission, bottom-up refit), %.2f s" % dtc}


def cpu_adjoint_baseline(cores):
    """The FP64 CPU twin of the differentiable estimator (forward, then the
    analytic adjoint) over the whole 256x256 x 16 spp configuration at the
    non-uniform state (a few seconds of CPU work)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    from fledge_amd import api
    P = orc._p
    L = orc.lib()
    theta = api.gen_fbm_grid(128, seed=3)
    th64 = theta.astype(np.float64)
    W = H = 256
    SPP, DEPTH = 16, 32
    mu = 8.0 * float(theta.max())
    params = np.array([8.0, 0.8, 0.3, mu], np.float32)
    lights = orc.lights11([1, 2, 2, -2, 50.0, 1, 0.5])  # grey: one channel covers all three
    cam10 = np.array([0.5, 0.5, -1.5, 0.5, 0.5, 0.0, 0.0, 1.0, 0.0, 45.0], np.float32)
    bounds = np.array([0, 0, 0, 1, 1, 1], np.float32)
    stride = 1
    lp = np.zeros(W * H * SPP, np.float64)
    img = np.zeros(W * H, np.float64)
    dl = np.full(W * H, 1.0 / (W * H), np.float64)
    grad = np.zeros(128 ** 3, np.float64)
    ga = np.zeros(1, np.float64)
    lk = np.zeros(1, np.uint64)
    t0 = time.perf_counter()
    L.orc_diff_render_strided(P(th64), P(theta), 128, 128, 128, P(bounds), P(params), P(lights), P(cam10), W, H,
                              SPP, DEPTH, stride, P(lp), P(img), None, None, None, cores, P(lk), 0)
    t1 = time.perf_counter()
    L.orc_diff_render_strided(P(th64), P(theta), 128, 128, 128, P(bounds), P(params), P(lights), P(cam10), W, H,
                              SPP, DEPTH, stride, P(lp), None, P(dl), P(grad), P(ga), cores, None, 0)
    t2 = time.perf_counter()
    npx = (W * H + stride - 1) // stride
    scale = W * H / npx
    return {"value": round((t2 - t0) * scale * 1e3, 1), "unit": "ms/iter", "cores": cores, "kind": "port",
            "forward_ms": round((t1 - t0) * scale * 1e3, 1), "adjoint_ms": round((t2 - t1) * scale * 1e3, 1),
            "sample": "%d of %d pixels (16 spp each) of the fBm state through the FP64 CPU twin "
                      "(forward %.2f s, adjoint %.2f s)" % (npx, W * H, t1 - t0, t2 - t1)}


if __name__ == "__main__":
    main()
