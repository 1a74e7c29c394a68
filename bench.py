"""Benchmark of the progressive path-tracing pass on one MI355X.

Workload (one step = one full render): displaced icosphere, subdivision level 8
(1,310,720 triangles, 3-octave value-noise displacement of amplitude 0.05, per-face
Beckmann alpha U[0.05, 0.5] with 30% Lambert faces), synthetic 1024x512 HDR environment
(log-normal sigma 1 + 1e4 sun lobe) importance-sampled through FP64-built alias tables,
1024x1024 pixels, 64 samples per pixel, depth 8, NEE + MIS, FP32 tracing, FP64
accumulation.  Synthetic data, seeded.

Prints one JSON line: Mrays/s (primary + extension + shadow rays actually traced, divided by
the wall time of the timed steps), with the roofline of the dominant kernel measured live
from HIP events on the renderer's stream, the CPU oracle timed on a bounded sample of the
same workload on the host cores, and (mse_ratio) the other half of the headline metric: the
equal-time mean squared error of the plain and the guided renderer on the hard-indirect room
against the FP64 oracle's 65,536-sample reference, as their ratio; (batch) the 48-render
batch's aggregate Mrays/s and per-render latency.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-pathed_amd"))

HBM_PEAK_GBS = 8000.0
WIDTH = HEIGHT = 1024
SPP = 64
DEPTH = 8


def log(msg):
    print(msg, file=sys.stderr, flush=True)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "wavefront_icosphere1M_1024sq_64spp_hbm_traffic.md")


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary of this workload (lines 'traffic_bytes_per_dispatch: <kernel> <bytes>'), or None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                parts = line.strip().strip("`").split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_dispatch:" and parts[1] == kernel:
                    return int(parts[2])
    except OSError:
        pass
    return None


def host_cpu():
    """Threads the CPU baseline uses and what the host reports: the cores this process may run
    on (affinity), capped at the per-GPU CPU share the environment sets (OMP_NUM_THREADS), the
    machine's logical CPU count and its model name (lscpu)."""
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = nproc
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = max(1, min(aff, share) if share > 0 else aff)
    model = None
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in txt.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    return threads, nproc, aff, model


def cpu_baseline(scene_arrays, budget_s):
    """The oracle (scalar C++ restatement of the reference's shading and intersection, with
    a median-split BVH) on the host cores: a band of rows of the same image at as many
    samples per pixel as fit in about budget_s / 3 seconds, timed three times (median)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    threads, nproc, aff, model = host_cpu()
    t0 = time.time()
    o = O.Scene(scene_arrays, WIDTH, HEIGHT, bvh=True)
    log(f"cpu baseline: oracle scene + BVH built in {time.time() - t0:.1f}s; {threads} threads "
        f"(nproc {nproc}, affinity {aff}, {model})")
    kw = dict(max_depth=DEPTH, nee=1, mode=1, tri_test=1, ref_env=0, precision=32, threads=threads)
    # probe 16 rows of the image centre at 1 spp (also the warm-up), then size rows x spp so
    # that one timed run takes about a third of the budget
    t0 = time.time()
    o.render(row_begin=504, row_end=520, spp=1, **kw)
    per_row = max(time.time() - t0, 1e-3) / 16
    run_s = budget_s / 3
    rows = int(min(HEIGHT, max(16, run_s / per_row)))
    spp = int(max(1, min(SPP, run_s / (per_row * rows))))
    rb = max(0, HEIGHT // 2 - rows // 2)
    rates, times = [], []
    for k in range(3):
        t0 = time.time()
        _, _, rays = o.render(row_begin=rb, row_end=rb + rows, spp=spp, sample_start=k * spp, **kw)
        dt = time.time() - t0
        rates.append(rays / dt / 1e6)
        times.append(round(dt, 2))
        log(f"cpu baseline run {k}: {rays} rays in {dt:.2f}s")
    med = sorted(rates)[1]
    return {
        "value": med,
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{spp} spp over rows [{rb}, {rb + rows}) of the {WIDTH}x{HEIGHT} image ({rows * WIDTH * spp} paths per run, "
                  f"{rays} rays), same scene, depth {DEPTH}, NEE+MIS, FP32, median-split BVH; median of 3 runs "
                  f"({times} s) on {threads} threads",
        "runs_mrays_s": [round(x, 2) for x in rates],
        "per_core_mrays_s": round(med / threads, 3),
        "host": {"nproc": nproc, "affinity_cpus": aff, "cpu_share": threads, "model": model},
    }


ROOM_W, ROOM_H, ROOM_DEPTH = 2048, 1024, 16
ROOM_REFERENCE = os.path.join(ROOT, "tests", "golden", "room_reference_1024px.npz")


def room_equal_time(P, budget_s, spp_per_pass, repeats=1):
    """Hard indirect lighting, equal wall time: the closed room of 4M triangles lit through a
    0.3-wide slot (2048x1024, depth 16, NEE + consistent MIS) rendered for `budget_s` seconds
    by the plain and by the guided renderer (network trained from scratch inside the budget).
    Mean squared error at a fixed seeded subset of 1024 pixels against the FP64 oracle's
    65,536-sample estimate there (the committed reference fixture).  With repeats > 1 each
    renderer runs that many times with different PCG32 seeds and the MSEs are averaged."""
    import numpy as np

    ref = np.load(ROOM_REFERENCE)
    scene = P.Scene.preset(2, seed=0)
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_room_reference import scene_fingerprint

    fp = scene_fingerprint(scene.export())
    if not np.allclose(fp, ref["fingerprint"], rtol=1e-9):
        raise SystemExit("bench.py: the room scene differs from the one the reference fixture was rendered for")
    pix = ref["pixels"]
    ref_mean = ref["sum"] / ref["spp"]
    ref_var = (ref["sumsq"] / ref["spp"] - ref_mean**2) / ref["spp"]
    res = {}
    for mode in ("plain", "guided"):
        runs = [room_run(P, scene, mode, budget_s, spp_per_pass, seed, pix, ref_mean) for seed in range(repeats)]
        res[mode] = runs[0] if repeats == 1 else {
            "mse": float(np.mean([x["mse"] for x in runs])), "rel_mse": float(np.mean([x["rel_mse"] for x in runs])),
            "spp": [x["spp"] for x in runs], "seconds": [x["seconds"] for x in runs], "mrays_s": [x["mrays_s"] for x in runs],
            "mse_runs": [x["mse"] for x in runs], "guide_steps": [x["guide_steps"] for x in runs]}
    return {
        "metric": "equal-time MSE ratio (plain / guided)", "value": round(res["plain"]["mse"] / res["guided"]["mse"], 3),
        "unit": "x", "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"closed room, {scene.info().n_triangles} tris, 0.3-wide slot, light 1e3, glass sphere r 0.5, "
                               f"2048x1024 env, {ROOM_W}x{ROOM_H}, depth {ROOM_DEPTH}, NEE + consistent MIS, "
                               f"{budget_s:.0f} s per renderer, {repeats} seed(s)",
                   "pixels": int(pix.size), "reference_spp": int(ref["spp"]),
                   "reference_noise_mse": float(ref_var.mean())},
        "plain": res["plain"], "guided": res["guided"],
    }


def room_run(P, scene, mode, budget_s, spp_per_pass, seed, pix, ref_mean):
    """One equal-time render of the room by the plain or the guided renderer."""
    import numpy as np

    r = P.Renderer(ROOM_W, ROOM_H)
    r.set_scene(scene)
    r.set_max_depth(ROOM_DEPTH)
    r.set_nee(True)
    r.set_option(P.OPT_CONSISTENT_MIS, 1)
    r.set_seed(seed)
    if mode == "guided":
        r.set_type(P.RENDERER_GUIDED)
    # load the kernels, then start again from nothing: fresh accumulation, fresh network
    r.render(1)
    if mode == "guided":
        r.set_option(P.OPT_GUIDE_RESET, 1)
    r.reset()
    r.reset_stats()
    r.synchronize()
    t0 = time.perf_counter()
    curve = []
    while True:
        r.render(spp_per_pass)
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
        if len(curve) < 64 and r.spp() & (r.spp() - 1) == 0:
            img = r.radiance_sum().reshape(-1, 3)[pix] / r.spp()
            curve.append((round(el, 3), r.spp(), float(((img - ref_mean) ** 2).mean())))
            t0 += time.perf_counter() - t0 - el  # the error readback is not render time
    el = time.perf_counter() - t0
    spp = r.spp()
    img = r.radiance_sum().reshape(-1, 3)[pix] / spp
    st = r.stats()
    mse = float(((img - ref_mean) ** 2).mean())
    out = {"mse": mse, "rel_mse": float((((img - ref_mean) ** 2) / (ref_mean**2 + 1e-2)).mean()),
           "spp": spp, "seconds": round(el, 3), "mrays_s": round(st.rays_total / el / 1e6, 2),
           "guide_steps": int(st.guide_steps), "curve": curve}
    log(f"room {mode} seed {seed}: {spp} spp in {el:.1f}s, MSE {mse:.5g}, {out['mrays_s']} Mrays/s")
    r.close()
    return out


def batch_renders(P, steps, probe_divisor, seed=2024, schedule="wavefront"):
    """48 independent renders in one captured graph: resolution uniform in {128, 256, 512,
    1024}^2, spp uniform in {4, 16, 64, 256}, each a displaced geodesic sphere with a
    triangle count log-uniform in [1e4, 2e6] and per-face materials and environment as the
    icosphere workload, depth 8.  Launch shape tuned by the autotuner on probe runs at
    spp / probe_divisor; aggregate Mrays/s over `steps` timed runs, per-render latency
    (device completion stamp - start of the run) p50 / p99 of the last run."""
    import numpy as np

    rng = np.random.default_rng(seed)
    b = P.Batch()
    b.set_option(P.OPT_BATCH_WAVEFRONT, int(schedule == "wavefront"))
    specs = []
    t0 = time.time()
    for i in range(48):
        res = int(rng.choice([128, 256, 512, 1024]))
        spp = int(rng.choice([4, 16, 64, 256]))
        tris = int(round(math.exp(rng.uniform(math.log(1e4), math.log(2e6)))))
        sc = P.Scene.preset(4, seed=1000 + i, param=tris)
        n = sc.info().n_triangles
        b.add(sc, res, res, spp, DEPTH)
        del sc
        specs.append((res, spp, n))
    log(f"batch: 48 scenes generated and uploaded in {time.time() - t0:.1f}s "
        f"({sum(x[2] for x in specs)} triangles, {sum(x[0] * x[0] * x[1] for x in specs)} paths per run)")
    t0 = time.time()
    tune = b.autotune(probe_divisor)
    log(f"batch: autotune {tune} in {time.time() - t0:.1f}s")
    b.run()  # capture + first launch
    log("batch: warm run done")
    wall = 0.0
    rays = 0
    for k in range(steps):
        b.run()
        st = b.stats()
        wall += st["wall_ms"]
        rays += st["rays"]
        log(f"batch: run {k} {st['wall_ms']:.1f} ms, {st['rays']} rays")
    lat = np.array([b.render_stats(i)["latency_ms"] for i in range(48)])
    out = {
        "metric": "Mrays/s (48-render batch)", "value": round(rays / (wall * 1e3), 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": steps, "ms_per_step": round(wall / steps, 2), "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "48 renders, res in {128,256,512,1024}^2, spp in {4,16,64,256}, geodesic spheres with "
                               "1e4-2e6 triangles (log-uniform), depth 8, NEE+MIS, one hipGraph launch per run",
                   "schedule": schedule,
                   "paths_per_run": int(sum(x[0] * x[0] * x[1] for x in specs)), "triangles": int(sum(x[2] for x in specs)),
                   "chunks": st["chunks"], "graph_nodes": st["graph_nodes"], "seed": seed},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2), "p99": round(float(np.percentile(lat, 99)), 2),
                       "max": round(float(lat.max()), 2)},
        "autotune": tune,
        "stack_overflows": st["stack_overflows"],
    }
    b.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["icosphere", "room", "batch"], default="icosphere",
                    help="icosphere: Mrays/s of the 1M-triangle render (default); room: equal-time MSE, guided vs plain; "
                         "batch: 48 renders in one captured graph")
    ap.add_argument("--probe-divisor", type=int, default=16, help="batch: autotuner probes run spp / this")
    ap.add_argument("--batch-schedule", choices=["wavefront", "sequential"], default="wavefront",
                    help="batch: one wavefront over all renders (default) or render after render")
    ap.add_argument("--budget", type=float, default=20.0, help="room: seconds per renderer")
    ap.add_argument("--spp-per-pass", type=int, default=1, help="room: samples per pixel per progressive pass")
    ap.add_argument("--repeats", type=int, default=3, help="room: seeds per renderer (MSE averaged)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU-oracle work for cpu_baseline (0: skip)")
    ap.add_argument("--room-budget", type=float, default=20.0,
                    help="default workload: seconds per renderer of the equal-time guided-vs-plain room measurement "
                         "reported as mse_ratio (0: skip)")
    ap.add_argument("--room-repeats", type=int, default=2, help="default workload: seeds per renderer for mse_ratio")
    ap.add_argument("--batch-steps", type=int, default=2,
                    help="default workload: timed runs of the 48-render batch reported as batch (0: skip)")
    ap.add_argument("--level", type=int, default=8, help="icosphere subdivision level")
    ap.add_argument("--pool", type=int, default=0, help="path pool size (0: library default)")
    ap.add_argument("--trace-blocks", type=int, default=0, help="trace blocks per CU (0: library default)")
    ap.add_argument("--bvh", choices=["default", "lbvh", "ploc"], default="default",
                    help="binary tree under the 4-wide BVH (default: the library's)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (torch) to calibrate HBM byte counters, then exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copied {args.copy_calibration} MiB three times")
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU path is a design, not part of this build", file=sys.stderr)
        sys.exit(2)

    import pathed_capi as P

    if P.device_count() == 0:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.workload == "room":
        print(json.dumps(room_equal_time(P, args.budget, args.spp_per_pass, args.repeats)), flush=True)
        return
    if args.workload == "batch":
        print(json.dumps(batch_renders(P, args.steps, args.probe_divisor, schedule=args.batch_schedule)), flush=True)
        return
    t0 = time.time()
    scene = P.Scene.preset(1, seed=0, param=args.level)
    info = scene.info()
    log(f"scene: {info.n_triangles} triangles, env {info.env_width}x{info.env_height} ({time.time() - t0:.1f}s)")
    r = P.Renderer(WIDTH, HEIGHT)
    if args.pool:
        r.set_option(P.OPT_POOL_PATHS, args.pool)
    if args.trace_blocks:
        r.set_option(P.OPT_TRACE_BLOCKS_PER_CU, args.trace_blocks)
    if args.bvh != "default":
        r.set_option(P.OPT_BVH_BUILDER, P.BVH_LBVH if args.bvh == "lbvh" else P.BVH_PLOC)
    t0 = time.time()
    r.set_scene(scene)
    build_s = time.time() - t0
    log(f"upload + GPU BVH + alias tables: {build_s:.2f}s")
    r.set_max_depth(DEPTH)
    r.set_nee(True)
    for i in range(args.warmup):
        r.render(SPP)
        log(f"warmup step {i} done")
    r.set_option(P.OPT_PROFILE_KERNELS, 1)
    r.reset_stats()
    r.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        r.render_async(SPP)
        r.synchronize()
        log(f"step {i} done")
    elapsed = time.perf_counter() - t0
    st = r.stats()
    ks = r.kernel_stats()
    assert st.stack_overflows == 0, "traversal stack overflow"
    mrays = st.rays_total / elapsed / 1e6
    samples = WIDTH * HEIGHT * SPP * args.steps / elapsed
    kernels = {}
    for name, k in ks.items():
        if k["launches"] == 0:
            continue
        avg_ms = k["ms"] / k["launches"]
        kernels[name] = {
            "launches": k["launches"], "total_ms": round(k["ms"], 3), "avg_us": round(avg_ms * 1e3, 2),
            "units": k["units"], "bytes_per_launch": k["bytes"] / k["launches"],
            "gbs": (k["bytes"] / (k["ms"] * 1e-3) / 1e9) if k["ms"] > 0 else 0.0,
        }
    # traversal statistics: one more pass, untimed, with the instrumented trace kernels
    r.set_option(P.OPT_PROFILE_KERNELS, 0)
    r.set_option(P.OPT_TRACE_STATS, 1)
    r.reset_stats()
    r.render(SPP)
    ts = r.stats()
    r.set_option(P.OPT_TRACE_STATS, 0)
    traversal = {}
    for q, name in ((0, "closest_hit"), (1, "shadow")):
        n = max(1, ts.trace_rays[q])
        traversal[name] = {"rays": ts.trace_rays[q], "node_visits_per_ray": round(ts.trace_nodes[q] / n, 2),
                           "prim_tests_per_ray": round(ts.trace_prims[q] / n, 2)}
    dom = max(kernels, key=lambda n: kernels[n]["total_ms"])
    d = kernels[dom]
    roof = {
        "kernel": dom, "bound": "hbm", "achieved": round(d["gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        # the traversal kernels are bound by neither: they issue VALU instructions most of their
        # time over divergent lanes (rocprofv3 counters in profiles/)
        "limiter": "VALU issue over divergent lanes (node and primitive tests), not HBM or MFMA",
        "frac": round(d["gbs"] / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom),
        "bytes_per_launch": d["bytes_per_launch"], "avg_launch_us": d["avg_us"],
    }
    # queue / path-state kernels (no traversal): algorithmic bytes over their summed time
    qk = [k for k in ("raygen", "shade", "accumulate", "tonemap") if k in kernels]
    qbytes = sum(kernels[k]["bytes_per_launch"] * kernels[k]["launches"] for k in qk)
    qms = sum(kernels[k]["total_ms"] for k in qk)
    queue_roof = {"kernels": qk, "achieved": round(qbytes / (qms * 1e-3) / 1e9, 1) if qms > 0 else 0.0, "peak": HBM_PEAK_GBS,
                  "unit": "GB/s", "frac": round(qbytes / (qms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if qms > 0 else 0.0}
    out = {
        "metric": "Mrays/s", "value": round(mrays, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"displaced icosphere level {args.level} ({info.n_triangles} tris) + {info.env_width}x{info.env_height} "
                               f"synthetic HDR env, {WIDTH}x{HEIGHT}, {SPP} spp, depth {DEPTH}, NEE+MIS, FP64 accumulation",
                   "rays_per_step": st.rays_total // args.steps, "samples_per_s": round(samples),
                   "bvh_depth": st.bvh_depth, "bvh_nodes": st.bvh_nodes,
                   "bvh_builder": "lbvh" if r.get_option(P.OPT_BVH_BUILDER) == P.BVH_LBVH else "ploc",
                   "scene_upload_s": round(build_s, 3), "parallelism": "1 GPU"},
        "roofline": roof,
        "queue_roofline": queue_roof,
        "traversal": traversal,
        "kernels": kernels,
    }
    if args.cpu_budget > 0:
        out["cpu_baseline"] = cpu_baseline(scene.export(), args.cpu_budget)
    r.close()
    if args.room_budget > 0:
        out["mse_ratio"] = room_summary(room_equal_time(P, args.room_budget, 1, args.room_repeats))
    if args.batch_steps > 0:
        b = batch_renders(P, args.batch_steps, args.probe_divisor)
        out["batch"] = {k: b[k] for k in ("metric", "value", "unit", "ms_per_step", "steps", "latency_ms", "config", "autotune")}
    print(json.dumps(out), flush=True)


def room_summary(res):
    """The equal-time room measurement without its convergence curves, with RMSE per renderer."""
    keep = ("mse", "rel_mse", "spp", "seconds", "mrays_s", "guide_steps", "mse_runs")
    out = {"metric": res["metric"], "value": res["value"], "unit": res["unit"], "config": res["config"]}
    for m in ("plain", "guided"):
        out[m] = {k: res[m][k] for k in keep if k in res[m]}
        out[m]["rmse"] = round(math.sqrt(res[m]["mse"]), 5)
    return out


if __name__ == "__main__":
    main()
