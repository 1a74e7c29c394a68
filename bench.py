"""Benchmark of the progressive path-tracing pass on one MI355X.

Workload (one step = one full render): displaced icosphere, subdivision level 8
(1,310,720 triangles, 3-octave value-noise displacement of amplitude 0.05, per-face
Beckmann alpha U[0.05, 0.5] with 30% Lambert faces), synthetic 1024x512 HDR environment
(log-normal sigma 1 + 1e4 sun lobe) importance-sampled through FP64-built alias tables,
1024x1024 pixels, 64 samples per pixel, depth 8, NEE + MIS, FP32 tracing, FP64
accumulation.  Synthetic data, seeded.

Prints one JSON line: Mrays/s (primary + extension + shadow rays actually traced, divided by
the wall time of the timed steps), with the roofline of the dominant kernel measured live
from HIP events on the renderer's stream, and the CPU oracle timed on a bounded sample of the
same workload on the host cores.
"""
import argparse
import json
import os
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


def cpu_baseline(scene_arrays, budget_s):
    """The oracle (scalar C++ restatement of the reference's shading and intersection, with
    a median-split BVH) on the host cores: 1 spp over a band of rows of the same image,
    sized to about `budget_s` seconds."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np  # noqa
    import oracle as O

    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.time()
    o = O.Scene(scene_arrays, WIDTH, HEIGHT, bvh=True)
    log(f"cpu baseline: oracle scene + BVH built in {time.time() - t0:.1f}s")
    kw = dict(spp=1, max_depth=DEPTH, nee=1, mode=1, tri_test=1, ref_env=0, precision=32, threads=threads)
    # probe 8 rows of the image centre, then size the band to the budget
    t0 = time.time()
    _, _, rays = o.render(row_begin=508, row_end=516, **kw)
    dt = max(time.time() - t0, 1e-3)
    rows = int(min(HEIGHT, max(8, 8 * budget_s / dt)))
    rb = max(0, HEIGHT // 2 - rows // 2)
    t0 = time.time()
    _, _, rays = o.render(row_begin=rb, row_end=rb + rows, **kw)
    dt = time.time() - t0
    return {
        "value": rays / dt / 1e6,
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"1 spp over rows [{rb}, {rb + rows}) of the 1024x1024 image ({rows * WIDTH} paths, {rays} rays, {dt:.1f} s), "
                  f"same scene, depth 8, NEE+MIS, FP32, median-split BVH",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU-oracle work for cpu_baseline (0: skip)")
    ap.add_argument("--level", type=int, default=8, help="icosphere subdivision level")
    ap.add_argument("--pool", type=int, default=0, help="path pool size (0: library default)")
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
    t0 = time.time()
    scene = P.Scene.preset(1, seed=0, param=args.level)
    info = scene.info()
    log(f"scene: {info.n_triangles} triangles, env {info.env_width}x{info.env_height} ({time.time() - t0:.1f}s)")
    r = P.Renderer(WIDTH, HEIGHT)
    if args.pool:
        r.set_option(P.OPT_POOL_PATHS, args.pool)
    t0 = time.time()
    r.set_scene(scene)
    log(f"upload + GPU BVH + alias tables: {time.time() - t0:.2f}s")
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
    dom = max(kernels, key=lambda n: kernels[n]["total_ms"])
    d = kernels[dom]
    roof = {
        "kernel": dom, "bound": "hbm", "achieved": round(d["gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(d["gbs"] / HBM_PEAK_GBS, 4), "traffic": None,
        "bytes_per_launch": d["bytes_per_launch"], "avg_launch_us": d["avg_us"],
    }
    out = {
        "metric": "Mrays/s", "value": round(mrays, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"displaced icosphere level {args.level} ({info.n_triangles} tris) + {info.env_width}x{info.env_height} "
                               f"synthetic HDR env, {WIDTH}x{HEIGHT}, {SPP} spp, depth {DEPTH}, NEE+MIS, FP64 accumulation",
                   "rays_per_step": st.rays_total // args.steps, "samples_per_s": round(samples),
                   "bvh_depth": st.bvh_depth, "parallelism": "1 GPU"},
        "roofline": roof,
        "kernels": kernels,
    }
    if args.cpu_budget > 0:
        out["cpu_baseline"] = cpu_baseline(scene.export(), args.cpu_budget)
    r.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
