"""Writes room_reference_<N>px.npz: the FP64 oracle's 65,536-sample estimate at a fixed seeded
subset of N pixels (default 1024) of the hard-indirect room (2048x1024, depth 16, NEE + consistent MIS,
alias-table environment sampling).  The room renderer benchmark measures the mean squared
error of the GPU renderers against it.

The oracle runs with its own PCG32 seed, so its samples are independent of the GPU renders
(seed 0).  Run from the repository root after building the oracle and the package:

    python tests/golden/make_room_reference.py [--pixels 1024] [--spp 65536] [--threads 0]
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "gpu-pathed_amd"), os.path.join(ROOT, "oracle")]
import oracle as O  # noqa: E402
import pathed_capi as P  # noqa: E402

W, H, DEPTH, PIXEL_SEED, ORACLE_SEED = 2048, 1024, 16, 20240611, 0x5EED0F0A


def scene_fingerprint(arrays):
    """Sums of the scene arrays that change if the room generator changes."""
    return np.array([arrays["tri_pos"].shape[0], float(np.sum(arrays["tri_pos"], dtype=np.float64)),
                     float(np.sum(arrays["mat_alpha"], dtype=np.float64)),
                     float(np.sum(arrays["env_rgba"], dtype=np.float64))], np.float64)


def room_pixels(n):
    rng = np.random.default_rng(PIXEL_SEED)
    return np.sort(rng.choice(W * H, n, replace=False)).astype(np.int32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pixels", type=int, default=1024)
    ap.add_argument("--spp", type=int, default=65536)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--out", default="")
    a = ap.parse_args()
    out = a.out or os.path.join(os.path.dirname(os.path.abspath(__file__)), f"room_reference_{a.pixels}px.npz")
    s = P.Scene.preset(2, 0)
    arr = s.export()
    o = O.Scene(arr, W, H, bvh=True)
    pix = room_pixels(a.pixels)
    t = time.time()
    tot, sq, rays = o.render_pixels(pix, spp=a.spp, max_depth=DEPTH, nee=1, mode=1, tri_test=1, ref_env=0, precision=64,
                                    seed=ORACLE_SEED, mis_sample=1, threads=a.threads)
    dt = time.time() - t
    np.savez(out, pixels=pix, sum=tot, sumsq=sq, spp=np.int64(a.spp), rays=np.int64(rays),
             fingerprint=scene_fingerprint(arr), size=np.array([W, H, DEPTH], np.int32), seconds=np.float64(dt))
    print(f"{a.spp} spp x {pix.size} pixels, {rays} rays in {dt:.1f} s -> {out}", flush=True)


if __name__ == "__main__":
    main()
