"""Parity of the GPU renderer with the oracle on the benchmark workloads themselves: the
Cornell box at its full size (256x256, 16 spp, depth 6), the displaced icosphere of 1.31M
triangles with its HDR environment (depth 8), and the 4M-triangle room (depth 16), path by
path on a pixel subset with the same PCG32 streams, plus the Cornell image against the FP64
reference-semantics oracle."""
import sys

import numpy as np
import pytest

import oracle as O
import pathed_capi as P

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if P.device_count() == 0:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")


def _per_path(scene, W, H, spp, depth, rows, consistent=False, min_frac=0.99):
    """Renders W x H x spp on the GPU; compares the paths of every pixel of `rows` (row
    indices) with the oracle's product-mode integrator."""
    r = P.Renderer(W, H)
    r.set_scene(scene)
    r.set_max_depth(depth)
    if consistent:
        r.set_option(P.OPT_CONSISTENT_MIS, 1)
    r.render(spp)
    gl = r.path_radiance(W * H * spp).reshape(spp, H * W, 3)
    st = r.stats()
    r.close()
    pix = np.concatenate([np.arange(W) + y * W for y in rows])
    o = O.Scene(scene.export(), W, H, bvh=True)
    pp = np.tile(pix, spp)
    ss = np.repeat(np.arange(spp), pix.size)
    cl = o.trace_paths(pp, ss, max_depth=depth, nee=1, mode=1, tri_test=1, ref_env=0, precision=32, mis_sample=int(consistent))
    g = gl[:, pix].reshape(-1, 3)
    err = np.abs(g - cl).max(1) / (np.abs(cl).max(1) + 1e-3)
    frac = (err < 1e-3).mean()
    return frac, g, cl, st


def test_cornell_full_size_c0():
    """256x256, 16 spp, depth 6: per-path parity on 8 rows and the whole image in the 99% band
    of the FP64 oracle restating the reference integrator."""
    say("case: Cornell 256x256x16 depth 6")
    s = P.Scene.preset(0)
    W = H = 256
    frac, g, cl, st = _per_path(s, W, H, 16, 6, rows=range(0, 256, 32))
    say(f"per-path agreement {frac:.5f} over {g.shape[0]} paths; rays per pass {st.rays_total}")
    assert frac >= 0.998 and st.stack_overflows == 0  # measured 0.99915
    # image: GPU 16 spp x 8 passes vs the FP64 oracle at 256 spp (reference semantics)
    r = P.Renderer(W, H)
    r.set_scene(s)
    r.set_max_depth(6)
    passes, per = 8, 16
    imgs, prev = [], np.zeros((H, W, 3))
    for _ in range(passes):
        r.render(per)
        cur = r.radiance_sum()
        imgs.append((cur - prev) / per)
        prev = cur
    r.close()
    imgs = np.array(imgs)
    o = O.Scene(s.export(), W, H)
    No = 256
    tot, sq, _ = o.render(want_sq=True, spp=No, max_depth=6, nee=1, mode=0, tri_test=0, ref_env=1, precision=64, seed=11)
    Ng = passes * per
    om = tot / No
    v = (imgs.var(0, ddof=1) * per * Ng + (sq / No - om**2) * No) / (Ng + No)
    se2 = v * (1 / Ng + 1 / No)
    z = (imgs.mean(0) - om) / np.sqrt(se2 + 1e-12)
    outside = (np.abs(z) > 2.576).mean()
    zt = (imgs.mean(0) - om).sum((0, 1)) / np.sqrt(se2.sum((0, 1)))
    say(f"outside the 99% band {outside:.4f}; whole-image z {np.round(zt, 2)}")
    assert outside < 0.025 and np.all(np.abs(zt) < 4)


def test_icosphere_c1_scene_per_path():
    """The 1.31M-triangle displaced icosphere with its 1024x512 environment (sun lobe), 64x64
    pixels, 2 spp, depth 8."""
    say("case: icosphere level 8 per-path (64x64x2, depth 8)")
    s = P.Scene.preset(1, seed=0, param=8)
    frac, g, cl, st = _per_path(s, 64, 64, 2, 8, rows=range(0, 64, 2))
    say(f"per-path agreement {frac:.5f} over {g.shape[0]} paths; mean radiance {g.mean():.4f}")
    assert frac >= 0.9995 and st.stack_overflows == 0  # measured 1.00000
    assert np.isfinite(g).all() and g.mean() > 0


def test_icosphere_camera_packets_per_path():
    """A chunk of 64 sample rows traces each camera run (the 64 samples of one pixel) as one
    packet on one wave stack: per-path parity with the oracle on the 1.31M-triangle icosphere
    (16x16 pixels, 64 spp, depth 8, every other row compared), and the same path radiances as
    the per-lane traversal (the instrumented trace kernels, PT_OPT_TRACE_STATS, trace per
    lane), up to equal-distance ties."""
    say("case: icosphere camera packets per-path (16x16x64, depth 8)")
    s = P.Scene.preset(1, seed=0, param=8)
    frac, g, cl, st = _per_path(s, 16, 16, 64, 8, rows=range(0, 16, 2))
    say(f"per-path agreement {frac:.5f} over {g.shape[0]} paths; mean radiance {g.mean():.4f}")
    assert frac >= 0.9995 and st.stack_overflows == 0
    lanes = []
    for stats in (0, 1):
        r = P.Renderer(16, 16)
        r.set_scene(s)
        r.set_max_depth(8)
        r.set_option(P.OPT_TRACE_STATS, stats)
        r.render(64)
        lanes.append(r.path_radiance(16 * 16 * 64))
        assert r.stats().stack_overflows == 0
        r.close()
    same = np.all(lanes[0] == lanes[1], axis=1).mean()
    say(f"packet vs per-lane traversal: {same:.5f} of paths identical")
    assert same >= 0.9995


def test_room_c4_scene_per_path():
    """The 4M-triangle room (slot light 1e3, glass sphere, environment), 64x32 pixels, 2 spp,
    depth 16, consistent MIS."""
    say("case: room per-path (64x32x2, depth 16)")
    s = P.Scene.preset(2, seed=0)
    frac, g, cl, st = _per_path(s, 64, 32, 2, 16, rows=range(0, 32, 2), consistent=True)
    say(f"per-path agreement {frac:.5f} over {g.shape[0]} paths; mean radiance {g.mean():.4f}")
    assert frac >= 0.995 and st.stack_overflows == 0  # measured 0.99951
    assert np.isfinite(g).all() and g.mean() > 0
