"""GPU tests of the consistent-MIS option and the learned-guiding renderer.

Per-path parity: with consistent MIS the light-sample weight uses the density sample() draws
with; the wavefront's paths match the oracle's restatement of that estimator path by path.
Guided rendering draws directions from a network that changes between passes, so it has no
per-path counterpart: its image is checked against the FP64 oracle within a confidence band
(the one-sample mixture with the mixture density in both MIS weights is unbiased), its
training loss must fall as it learns, and on the hard-indirect room its error against the
FP64 oracle's 65,536-sample reference is lower than the plain renderer's at the same sample
count: what the trained network gains over the method without it.
"""
import math
import os
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


def _quad(a, b, c, d):
    return [np.concatenate([a, b, c]), np.concatenate([a, c, d])]


def glossy_scene(env=True):
    """A Beckmann floor, a small downward quad light, a diffuse and a glass sphere and a low-
    resolution environment map."""
    s = P.Scene()
    floor = s.add_material(P.MICROFACET, alpha=0.25)
    light = s.add_material(P.LAMBERTIAN, albedo=(0.5, 0.5, 0.5), emit=(12, 10, 8))
    diff = s.add_material(P.LAMBERTIAN, albedo=(0.7, 0.3, 0.2))
    glass = s.add_material(P.GLASS, ior=1.5)
    f = np.array([(-3, 0, -3), (3, 0, -3), (3, 0, 3), (-3, 0, 3)], np.float64)
    q = np.array([(-0.3, 1.6, -0.3), (0.3, 1.6, -0.3), (0.3, 1.6, 0.3), (-0.3, 1.6, 0.3)], np.float64)
    # floor wound to face +y, light quad wound to face -y
    tris = _quad(f[0], f[3], f[2], f[1]) + _quad(q[0], q[1], q[2], q[3])
    s.add_triangles(np.array(tris), [floor, floor, light, light])
    s.add_spheres([(-0.6, 0.4, 0.0), (0.6, 0.4, 0.3)], [0.4, 0.4], [diff, glass])
    if env:
        rng = np.random.default_rng(5)
        h, w = 32, 64
        rgba = np.ones((h, w, 4), np.float32)
        rgba[..., :3] = np.exp(rng.normal(-1.0, 0.8, (h, w, 3)))
        rgba[4:7, 20:24, :3] = 40.0  # a small bright patch above the horizon
        s.set_environment(rgba)
    s.set_camera((0, 1.4, 3.6), (0, 0.3, 0), (0, 1, 0), 0.75)
    return s


def test_consistent_mis_per_path_parity():
    """Consistent MIS on a glossy + environment scene: per-path radiance matches the oracle's
    product-mode integrator with the same option (same PCG32 streams)."""
    say("case: consistent-MIS per-path parity (glossy scene, 48x48x4, depth 5)")
    W = H = 48
    s = glossy_scene()
    r = P.Renderer(W, H)
    r.set_scene(s)
    r.set_max_depth(5)
    r.set_option(P.OPT_CONSISTENT_MIS, 1)
    assert r.get_option(P.OPT_CONSISTENT_MIS) == 1
    r.render(4)
    gl = r.path_radiance(W * H * 4)
    o = O.Scene(s.export(), W, H)
    pix = np.tile(np.arange(W * H), 4)
    smp = np.repeat(np.arange(4), W * H)
    cl = o.trace_paths(pix, smp, max_depth=5, nee=1, mode=1, tri_test=1, ref_env=0, precision=32, mis_sample=1)
    err = np.abs(gl - cl).max(1) / (np.abs(cl).max(1) + 1e-3)
    frac = (err < 1e-3).mean()
    # the reference weighting differs on the glossy floor: the option must change the result
    cr = o.trace_paths(pix, smp, max_depth=5, nee=1, mode=1, tri_test=1, ref_env=0, precision=32, mis_sample=0)
    frac_ref = ((np.abs(gl - cr).max(1) / (np.abs(cr).max(1) + 1e-3)) < 1e-3).mean()
    say(f"per-path agreement {frac:.5f} (against the reference weighting: {frac_ref:.5f})")
    assert frac >= 0.999  # measured 0.99989
    assert frac_ref < frac - 0.05
    assert np.isfinite(gl).all()
    r.close()


def _pass_images(r, passes, per):
    imgs = []
    prev = np.zeros((r.H, r.W, 3))
    losses = []
    for _ in range(passes):
        r.render(per)
        cur = r.radiance_sum()
        imgs.append((cur - prev) / per)
        prev = cur
        losses.append(r.stats().guide_loss)
    return np.array(imgs), losses


def _band(imgs, per, o_tot, o_sq, No):
    Ng = imgs.shape[0] * per
    g_mean = imgs.mean(0)
    o_mean = o_tot / No
    v = (imgs.var(0, ddof=1) * per * Ng + (o_sq / No - o_mean**2) * No) / (Ng + No)
    se2 = v * (1 / Ng + 1 / No)
    z = (g_mean - o_mean) / np.sqrt(se2 + 1e-12)
    outside = (np.abs(z) > 2.576).mean()
    zt = (g_mean - o_mean).sum((0, 1)) / np.sqrt(se2.sum((0, 1)))
    return outside, zt


@pytest.mark.parametrize("scene", ["cornell", "glossy"])
def test_guided_image_within_confidence_band(scene):
    """The guided renderer (network trained between passes) converges to the FP64 oracle's
    image: 99% band per pixel, whole-image |z| < 4."""
    say(f"case: guided image band ({scene})")
    W = H = 32
    depth = 5
    s = P.Scene.preset(0) if scene == "cornell" else glossy_scene()
    r = P.Renderer(W, H)
    r.set_scene(s)
    r.set_max_depth(depth)
    r.set_type(P.RENDERER_GUIDED)
    r.set_option(P.OPT_CONSISTENT_MIS, 1)
    passes, per = 16, 64
    imgs, _ = _pass_images(r, passes, per)
    st = r.stats()
    say(f"guide steps {st.guide_steps}, last loss {st.guide_loss:.4f}")
    assert st.guide_steps > 0
    o = O.Scene(s.export(), W, H)
    No = 2048
    tot, sq, _ = o.render(want_sq=True, spp=No, max_depth=depth, nee=1, mode=1, tri_test=1, ref_env=0,
                          precision=64, seed=77, mis_sample=1)
    outside, zt = _band(imgs, per, tot, sq, No)
    say(f"fraction outside the 99% band: {outside:.4f}; whole-image z {np.round(zt, 2)}")
    assert outside < 0.025
    assert np.all(np.abs(zt) < 4)
    assert np.isfinite(imgs).all()
    r.close()


def test_guided_training_lowers_loss():
    """The weighted negative log-likelihood of the training records falls as the network
    trains on the glossy scene."""
    say("case: guided training loss (glossy scene 64x64)")
    s = glossy_scene()
    r = P.Renderer(64, 64)
    r.set_scene(s)
    r.set_max_depth(5)
    r.set_type(P.RENDERER_GUIDED)
    r.set_option(P.OPT_GUIDE_TRAIN_STEPS, 4)
    assert r.get_option(P.OPT_GUIDE_TRAIN_STEPS) == 4
    _, losses = _pass_images(r, 40, 16)
    losses = np.array(losses)
    say(f"loss first passes {np.round(losses[:5], 4)}, last {np.round(losses[-5:], 4)}")
    assert np.isfinite(losses).all()
    assert losses[-5:].mean() < losses[:5].mean() - 0.05 * abs(losses[:5].mean())
    steps = r.stats().guide_steps
    assert steps == 40 * 4
    # reset: the network is re-initialised and the step count starts again
    r.set_option(P.OPT_GUIDE_RESET, 1)
    r.render(16)
    assert r.stats().guide_steps == 4
    r.close()


ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROOM_REFERENCE = os.path.join(ROOT, "tests", "golden", "room_reference_1024px.npz")


def _room_pixel_errors(scene, guided, spp, seed):
    """Squared errors (per pixel, summed over RGB) at the reference fixture's 1024 pixels of a
    2048x1024 render of the room (depth 16, NEE + consistent MIS) after `spp` one-sample
    passes; the guided renderer starts from an untrained network and trains after each."""
    ref = np.load(ROOM_REFERENCE)
    pix = ref["pixels"]
    ref_mean = ref["sum"] / ref["spp"]
    r = P.Renderer(2048, 1024)
    r.set_scene(scene)
    r.set_max_depth(16)
    r.set_option(P.OPT_CONSISTENT_MIS, 1)
    r.set_seed(seed)
    if guided:
        r.set_type(P.RENDERER_GUIDED)
        r.set_option(P.OPT_GUIDE_RESET, seed + 1)
    for _ in range(spp):
        r.render(1)
    assert r.spp() == spp
    img = r.radiance_sum().reshape(-1, 3)[pix] / spp
    st = r.stats()
    r.close()
    return ((img - ref_mean) ** 2).sum(1), st


def test_guided_beats_plain_at_equal_samples():
    """The room of 4M triangles lit through a 0.3-wide slot (the equal-time benchmark's scene):
    at 256 samples per pixel each, the guided renderer's squared error against the FP64
    oracle's 65,536-sample estimate (the committed fixture, whose own noise is 0.0014) is at
    most half the plain renderer's, and a paired bootstrap over the 1024 pixels puts the ratio's
    lower 5% quantile (a one-sided 95% bound) above 1.2.  The per-pixel errors are heavy-tailed:
    single samples of the 1e3 light behind the slot, reached through specular chains neither
    method samples well, leave fireflies whose squared error can exceed that of all other
    pixels together (one guided run in about ten: raw MSE ratio 0.99 with a median pixel error
    half the plain one).  The test therefore compares the mean squared error over all but each
    image's 10 worst pixels (1%); over 8 guided runs that ratio measured 3.9-6.2, its
    bootstrap bound 2.1-3.4, the raw MSE ratio (reported, not asserted) 3.5-6.8.  Both use the
    same PCG32 seeds; the oracle used its own."""
    say("case: guided vs plain on the room at equal spp")
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_room_reference import scene_fingerprint

    ref = np.load(ROOM_REFERENCE)
    scene = P.Scene.preset(2, seed=0)
    assert np.allclose(scene_fingerprint(scene.export()), ref["fingerprint"], rtol=1e-9)
    spp = 256
    ep, sp = _room_pixel_errors(scene, False, spp, seed=11)
    eg, sg = _room_pixel_errors(scene, True, spp, seed=11)
    trim = 10

    def tmean(e):
        return np.sort(e, axis=-1)[..., :-trim].mean(-1)

    ratio = float(tmean(ep) / tmean(eg))
    rng = np.random.default_rng(0)
    idx = rng.integers(0, ep.size, (2000, ep.size))
    boot = tmean(ep[idx]) / tmean(eg[idx])
    lo = float(np.quantile(boot, 0.05))
    top = np.argsort(ep)[::-1][:8]
    say(f"trimmed MSE plain {tmean(ep) / 3:.5f}, guided {tmean(eg) / 3:.5f}: ratio {ratio:.2f} (bootstrap 5% quantile "
        f"{lo:.2f}); raw MSE plain {ep.mean() / 3:.5f}, guided {eg.mean() / 3:.5f}: ratio {ep.mean() / eg.mean():.2f}; "
        f"median pixel error plain {np.median(ep):.4g} guided {np.median(eg):.4g}; share of the 8 worst plain pixels "
        f"{ep[top].sum() / ep.sum():.2f}; guide steps {sg.guide_steps}, rays plain {sp.rays_total} guided {sg.rays_total}")
    assert sg.guide_steps >= spp and np.isfinite(eg).all()
    assert ratio > 2.0 and lo > 1.2
