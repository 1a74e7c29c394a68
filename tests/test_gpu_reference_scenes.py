"""The reference's own scene files on the GPU: the Cornell box OBJ (scene 0), the glossy
Cornell box with a mirror box and a glass ball (scene 1) and the Veach MIS plates with five
sphere lights and the wells6 environment (scene 3), loaded from the committed copies of its
scenes/ files, against the oracle on the same arrays; and the session's EXR dump at
power-of-two sample counts."""
import os
import sys

import numpy as np
import pytest

import oracle as O
import pathed_capi as P
from helpers import one_texel_plane_scene, path_agreement

pytestmark = pytest.mark.gpu

SCENES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if P.device_count() == 0:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")


@pytest.mark.parametrize("index", [0, 1, 2, 3])
def test_reference_scene_per_path_parity(index):
    say(f"case: reference scene {index} per-path parity (64x64x4, depth 6)")
    W = H = 64
    s = P.Scene.reference(index, SCENES)
    r = P.Renderer(W, H)
    r.set_scene(s)
    r.set_max_depth(6)
    r.render(4)
    gl = r.path_radiance(W * H * 4)
    o = O.Scene(s.export(), W, H)
    pix = np.tile(np.arange(W * H), 4)
    smp = np.repeat(np.arange(4), W * H)
    frac, self_frac, _ = path_agreement(o, gl, pix, smp, max_depth=6)
    say(f"per-path agreement {frac:.5f} (oracle FP32 vs FP64: {self_frac:.5f}), mean radiance {gl.mean():.4f}")
    # measured: scene 0 0.99982, scene 1 0.99957, scene 3 0.95166.  Scene 3's plates of
    # alpha 0.005 evaluate the reference's Beckmann terms next to their FP32 cancellation
    # (tan^2 = (1 - cos^2) / cos^2 for cos ~ 1): there the oracle's own FP32 and FP64 runs agree
    # on only 0.94 of the paths at 1e-3, and on 0.984 at 1e-2
    floor = {0: 0.9995, 1: 0.999, 2: 0.9995, 3: 0.945}[index]
    assert frac >= floor and frac >= self_frac - 0.002
    # scene 2 is lit by one environment texel: its plane is dim, the texel itself is 1e4
    assert gl.mean() > (1e-3 if index == 2 else 0.01) and np.isfinite(gl).all()
    assert r.stats().stack_overflows == 0
    r.close()


def test_reference_scene0_image_band():
    """Cornell OBJ image (FP64 accumulation) vs the FP64 oracle restating the reference
    integrator (two traces per bounce, Moeller-Trumbore, CMF environment): 99% band."""
    say("case: reference scene 0 image band")
    W = H = 32
    s = P.Scene.reference(0, SCENES)
    r = P.Renderer(W, H)
    r.set_scene(s)
    r.set_max_depth(6)
    passes, per = 16, 64
    imgs = []
    prev = np.zeros((H, W, 3))
    for _ in range(passes):
        r.render(per)
        cur = r.radiance_sum()
        imgs.append((cur - prev) / per)
        prev = cur
    imgs = np.array(imgs)
    Ng = passes * per
    o = O.Scene(s.export(), W, H)
    No = 2048
    tot, sq, _ = o.render(want_sq=True, spp=No, max_depth=6, nee=1, mode=0, tri_test=0, ref_env=1, precision=64, seed=5)
    om = tot / No
    v = (imgs.var(0, ddof=1) * per * Ng + (sq / No - om**2) * No) / (Ng + No)
    se2 = v * (1 / Ng + 1 / No)
    z = (imgs.mean(0) - om) / np.sqrt(se2 + 1e-12)
    outside = (np.abs(z) > 2.576).mean()
    zt = (imgs.mean(0) - om).sum((0, 1)) / np.sqrt(se2.sum((0, 1)))
    say(f"outside the 99% band {outside:.4f}; whole-image z {np.round(zt, 2)}")
    assert outside < 0.025 and np.all(np.abs(zt) < 4)
    r.close()


def test_session_exr_dump_at_powers_of_two(tmp_path):
    say("case: session EXR dump")
    import path_tracer as pt

    path = str(tmp_path / "out.exr")
    s = pt.RenderSession(40, 30)
    s.loadReferenceScene(0, SCENES)
    s.setExrDumpPath(path)
    s.init(0, 0)
    while s.getAccumulatedSpp() < 2:
        s.renderAsync()
        while s.pollRender().isRendering:
            pass
    assert os.path.exists(path)
    img = P.exr_read(path)[::-1, :, :3]  # EXR scanlines top-down, radiance rows bottom-up
    with np.errstate(over="ignore"):
        want = s.getRadianceBuffer().astype(np.float16).astype(np.float32)
    # the loop stops after the pass that crossed 2 spp, whose radiance the dump holds
    assert img.shape == (30, 40, 3) and np.isfinite(img).all() and img.mean() > 0
    assert np.array_equal(img, want)


def test_reference_scene2_one_texel_closed_form():
    """Scene 2 (the plane under the PIZ-compressed one-texel map) seen from straight above: the
    GPU image's mean and every pixel agree with the closed form albedo / pi x texel irradiance
    (NEE samples the texel through the alias tables; 256 spp)."""
    say("case: reference scene 2, one-texel environment closed form")
    s, want = one_texel_plane_scene(SCENES)
    W = H = 64
    r = P.Renderer(W, H)
    r.set_scene(s)
    r.set_max_depth(3)
    r.render(256)
    img = r.radiance()[..., 0].astype(np.float64)
    r.close()
    say(f"mean {img.mean():.6g} vs closed form {want:.6g}; pixel range [{img.min():.6g}, {img.max():.6g}]")
    assert abs(img.mean() - want) < 2e-3 * want
    assert np.all(np.abs(img - want) < 0.02 * want)
