"""GPU tests of batched renders (one captured graph) and of the autotuner.

A render inside a batch consumes the same PCG32 streams as the same render on its own
renderer, so with one chunk each the images agree to the last FP32 rounding of the mean.
That holds for both schedules: the wavefront (paths of every render in the same queues, chunks
of sample rows of several renders) and render after render.  The captured graph is replayed
several times, with eager work of the same program between the replays, and every replay must
reproduce the eager image.

At the batch workload's extremes (a 1024^2 x 256 spp render of ~2M triangles next to a
128^2 x 4 spp one) each render's pixel means on a seeded pixel subset match the oracle run
on the same PCG32 streams.
"""
import sys

import numpy as np
import pytest

import pathed_capi as P

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if P.device_count() == 0:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")


JOBS = [  # (preset, seed, param, width, height, spp, depth)
    (0, 0, 0, 48, 40, 8, 6),
    (4, 1, 12000, 64, 48, 4, 5),
    (4, 2, 40000, 32, 32, 16, 8),
]


def _scenes():
    return [P.Scene.preset(p, seed=s, param=q) for p, s, q, *_ in JOBS]


def _single(scene, w, h, spp, depth):
    r = P.Renderer(w, h)
    r.set_scene(scene)
    r.set_max_depth(depth)
    r.render(spp)
    img = r.radiance()
    rays = r.stats().rays_total
    r.close()
    return img, rays


def _batch(scenes, graph=True, wavefront=True, pool=0):
    b = P.Batch()
    for sc, (_, _, _, w, h, spp, depth) in zip(scenes, JOBS):
        b.add(sc, w, h, spp, depth)
    b.set_option(P.OPT_BATCH_GRAPH, int(graph))
    b.set_option(P.OPT_BATCH_WAVEFRONT, int(wavefront))
    if pool:
        b.set_option(P.OPT_POOL_PATHS, pool)
    return b


@pytest.mark.parametrize("wavefront", [1, 0])
def test_batch_matches_single_renders_and_graph_replays(wavefront):
    say(f"case: batch of 3 renders (graph, wavefront {wavefront}) vs single renderers; 3 graph replays")
    scenes = _scenes()
    ref = [_single(sc, w, h, spp, depth) for sc, (_, _, _, w, h, spp, depth) in zip(scenes, JOBS)]
    b = _batch(scenes, wavefront=bool(wavefront))
    assert b.get_option(P.OPT_BATCH_WAVEFRONT) == wavefront
    b.run()
    st = b.stats()
    say(f"batch stats {st}")
    assert st["graph_nodes"] > 3 * 10 and st["stack_overflows"] == 0 and st["renders"] == 3
    first = [b.radiance(i) for i in range(3)]
    for i, (img, rays) in enumerate(ref):
        assert np.isfinite(first[i]).all()
        assert np.allclose(first[i], img, rtol=1e-6, atol=1e-7), i
        assert b.render_stats(i)["rays"] == rays
    for rep in range(3):
        # eager work of the same program between the replays
        _single(scenes[0], 16, 16, 2, 3)
        b.run()
        for i in range(3):
            assert np.array_equal(b.radiance(i), first[i]), (rep, i)
    b.close()


@pytest.mark.parametrize("wavefront", [1, 0])
def test_batch_direct_launches_equal_graph(wavefront):
    say(f"case: batch direct launches vs graph, wavefront {wavefront}")
    scenes = _scenes()
    g = _batch(scenes, graph=True, wavefront=bool(wavefront))
    g.run()
    d = _batch(scenes, graph=False, wavefront=bool(wavefront))
    d.run()
    assert d.stats()["graph_nodes"] == 0
    for i in range(3):
        assert np.array_equal(g.radiance(i), d.radiance(i))
    g.close()
    d.close()


@pytest.mark.parametrize("wavefront,pool", [(0, 0), (1, 0), (1, 12288)])
def test_batch_latency_follows_order(wavefront, pool):
    """Renders complete in the batch's order (cheapest first by pixels x spp x log2 primitives).
    Render after render, each completes after the previous one; in the wavefront a chunk's
    renders complete together, so with the whole batch in one chunk the latencies are equal and
    with a pool of one render's paths they follow the order."""
    say(f"case: batch latencies, wavefront {wavefront}, pool {pool}")
    scenes = _scenes()
    b = _batch(scenes, wavefront=bool(wavefront), pool=pool)
    b.run()
    lat = np.array([b.render_stats(i)["latency_ms"] for i in range(3)])
    st = b.stats()
    say(f"latencies {lat}, chunks {st['chunks']}")
    assert (lat > 0).all()
    cost = [w * h * spp * np.log2(2 + sc.info().n_triangles) for sc, (_, _, _, w, h, spp, _) in zip(scenes, JOBS)]
    order = np.argsort(cost, kind="stable")
    assert np.all(np.diff(lat[order]) >= 0)
    if not wavefront or pool:
        assert np.all(np.diff(lat[order]) > 0)
    if wavefront and not pool:
        # one chunk: one completion stamp for all three
        assert st["chunks"] == 1 and np.ptp(lat) == 0
    b.close()


def test_autotune_batch_and_renderer():
    say("case: autotuners")
    scenes = _scenes()
    b = _batch(scenes)
    t = b.autotune(2)
    say(f"batch tune {t}")
    assert t["probes"] >= 4 and t["mrays_per_s"] > 0
    assert t["trace_blocks_per_cu"] in (8, 11, 12, 14) and t["lds_stack"] in (16, 32)
    assert b.get_option(P.OPT_TRACE_BLOCKS_PER_CU) == t["trace_blocks_per_cu"]
    b.run()
    ref = _single(scenes[1], *JOBS[1][3:])[0]
    assert np.allclose(b.radiance(1), ref, rtol=1e-6, atol=1e-7)
    b.close()
    r = P.Renderer(64, 64)
    r.set_scene(scenes[2])
    r.set_max_depth(6)
    t = r.autotune(4)
    say(f"renderer tune {t}")
    assert t["probes"] >= 4 and t["mrays_per_s"] > 0
    assert r.get_option(P.OPT_TRACE_LDS_STACK) == t["lds_stack"]
    r.render(4)
    assert r.spp() == 4  # probe samples were discarded
    assert r.stats().stack_overflows == 0
    r.close()


MIXED = [  # (seed, triangles, resolution, spp): one render larger than the pool, small ones to share chunks
    (4001, 300_000, 512, 32),
    (4002, 20_000, 128, 4),
    (4003, 60_000, 256, 16),
    (4004, 10_000, 96, 8),
    (4005, 150_000, 128, 64),
]


def test_batch_schedules_agree_bit_for_bit():
    """The wavefront schedule (chunks of several renders through the batch kernels, and a large
    render's own chunks through its renderer's kernels) and the render-after-render schedule
    give every render the same image and ray count, bit for bit, under a pool small enough that
    the large render spans several chunks of its own."""
    say("case: batch schedules, bit for bit")
    scenes = [P.Scene.preset(4, seed=s, param=t) for s, t, _, _ in MIXED]
    out = {}
    for wave in (1, 0):
        b = P.Batch()
        for sc, (_, _, res, spp) in zip(scenes, MIXED):
            b.add(sc, res, res, spp, 8)
        b.set_option(P.OPT_BATCH_WAVEFRONT, wave)
        b.set_option(P.OPT_POOL_PATHS, 1 << 21)
        b.run()
        st = b.stats()
        say(f"wavefront {wave}: {st}")
        assert st["stack_overflows"] == 0
        out[wave] = ([b.radiance(i) for i in range(len(MIXED))], [b.render_stats(i)["rays"] for i in range(len(MIXED))], st)
        b.close()
    (iw, rw, sw), (is_, rs, ss) = out[1], out[0]
    assert sw["chunks"] < ss["chunks"] and sw["rays"] == ss["rays"]
    for i in range(len(MIXED)):
        assert rw[i] == rs[i], i
        assert np.array_equal(iw[i], is_[i]), i


EXTREMES = [  # (seed, triangles, resolution, spp): the batch workload's largest and smallest renders
    (3001, 2_000_000, 1024, 256),
    (3002, 10_000, 128, 4),
    (3003, 200_000, 512, 16),
]


def test_batch_extremes_match_the_oracle():
    """One wavefront over a 1024^2 x 256 spp render of ~2M triangles (268M paths, several
    chunks), a 128^2 x 4 spp render of 10k triangles and a 512^2 x 16 spp one: at 48 seeded pixels
    of each render, the batch's pixel means equal the oracle's means of the same PCG32 streams
    (each path of the pixel traced by the product-mode FP32 integrator) to 1e-3 relative."""
    import oracle as O

    say("case: batch at the workload's extremes vs the oracle")
    scenes = [P.Scene.preset(4, seed=s, param=t) for s, t, _, _ in EXTREMES]
    b = P.Batch()
    for sc, (_, _, res, spp) in zip(scenes, EXTREMES):
        b.add(sc, res, res, spp, 8)
    b.run()
    st = b.stats()
    say(f"batch stats {st}")
    assert st["stack_overflows"] == 0 and st["chunks"] >= 4
    rng = np.random.default_rng(7)
    for i, (sc, (seed, tris, res, spp)) in enumerate(zip(scenes, EXTREMES)):
        img = b.radiance(i).reshape(-1, 3)
        pix = np.sort(rng.choice(res * res, 48, replace=False))
        o = O.Scene(sc.export(), res, res, bvh=True)
        sums, _, rays = o.render_pixels(pix, spp=spp, max_depth=8, nee=1, mode=1, tri_test=1, ref_env=0, precision=32, threads=16)
        want = sums / spp
        got = img[pix].astype(np.float64)
        rel = np.abs(got - want).max(1) / (np.abs(want).max(1) + 1e-3)
        frac = float((rel < 1e-3).mean())
        rs = b.render_stats(i)
        say(f"render {i}: {res}^2 x {spp} spp, {sc.info().n_triangles} tris: {frac:.4f} of pixels within 1e-3 "
            f"(max rel {rel.max():.2e}); {rs['rays']} rays, latency {rs['latency_ms']:.1f} ms")
        assert frac >= 0.97 and np.isfinite(img).all(), i
        assert rs["rays"] > res * res * spp
    b.close()
