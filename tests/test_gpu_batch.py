"""GPU tests of batched renders (one captured graph) and of the autotuner.

A render inside a batch consumes the same PCG32 streams as the same render on its own
renderer, so with one chunk each the images agree to the last FP32 rounding of the mean.
The captured graph is replayed several times, with eager work of the same program between the
replays, and every replay must reproduce the eager image.
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


def _batch(scenes, graph=True):
    b = P.Batch()
    for sc, (_, _, _, w, h, spp, depth) in zip(scenes, JOBS):
        b.add(sc, w, h, spp, depth)
    b.set_option(P.OPT_BATCH_GRAPH, int(graph))
    return b


def test_batch_matches_single_renders_and_graph_replays():
    say("case: batch of 3 renders (graph) vs single renderers; 3 graph replays")
    scenes = _scenes()
    ref = [_single(sc, w, h, spp, depth) for sc, (_, _, _, w, h, spp, depth) in zip(scenes, JOBS)]
    b = _batch(scenes)
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


def test_batch_direct_launches_equal_graph():
    say("case: batch direct launches vs graph")
    scenes = _scenes()
    g = _batch(scenes, graph=True)
    g.run()
    d = _batch(scenes, graph=False)
    d.run()
    assert d.stats()["graph_nodes"] == 0
    for i in range(3):
        assert np.array_equal(g.radiance(i), d.radiance(i))
    g.close()
    d.close()


def test_batch_latency_follows_order():
    say("case: batch latencies")
    b = _batch(_scenes())
    b.run()
    lat = [b.render_stats(i)["latency_ms"] for i in range(3)]
    say(f"latencies {lat}")
    assert all(x > 0 for x in lat)
    # renders complete shortest-first by the cost model pixels x spp x log2(primitives)
    cost = [w * h * spp * np.log2(2 + P.Scene.preset(p, seed=s, param=q).info().n_triangles)
            for p, s, q, w, h, spp, _ in JOBS]
    assert list(np.argsort(lat)) == list(np.argsort(cost, kind="stable"))
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
