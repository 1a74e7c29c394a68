"""Host-side checks of libpathed (no GPU): every declared symbol is exported, the scene
generators produce the documented workloads, errors come back as status codes, and the
`path_tracer` module keeps the reference's names and edit semantics."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import pathed_capi as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pathed.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pt_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(P.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(P.EXPORTS) == names


@pytest.mark.parametrize("pool", [0, 4096, 5000, 1_000_000, 1_003_455, (1 << 20) + 1, 3 * 4096 + 64, 123_457])
def test_pool_holds_every_queue_slot(pool):
    """Every queue slot a pass's chunks address (camera runs dealt to 64 segments, linear
    queues, appended continuation / shadow / guided-record queues) lies inside the per-slot
    buffers the renderer allocates, for any pool setting: image sizes that fill a chunk
    exactly, leave a ragged last chunk, need one chunk per pixel run, or one pixel."""
    rng = np.random.default_rng(pool)
    shapes = [(1, 1, 1), (1024, 1024, 1), (1024, 1024, 64), (1000, 1000, 3), (777, 333, 5), (64, 64, 4), (4096, 4096, 1)]
    shapes += [(int(rng.integers(1, 3000)), int(rng.integers(1, 3000)), int(rng.integers(1, 300))) for _ in range(20)]
    for w, h, spp in shapes:
        cap, chunks, hi = P.pool_layout(w, h, spp, pool)
        assert hi < cap, (w, h, spp, pool, cap, hi)
        assert cap % (64 * 64) == 0 and chunks >= 1
        assert cap >= min(w * h * spp, 4096)
        if pool:
            # the configured capacity is honoured up to the segment rounding
            assert cap < max(pool, 4096) + 64 * 64


def test_pool_layout_rejects_bad_sizes():
    with pytest.raises(P.PtError):
        P.pool_layout(0, 10, 1)
    with pytest.raises(P.PtError):
        P.pool_layout(10, 10, 1, (1 << 27) + 1)


def test_abi_version_and_errors():
    L = P.lib()
    assert L.pt_abi_version() == 1
    r = ctypes.c_void_p()
    # invalid size is rejected before any device call
    assert L.pt_renderer_create(0, 10, ctypes.byref(r)) == -1
    assert b"bad arguments" in L.pt_last_error()
    assert L.pt_scene_preset(None, 0, 0, 0) == -1


def test_no_device_is_an_error_not_a_fallback():
    if P.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(P.PtError, match="no HIP device"):
        P.Renderer(8, 8)
    with pytest.raises(P.PtError, match="no HIP device"):
        P.path_uniforms(0, [0], [0], 2)


def test_cornell_preset():
    s = P.Scene.preset(0)
    i = s.info()
    assert (i.n_triangles, i.n_spheres, i.n_materials, i.n_emitters) == (12, 2, 6, 2)
    a = s.export()
    # walls face into the box: the flat normal points towards the box centre
    pos = a["tri_pos"].reshape(-1, 3, 3).astype(np.float64)
    n = np.cross(pos[:, 1] - pos[:, 0], pos[:, 2] - pos[:, 0])
    to_centre = np.array([0, 1, 0]) - pos.mean(1)
    assert np.all(np.sum(n[:10] * to_centre[:10], 1) > 0)
    # the light quad faces down
    assert np.all(n[10:, 1] < 0)
    assert np.allclose(a["mat_emit"][a["tri_mat"][10]], [17, 12, 4])
    assert a["mat_type"][a["sph_mat"][1]] == P.GLASS and a["mat_ior"][a["sph_mat"][1]] == pytest.approx(1.5)
    assert np.allclose(a["sph"][:, 3], 0.3)
    c = a["camera"]
    assert np.allclose(c["origin"], (0, 1, 6.8)) and c["vfov"] == pytest.approx(19.5 / 180 * math.pi)


def test_icosphere_preset_small_level():
    s = P.Scene.preset(1, seed=3, param=3)
    i = s.info()
    assert i.n_triangles == 20 * 4**3 and i.n_materials == i.n_triangles
    assert (i.env_width, i.env_height) == (1024, 512)
    a = s.export()
    v = a["tri_pos"].reshape(-1, 3)
    r = np.linalg.norm(v, axis=1)
    assert r.min() >= 0.95 - 1e-6 and r.max() <= 1.05 + 1e-6 and r.std() > 1e-3
    n = np.cross(a["tri_pos"][:, 3:6] - a["tri_pos"][:, 0:3], a["tri_pos"][:, 6:9] - a["tri_pos"][:, 0:3])
    assert np.mean(np.sum(n * a["tri_pos"][:, 0:3], 1) > 0) > 0.99  # outward winding
    lam = a["mat_type"] == P.LAMBERTIAN
    assert 0.2 < lam.mean() < 0.4
    al = a["mat_alpha"][~lam]
    assert al.min() >= 0.05 and al.max() <= 0.5
    env = a["env_rgba"]
    assert env.shape == (512, 1024, 4) and env[..., :3].max() > 5e3  # sun lobe
    lum = env[..., 0]
    body = lum[lum < 50]
    assert abs(np.median(np.log(body))) < 0.05 and abs(np.std(np.log(body)) - 1.0) < 0.05


def test_icosphere_full_level_counts():
    s = P.Scene.preset(1, seed=0)
    assert s.info().n_triangles == 1310720


def test_room_preset():
    s = P.Scene.preset(2, seed=1, param=200000)
    i = s.info()
    assert 150000 <= i.n_triangles <= 260000 and i.n_spheres == 1
    a = s.export()
    assert i.n_emitters == 2 and np.allclose(a["mat_emit"][a["tri_mat"][a["mat_emit"][a["tri_mat"]].sum(1) > 0][0]], 1000)
    assert (i.env_width, i.env_height) == (2048, 1024)
    clutter = a["mat_type"] == P.MICROFACET
    assert a["mat_alpha"][clutter].min() >= 0.02 and a["mat_alpha"][clutter].max() <= 0.3


def test_scene_building_api_validates():
    s = P.Scene()
    m = s.add_material(P.LAMBERTIAN, (0.5, 0.5, 0.5))
    with pytest.raises(P.PtError, match="material id"):
        s.add_triangles(np.zeros((1, 9)), [m + 1])
    with pytest.raises(P.PtError, match="radius"):
        s.add_spheres([[0, 0, 0]], [0.0], [m])
    with pytest.raises(P.PtError, match="up vector"):
        s.set_camera((0, 0, 0), (0, -1, 0), (0, 1, 0), 0.5)
    with pytest.raises(P.PtError, match="zero total"):
        s.set_environment(np.zeros((2, 4, 4)))
    s.add_triangles(np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]]), [m], normals=np.array([[0, 0, 1] * 3]))
    s.add_triangles(np.array([[0, 0, 1, 1, 0, 1, 0, 1, 1]]), [m])
    e = s.export()
    assert e["tri_nrm"].shape == (2, 9) and np.allclose(e["tri_nrm"][1], [0, 0, 1] * 3)


# ------------------------------------------------------------------------ path_tracer module
pt = pytest.importorskip("path_tracer")


def test_module_names_match_reference():
    for name in ["Vec3", "RenderState", "RenderSession", "RendererType", "MaterialType", "MaterialIndex", "SceneModel"]:
        assert hasattr(pt, name)
    for v in ["CUDA", "Optix", "Normals", "Lambertian", "Mirror", "Glass", "Microfacet", "None"]:
        assert hasattr(pt, v)
    methods = ["getRendererType", "setRendererType", "setColor", "getColor", "setEmit", "getEmit", "getIOR", "setIOR", "getAlpha",
               "setAlpha", "getMaterialID", "getMaterialType", "setMaterialType", "setLightPosition", "getLightPosition",
               "getCameraOrigin", "setCameraOrigin", "getCameraTarget", "setCameraTarget", "getCameraUp", "setCameraUp",
               "getMaxDepth", "setMaxDepth", "getNextEventEstimation", "setNextEventEstimation", "zoomCamera", "getSpp"]
    for mth in methods:
        assert hasattr(pt.SceneModel, mth), mth
    for mth in ["init", "renderAsync", "pollRender", "hitTest", "getWidth", "getHeight", "getSceneModel"]:
        assert hasattr(pt.RenderSession, mth), mth


def test_scene_model_semantics_without_gpu():
    s = pt.RenderSession(40, 30)
    assert (s.getWidth(), s.getHeight()) == (40, 30)
    m = s.getSceneModel()
    # scene_model.cu: no picked material -> neutral getters
    assert m.getMaterialID() == -1 and m.getMaterialType() == getattr(pt.MaterialType, "None")
    assert m.getIOR() == -1.0 and m.getAlpha() == -1.0 and m.getColor().r() == 0.0
    # scene.h defaults: depth 3, NEE on; light position -0.6 (stored only)
    assert m.getMaxDepth() == 3 and m.getNextEventEstimation() is True
    assert m.getLightPosition() == pytest.approx(-0.6)
    m.setMaxDepth(5)
    m.setNextEventEstimation(False)
    assert m.getMaxDepth() == 5 and m.getNextEventEstimation() is False
    o = m.getCameraOrigin()
    assert (o.x(), o.y(), o.z()) == pytest.approx((0, 1, 6.8))
    m.zoomCamera(1.0)
    o = m.getCameraOrigin()
    assert (o.x(), o.y(), o.z()) == pytest.approx((0, 1, 5.8))
    m.setCameraTarget(0.0, 1.5, 0.0)
    assert m.getCameraTarget().y() == pytest.approx(1.5)
    assert m.getRendererType() == pt.RendererType.Optix
    m.setRendererType(pt.RendererType.CUDA)
    assert m.getRendererType() == pt.RendererType.CUDA
    # renderer type changes do not touch other state (render_session.cu:83-87)
    assert m.getMaxDepth() == 5


def test_material_edits_follow_reference_rules():
    # without a GPU the pick is simulated through the material id of the scene model
    s = pt.RenderSession(16, 16)
    m = s.getSceneModel()
    with pytest.raises(RuntimeError):
        s.hitTest(8, 8)
    with pytest.raises(RuntimeError):
        s.renderAsync()
    st = s.pollRender()
    assert st.isRendering is False


def test_env_tables_noalloc_rejects_bad_sizes_on_the_host():
    """The allocation-free table build refuses maps beyond its fixed scratch, and NULL pointers,
    before it touches a device (the pointers here are never dereferenced)."""
    L = P.lib()
    fake = ctypes.c_void_p(0x1000)
    assert L.pt_env_tables_noalloc(fake, 4096, 1024, 0, fake, fake, fake, None) == P.ERR_INVALID
    assert b"exceeds the fixed scratch" in L.pt_last_error()
    assert L.pt_env_tables_noalloc(fake, 1, (1 << 12) + 1, 0, fake, fake, fake, None) == P.ERR_INVALID
    assert L.pt_env_tables_noalloc(None, 4, 4, 0, fake, fake, fake, None) == P.ERR_INVALID
    assert L.pt_env_tables_noalloc(fake, 0, 4, 0, fake, fake, fake, None) == P.ERR_INVALID
