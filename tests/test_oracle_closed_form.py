"""Closed-form and statistical pins of the oracle: the parts of the reference with no
known-answer test of their own (BSDFs, Fresnel/Snell, intersection, camera, light sampling,
alias tables, framebuffer, integrator)."""
import math

import numpy as np
import pytest
from scipy import stats

import oracle as O
import os

from helpers import one_texel_plane_scene, sphere_light_scene, srgb_sweep

LAMB, MIRROR, GLASS, MICRO = 0, 1, 2, 3


def mat(t, albedo=(0.5, 0.5, 0.5), ior=1.5, alpha=0.3):
    return [t, albedo[0], albedo[1], albedo[2], ior, alpha]


def sph_dirs(rng, n, upper=False):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    if upper:
        v[:, 2] = np.abs(v[:, 2])
    return v.astype(np.float32)


# ----------------------------------------------------------------------------- Fresnel/Snell
def test_fresnel_normal_incidence():
    for eta in (1.33, 1.5, 2.4):
        assert O.fresnel(1.0, 1.0, eta) == pytest.approx(((eta - 1) / (eta + 1)) ** 2, rel=1e-12)


def test_fresnel_tir_and_grazing():
    # from glass to air beyond the critical angle: total internal reflection
    crit = math.asin(1 / 1.5)
    assert O.fresnel(math.cos(crit + 0.05), 1.5, 1.0) == 1.0
    assert O.fresnel(1e-9, 1.0, 1.5) == pytest.approx(1.0, abs=1e-6)


def test_fresnel_matches_textbook_formula():
    # pbrt 3e eq. 8.3: r_par, r_perp from Snell's law
    for cos_i in np.linspace(0.05, 1, 20):
        ei, et = 1.0, 1.5
        sin_t = ei / et * math.sqrt(1 - cos_i * cos_i)
        cos_t = math.sqrt(1 - sin_t * sin_t)
        rpar = (et * cos_i - ei * cos_t) / (et * cos_i + ei * cos_t)
        rper = (ei * cos_i - et * cos_t) / (ei * cos_i + et * cos_t)
        assert O.fresnel(cos_i, ei, et) == pytest.approx(0.5 * (rpar**2 + rper**2), rel=1e-12)


# ----------------------------------------------------------------------------- BSDFs
def test_lambert_eval_and_pdf():
    rng = np.random.default_rng(1)
    wo = sph_dirs(rng, 200, True)
    wi = sph_dirs(rng, 200)
    m = np.array([mat(LAMB, (0.2, 0.5, 0.9))] * 200, np.float32)
    f, pdf = O.bsdf_eval(m, wo, wi, 64)
    up = wi[:, 2] >= 0
    assert np.allclose(f[up], np.array([0.2, 0.5, 0.9]) / np.pi, rtol=1e-6)
    assert np.all(f[~up] == 0)
    assert np.allclose(pdf[up], 1 / (2 * np.pi)) and np.all(pdf[~up] == 0)


def test_lambert_sample_uniform_hemisphere():
    """lambertian.h:67-76 uses uniform (not cosine) hemisphere sampling: z = xi1."""
    rng = np.random.default_rng(2)
    n = 20000
    xi = rng.random((n, 2)).astype(np.float32)
    m = np.array([mat(LAMB)] * n, np.float32)
    wo = np.tile([[0, 0, 1]], (n, 1)).astype(np.float32)
    s = O.bsdf_sample(m, wo, xi, 64)
    assert np.allclose(s[:, 2], xi[:, 0], atol=1e-7)
    assert np.allclose(np.linalg.norm(s[:, :3], axis=1), 1, atol=1e-6)
    assert stats.kstest(s[:, 2], "uniform").pvalue > 1e-3
    assert np.allclose(s[:, 3], 1 / (2 * np.pi))


def test_mirror_sample_throughput_one():
    rng = np.random.default_rng(3)
    n = 500
    wo = sph_dirs(rng, n, True)
    m = np.array([mat(MIRROR)] * n, np.float32)
    s = O.bsdf_sample(m, wo, np.zeros((n, 2), np.float32), 64)
    wi = s[:, :3]
    assert np.allclose(wi[:, :2], -wo[:, :2], atol=1e-7) and np.allclose(wi[:, 2], wo[:, 2], atol=1e-7)
    assert np.allclose(s[:, 4] * np.abs(wi[:, 2]) / s[:, 3], 1.0, rtol=1e-12)
    assert np.all(s[:, 7] == 1)
    f, pdf = O.bsdf_eval(m, wo, wi.astype(np.float32), 64)
    assert np.all(f == 0) and np.all(pdf == 0)


def test_glass_entering_snell_and_energy():
    """Entering the dielectric (wo.z > 0): refracted direction obeys Snell's law; the
    reflect/refract pdfs are F and 1-F; f*cos/pdf is 1 (reflection) or 1/eta^2."""
    rng = np.random.default_rng(4)
    n = 300
    wo = sph_dirs(rng, n, True)
    wo[:, 2] = np.maximum(wo[:, 2], 0.05)
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    m = np.array([mat(GLASS, ior=1.5)] * n, np.float32)
    xi_t = np.full((n, 2), 0.9999, np.float32)
    s = O.bsdf_sample(m, wo, xi_t, 64)
    wi = s[:, :3]
    sin_o = np.sqrt(1 - wo[:, 2].astype(np.float64) ** 2)
    sin_i = np.sqrt(1 - wi[:, 2] ** 2)
    F = np.array([O.fresnel(float(c), 1.0, 1.5) for c in wo[:, 2]])
    refr = s[:, 3] == pytest.approx(1 - F)  # noqa
    assert np.allclose(s[:, 3], 1 - F, rtol=1e-9)
    assert np.allclose(1.0 * sin_o, 1.5 * sin_i, rtol=1e-5, atol=1e-6)
    assert np.all(wi[:, 2] < 0)
    assert np.allclose(s[:, 4] * np.abs(wi[:, 2]) / s[:, 3], (1 / 1.5) ** 2, rtol=1e-9)
    s0 = O.bsdf_sample(m, wo, np.zeros((n, 2), np.float32), 64)
    assert np.allclose(s0[:, 3], F, rtol=1e-9)
    assert np.allclose(s0[:, 4] * np.abs(s0[:, 2]) / s0[:, 3], 1.0, rtol=1e-9)


def test_glass_leaving_quirk_and_tir():
    """Leaving the dielectric (wo.z < 0): Fresnel uses eta 1.5 -> 1 (TIR beyond the critical
    angle), while the refracted direction goes through the reference's second eta swap in
    snell.h:17-23 and bends with ratio 1/ior (kept for parity)."""
    cos_o = 0.3  # beyond the critical angle for 1.5 -> 1 (cos_c = 0.745)
    wo = np.array([[math.sqrt(1 - cos_o**2), 0, -cos_o]], np.float32)
    m = np.array([mat(GLASS, ior=1.5)], np.float32)
    s = O.bsdf_sample(m, wo, np.array([[0.99999, 0]], np.float32), 64)
    assert s[0, 3] == 1.0 and s[0, 2] < 0  # TIR: reflected back inside, pdf 1
    cos_o = 0.95
    wo = np.array([[math.sqrt(1 - cos_o**2), 0, -cos_o]], np.float32)
    s = O.bsdf_sample(m, wo, np.array([[0.99999, 0]], np.float32), 64)
    sin_o = math.sqrt(1 - float(wo[0, 2]) ** 2)
    sin_i = math.sqrt(1 - s[0, 2] ** 2)
    assert s[0, 2] > 0 and sin_i == pytest.approx(sin_o / 1.5, rel=1e-5)
    F = O.fresnel(float(-wo[0, 2]), 1.5, 1.0)
    assert s[0, 3] == pytest.approx(1 - F, rel=1e-9)
    assert s[0, 4] * abs(s[0, 2]) / s[0, 3] == pytest.approx(1.5**2, rel=1e-9)


def _beckmann_ref(alpha, wo, wi):
    """Independent evaluation of microfacet.h:98-134 from the published formulas."""
    wo = np.asarray(wo, np.float64)
    wi = np.asarray(wi, np.float64)
    if wo[2] < 0 or wi[2] < 0:
        return 0.0, 0.0
    wh = wo + wi
    wh /= np.linalg.norm(wh)

    def alpha_w(c):
        return (1.2 - 0.2 * math.sqrt(abs(c))) * alpha

    def D(wv):
        a = alpha_w(wv[2])
        c2 = wh[2] ** 2
        t2 = (1 - c2) / c2
        return math.exp(-t2 / a**2) / (math.pi * a**2 * c2 * c2)

    def G1(v):
        tan = math.sqrt(max(0.0, 1 - v[2] ** 2)) / v[2]
        a = 1 / (alpha * abs(tan))
        if a >= 1.6:
            return 1.0
        return (3.535 * a + 2.181 * a * a) / (1 + 2.276 * a + 2.577 * a * a)

    G = G1(wi) * G1(wo) if (np.dot(wo, wh) >= 0) == (wo[2] >= 0) and (np.dot(wi, wh) >= 0) == (wi[2] >= 0) else 0.0
    F = O.fresnel(min(1.0, max(0.0, float(np.dot(wi, wh)))), 1.0, 1.5)
    f = D(wi) * G * F / (4 * wi[2] * wo[2])
    pdf = D(wi) * abs(wh[2]) / (4 * abs(np.dot(wo, wh)))
    return f, pdf


def test_microfacet_eval_matches_published_formulas():
    rng = np.random.default_rng(5)
    n = 400
    wo = sph_dirs(rng, n, True)
    wi = sph_dirs(rng, n, True)
    alphas = rng.uniform(0.05, 0.5, n).astype(np.float32)
    m = np.array([mat(MICRO, alpha=a) for a in alphas], np.float32)
    f, pdf = O.bsdf_eval(m, wo, wi, 64)
    for k in range(n):
        fr, pr = _beckmann_ref(float(alphas[k]), wo[k], wi[k])
        assert f[k, 0] == pytest.approx(fr, rel=1e-9, abs=1e-300)
        assert pdf[k] == pytest.approx(pr, rel=1e-9, abs=1e-300)
    assert np.all(f[:, 0] == f[:, 1]) and np.all(f[:, 1] == f[:, 2])


def test_walter_g1_close_to_exact_smith():
    """The G1 constants (microfacet.h:74-84) are Walter et al. 2007's rational fit to the exact
    Beckmann-Smith G1 = 2 / (1 + erf(a) + exp(-a^2) / (a sqrt(pi))); the fit is within 0.4%.
    With wi = wo (so wh = wo), f = D G1(wo)^2 F / (4 cos^2), which isolates G1."""
    alpha = 0.3
    errs = []
    for theta in np.linspace(0.2, 1.45, 30):
        w = np.array([[math.sin(theta), 0, math.cos(theta)]], np.float32)
        m = np.array([mat(MICRO, alpha=alpha)], np.float32)
        f, _ = O.bsdf_eval(m, w, w, 64)
        wd = w[0].astype(np.float64)
        c = wd[2]
        aw = (1.2 - 0.2 * math.sqrt(c)) * alpha
        D = math.exp(-((1 - c * c) / (c * c)) / aw**2) / (math.pi * aw**2 * c**4)
        F = O.fresnel(1.0, 1.0, 1.5)
        g1 = math.sqrt(f[0, 0] * 4 * c * c / (D * F))
        a = 1 / (alpha * math.sqrt(1 - c * c) / c)
        exact = 1.0 if a > 8 else 2 / (1 + math.erf(a) + math.exp(-a * a) / (a * math.sqrt(math.pi)))
        errs.append(abs(g1 - exact) / exact)
    assert max(errs) < 4e-3, max(errs)


@pytest.mark.parametrize("alpha,cos_o", [(0.1, 0.9), (0.3, 0.5), (0.5, 0.2)])
def test_microfacet_sample_matches_its_pdf(alpha, cos_o):
    """microfacet.h:53-63, 148-158: half vectors are Beckmann-distributed with the widened
    alpha a' = (1.2 - 0.2 sqrt|cos wo|) alpha (tan^2 / a'^2 ~ Exp(1), phi uniform), D with
    that alpha integrates to one against cos(theta_h), and the returned pdf is the
    half-vector density times the reflection Jacobian 1 / (4 |wi.wh|) (Walter 2007 eq. 14)."""
    rng = np.random.default_rng(6)
    n = 100000
    wo1 = np.array([math.sqrt(1 - cos_o**2), 0, cos_o])
    wo = np.tile(wo1, (n, 1)).astype(np.float32)
    m = np.array([mat(MICRO, alpha=alpha)] * n, np.float32)
    xi = rng.random((n, 2)).astype(np.float32)
    s = O.bsdf_sample(m, wo, xi, 64)
    wi = s[:, :3]
    wh = wi + wo.astype(np.float64)
    wh /= np.linalg.norm(wh, axis=1, keepdims=True)
    wh[wh[:, 2] < 0] *= -1
    a = (1.2 - 0.2 * math.sqrt(cos_o)) * alpha
    t2 = (1 - wh[:, 2] ** 2) / wh[:, 2] ** 2
    assert stats.kstest(t2 / a**2, "expon").pvalue > 1e-3
    assert stats.kstest((np.arctan2(wh[:, 1], wh[:, 0]) + np.pi) / (2 * np.pi), "uniform").pvalue > 1e-3
    D = np.exp(-t2 / a**2) / (np.pi * a**2 * wh[:, 2] ** 4)
    good = np.abs(np.sum(wi * wh, 1)) > 1e-3
    assert np.allclose(s[good, 3] * 4 * np.abs(np.sum(wi * wh, 1))[good], (D * wh[:, 2])[good], rtol=1e-6)
    th = (np.arange(200000) + 0.5) / 200000 * (np.pi / 2)
    c = np.cos(th)
    Dn = np.exp(-np.tan(th) ** 2 / a**2) / (np.pi * a**2 * c**4)
    assert np.sum(Dn * c * np.sin(th)) * (np.pi / 2 / 200000) * 2 * np.pi == pytest.approx(1.0, rel=1e-6)


@pytest.mark.parametrize("alpha,cos_o", [(0.1, 0.9), (0.3, 0.5), (0.5, 0.2)])
def test_microfacet_sampling_density(alpha, cos_o):
    """The density the consistent-MIS weights use: at the directions sample() draws it equals
    the pdf sample() returns, it integrates to one over the sphere (sample() may reflect below
    the horizon), and it differs from pdf(), which widens alpha at |cos wi|."""
    rng = np.random.default_rng(7)
    n = 20000
    wo1 = np.array([math.sqrt(1 - cos_o**2), 0, cos_o])
    wo = np.tile(wo1, (n, 1)).astype(np.float32)
    m = np.array([mat(MICRO, alpha=alpha)] * n, np.float32)
    s = O.bsdf_sample(m, wo, rng.random((n, 2)).astype(np.float32), 64)
    _, ps = O.bsdf_eval(m, wo, s[:, :3], precision=64, sample_pdf=True)
    # the half vector is ill-conditioned where wi approaches -wo (|wo + wi| -> 0)
    ok = (s[:, 3] > 1e-6) & (np.linalg.norm(s[:, :3] + wo1, axis=1) > 0.05)
    assert np.allclose(ps[ok], s[ok, 3], rtol=1e-4)
    # integral over the sphere by stratified uniform directions: E[4 pi p(w)] = 1
    k = 1000
    u = (np.arange(k)[:, None] + rng.random((k, 400))) / k
    z = (1 - 2 * u).reshape(-1)
    ph = rng.random(z.size) * 2 * np.pi
    r = np.sqrt(1 - z * z)
    w = np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)
    mm = np.array([mat(MICRO, alpha=alpha)] * z.size, np.float32)
    wwo = np.tile(wo1, (z.size, 1))
    _, pa = O.bsdf_eval(mm, wwo, w, precision=64, sample_pdf=True)
    integral = 4 * np.pi * pa.mean()
    se = 4 * np.pi * pa.std() / math.sqrt(pa.size)
    assert abs(integral - 1) < max(5 * se, 2e-3), (integral, se)
    _, pr = O.bsdf_eval(m, wo, s[:, :3], precision=64)
    assert np.abs(pr[ok] / ps[ok] - 1).max() > 1e-3


def test_consistent_mis_matches_naive_on_glossy_floor():
    """With the light-sample MIS weight taken from the sampling density the weights of the two
    techniques sum to one, so NEE+MIS converges to the naive estimator's image on a Beckmann
    floor lit by a small quad."""
    tris = []
    f = np.array([(-2, 0, -2), (2, 0, -2), (2, 0, 2), (-2, 0, 2)], np.float64)
    tris += [np.concatenate([f[0], f[2], f[1]]), np.concatenate([f[0], f[3], f[2]])]
    L = np.array([(-0.25, 1.0, -0.25), (0.25, 1.0, -0.25), (0.25, 1.0, 0.25), (-0.25, 1.0, 0.25)])
    tris += [np.concatenate([L[0], L[1], L[2]]), np.concatenate([L[0], L[2], L[3]])]
    arr = _scene_arrays(tris=np.array(tris),
                        mats=[dict(type=3, alpha=0.3), dict(type=0, albedo=(0.5, 0.5, 0.5), emit=(8, 8, 8))],
                        cam=dict(origin=(0, 1.2, 3.0), target=(0, 0, 0), up=(0, 1, 0), vfov=0.6))
    arr["tri_mat"] = np.array([0, 0, 1, 1], np.int32)
    s = O.Scene(arr, 16, 16)
    spp = 256
    r = {}
    for nee, mis in ((0, 0), (1, 1)):
        tot, sq, _ = s.render(want_sq=True, spp=spp, max_depth=3, nee=nee, mode=1, tri_test=1, precision=64, mis_sample=mis)
        r[nee] = (tot / spp, (sq / spp - (tot / spp) ** 2) / spp)
    d = r[1][0].sum(axis=(0, 1)) - r[0][0].sum(axis=(0, 1))
    se = np.sqrt(r[1][1].sum(axis=(0, 1)) + r[0][1].sum(axis=(0, 1)))
    assert np.all(np.abs(d) < 4.5 * se), (d, se)
    assert r[0][0].sum() > 0


# ----------------------------------------------------------------------------- intersection
def _tri_rays(rng, n):
    tris = rng.uniform(-1, 1, (n, 9)).astype(np.float32)
    o = rng.uniform(-3, 3, (n, 3)).astype(np.float32)
    # aim at a random barycentric point inside/outside
    b = rng.uniform(-0.2, 1.2, (n, 2))
    tgt = tris[:, 0:3] * (1 - b[:, :1] - b[:, 1:]) + tris[:, 3:6] * b[:, :1] + tris[:, 6:9] * b[:, 1:]
    d = (tgt - o).astype(np.float32)
    rays = np.concatenate([o, np.zeros((n, 1)), d, np.full((n, 1), 1e30)], 1).astype(np.float32)
    return tris, rays, b


def test_triangle_tests_closed_form():
    rng = np.random.default_rng(7)
    n = 20000
    tris, rays, b = _tri_rays(rng, n)
    inside = (b[:, 0] > 1e-3) & (b[:, 1] > 1e-3) & (b.sum(1) < 1 - 1e-3)
    outside = (b[:, 0] < -1e-3) | (b[:, 1] < -1e-3) | (b.sum(1) > 1 + 1e-3)
    t64 = tris.astype(np.float64)
    nrm = np.cross(t64[:, 3:6] - t64[:, 0:3], t64[:, 6:9] - t64[:, 0:3])
    d64 = rays[:, 4:7].astype(np.float64)
    cosang = np.abs(np.sum(nrm * d64, 1)) / (np.linalg.norm(nrm, axis=1) * np.linalg.norm(d64, axis=1))
    well = (cosang > 0.2) & (np.linalg.norm(nrm, axis=1) > 0.05)
    for mode in (0, 1):
        out = O.tri_test(tris, rays, mode)
        hit = out[:, 0] == 1
        assert np.all(hit[inside]) and not np.any(hit[outside])
        # the target point was built at t = 1 with barycentrics b
        sel = inside & well
        assert np.allclose(out[sel, 1], 1.0, atol=2e-4)
        assert np.allclose(out[sel, 2], b[sel, 0], atol=2e-3)
        assert np.allclose(out[sel, 3], b[sel, 1], atol=2e-3)


def test_watertight_agrees_with_mt_off_edges():
    rng = np.random.default_rng(8)
    tris, rays, b = _tri_rays(rng, 50000)
    m = np.minimum(np.minimum(b[:, 0], b[:, 1]), 1 - b.sum(1))
    far = np.abs(m) > 1e-6
    a = O.tri_test(tris, rays, 0)
    w = O.tri_test(tris, rays, 1)
    assert np.array_equal(a[far, 0], w[far, 0])


def test_watertight_shared_edge_no_leak():
    """Rays through the shared edge of two triangles hit at least one of them (watertight)."""
    rng = np.random.default_rng(9)
    n = 20000
    p0 = np.array([-1, -1, 0.0]); p1 = np.array([1, -1, 0.0]); p2 = np.array([1, 1, 0.0]); p3 = np.array([-1, 1, 0.0])
    s = rng.uniform(0, 1, n)
    pts = p0 + (p2 - p0) * s[:, None]
    o = pts + rng.normal(size=(n, 3)) * np.array([0.3, 0.3, 0]) + np.array([0, 0, 2.0])
    d = pts - o
    rays = np.concatenate([o, np.zeros((n, 1)), d, np.full((n, 1), 1e30)], 1).astype(np.float32)
    t1 = np.tile(np.concatenate([p0, p1, p2]), (n, 1)).astype(np.float32)
    t2 = np.tile(np.concatenate([p0, p2, p3]), (n, 1)).astype(np.float32)
    w1 = O.tri_test(t1, rays, 1)[:, 0]
    w2 = O.tri_test(t2, rays, 1)[:, 0]
    assert np.all((w1 + w2) >= 1)


def _scene_arrays(tris=None, sph=None, mats=None, env=None, cam=None):
    tris = np.zeros((0, 9), np.float32) if tris is None else np.asarray(tris, np.float32)
    mats = mats or [dict(type=0, albedo=(0.5, 0.5, 0.5), emit=(0, 0, 0))]
    return dict(
        tri_pos=tris.reshape(-1, 9), tri_nrm=None, tri_mat=np.zeros(len(tris), np.int32) if len(tris) else np.zeros(0, np.int32),
        sph=np.asarray(sph if sph is not None else np.zeros((0, 4)), np.float32), sph_mat=np.zeros(max(1, 0 if sph is None else len(sph)), np.int32),
        mat_type=np.array([m["type"] for m in mats], np.int32),
        mat_albedo=np.array([m.get("albedo", (0, 0, 0)) for m in mats], np.float32),
        mat_emit=np.array([m.get("emit", (0, 0, 0)) for m in mats], np.float32),
        mat_ior=np.array([m.get("ior", 1.5) for m in mats], np.float32),
        mat_alpha=np.array([m.get("alpha", 0.1) for m in mats], np.float32),
        env_rgba=env,
        camera=cam or dict(origin=(0, 0, 5), target=(0, 0, 0), up=(0, 1, 0), vfov=0.6, flip=0),
    )


def test_sphere_both_roots():
    a = _scene_arrays(sph=[[0, 0, 0, 1]])
    s = O.Scene(a, 4, 4)
    rays = np.array([[0, 0, 5, 0, 0, 0, -1, 1e30],   # outside: near root t = 4
                     [0, 0, 0, 0, 0, 0, -1, 1e30],   # inside: far root t = 1
                     [0, 0, 5, 4.5, 0, 0, -1, 1e30],  # tmin past the near root: far root t = 6
                     [0, 2, 5, 0, 0, 0, -1, 1e30]], np.float32)  # miss
    t, u, v, prim = s.intersect(rays, 1)
    assert t[0] == pytest.approx(4) and t[1] == pytest.approx(1) and t[2] == pytest.approx(6)
    assert list(prim[:3]) == [-1, -1, -1] and prim[3] == np.iinfo(np.int32).min


def test_bvh_matches_brute_force():
    rng = np.random.default_rng(10)
    n = 3000
    c = rng.uniform(-2, 2, (n, 3))
    tris = np.concatenate([c + rng.normal(size=(n, 3)) * 0.2 for _ in range(3)], 1).astype(np.float32)
    a = _scene_arrays(tris=tris, sph=[[0, 0, 0, 0.5], [1, 1, 1, 0.3]])
    a["sph_mat"] = np.zeros(2, np.int32)
    sb = O.Scene(a, 4, 4, bvh=False)
    sv = O.Scene(a, 4, 4, bvh=True)
    m = 4000
    o = rng.uniform(-4, 4, (m, 3))
    d = rng.normal(size=(m, 3))
    rays = np.concatenate([o, np.zeros((m, 1)), d, np.full((m, 1), 1e30)], 1).astype(np.float32)
    for mode in (0, 1):
        tb, _, _, pb = sb.intersect(rays, mode)
        tv, _, _, pv = sv.intersect(rays, mode)
        assert np.array_equal(pb == np.iinfo(np.int32).min, pv == np.iinfo(np.int32).min)
        hit = pb != np.iinfo(np.int32).min
        assert np.array_equal(tb[hit], tv[hit])


# ----------------------------------------------------------------------------- camera
def test_camera_centre_ray_points_at_target():
    cam = dict(origin=(0.0, 1.0, 6.8), target=(0.0, 1.0, 0.0), up=(0.0, 1.0, 0.0), vfov=19.5 / 180 * math.pi)
    W, H = 256, 256
    # the image centre is the corner between pixels (127,127)...(128,128): col=127, sx=1
    out = O.camera_rays(cam, W, H, [[127, 127, 1.0, 1.0], [0, 0, 0.0, 0.0], [H - 1, W - 1, 1.0, 1.0]])
    assert np.allclose(out[0, :3], cam["origin"])
    assert np.allclose(out[0, 3:], [0, 0, -1], atol=1e-6)
    # bottom-left corner (row 0 = bottom): direction tilted by the half fov both ways
    t = math.tan(cam["vfov"] / 2)
    dl = np.array([-t, -t, -1.0]) / math.sqrt(2 * t * t + 1)
    assert np.allclose(out[1, 3:], dl, atol=1e-6)
    assert np.allclose(out[2, 3:], [t / math.sqrt(2 * t * t + 1), t / math.sqrt(2 * t * t + 1), -1 / math.sqrt(2 * t * t + 1)], atol=1e-6)


def test_camera_flip_handedness_mirrors_x():
    cam = dict(origin=(0.0, 2.0, 15.0), target=(0.0, 1.69521, 14.0476), up=(0.0, 0.952421, -0.304787), vfov=0.5)
    a = O.camera_rays(dict(cam, flip=0), 64, 64, [[10, 3, 0.5, 0.5]])
    b = O.camera_rays(dict(cam, flip=1), 64, 64, [[10, 60, 0.5, 0.5]])
    assert np.allclose(a[0, 3:], b[0, 3:], atol=1e-5)


# ----------------------------------------------------------------------------- alias tables
def _alias_pmf(prob, alias):
    n = prob.size
    pmf = prob / n
    np.add.at(pmf, alias, (1 - prob) / n)
    return pmf


@pytest.mark.parametrize("seed", range(6))
def test_alias_pmf_equals_weights(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 3000))
    w = rng.lognormal(0, 2, n)
    w[rng.random(n) < 0.2] = 0
    if seed == 0:
        w[:] = 1.0
    if seed == 1:
        w[:] = 0; w[n // 2] = 5.0
    prob, alias, s = O.alias_build(w)
    assert s == pytest.approx(w.sum(), rel=1e-12)
    assert np.all((prob >= 0) & (prob <= 1 + 1e-12)) and np.all((alias >= 0) & (alias < n))
    assert np.allclose(_alias_pmf(prob, alias), w / w.sum(), rtol=1e-9, atol=1e-12)


def test_env_alias_pmf_and_pdf_match_reference_cmf():
    """The product's FP64 alias tables give the same texel pmf (and so the same solid-angle
    pdf) as the reference's FP32 CMF tables (distribution.cu)."""
    rng = np.random.default_rng(11)
    h, w = 32, 64
    rgba = np.zeros((h, w, 4), np.float32)
    rgba[..., :3] = rng.lognormal(0, 1, (h, w, 3))
    rp, ra, mp, ma, tp = O.env_alias(rgba)
    pmf = _alias_pmf(mp, ma)[:, None] * np.stack([_alias_pmf(rp[y], ra[y]) for y in range(h)])
    wsum = rgba[..., :3].astype(np.float64).sum(-1)
    assert np.allclose(pmf, wsum / wsum.sum(), rtol=1e-9)
    ref_pmf, _, _ = O.dist2d(rgba[..., :3].sum(-1), w, h)
    assert np.allclose(pmf, ref_pmf, rtol=2e-4, atol=1e-8)
    dirs = sph_dirs(rng, 5000)
    assert np.allclose(O.env_pdf(rgba, dirs, ref=False), O.env_pdf(rgba, dirs, ref=True), rtol=3e-4)


def test_env_sampling_matches_pdf():
    """Alias sampling with continuous in-texel offsets: E[g(w)/p(w)] = integral of g."""
    rng = np.random.default_rng(12)
    h, w = 16, 32
    rgba = np.zeros((h, w, 4), np.float32)
    rgba[..., :3] = rng.lognormal(0, 1, (h, w, 3))
    n = 400000
    d, pdf, em = O.env_sample(rgba, rng.random((n, 2)), ref=False)
    # integral of the map's luminance over the sphere, computed per texel analytically
    th0 = np.arange(h) / h * np.pi
    th1 = (np.arange(h) + 1) / h * np.pi
    area = (np.cos(th0) - np.cos(th1)) * 2 * np.pi / w
    exact = (rgba[..., 0].astype(np.float64) * area[:, None]).sum()
    # a sample that lands exactly on a pole has density 0 and contributes nothing; next to
    # the poles (y rounded to +-1) the density stays finite
    zero = pdf == 0
    assert zero.sum() <= 2 and np.all(np.abs(d[zero, 1]) == 1.0)
    assert np.isfinite(pdf).all()
    est = np.mean(np.where(zero, 0.0, em[:, 0] / np.where(zero, 1.0, pdf)))
    assert est == pytest.approx(exact, rel=0.01)
    # samples land on the texel whose emission they return
    assert np.allclose(pdf, O.env_pdf(rgba, d, ref=False))


# ----------------------------------------------------------------------------- framebuffer
def test_framebuffer_running_mean():
    rng = np.random.default_rng(13)
    passes = [rng.random(30).astype(np.float32) for _ in range(5)]
    spp = [1, 3, 2, 7, 4]
    mean = np.zeros(30, np.float32)
    prev = 0
    for p, s in zip(passes, spp):
        mean, rgba = O.framebuffer_update(p, mean, s, prev)
        prev += s
    exact = sum(p.astype(np.float64) * s for p, s in zip(passes, spp)) / sum(spp)
    assert np.allclose(mean, exact, rtol=1e-5)
    v = mean.reshape(-1, 3)
    srgb = np.where(v <= 0.0031308, v * 12.92, 1.055 * np.power(v, 1 / 2.2) - 0.055)
    assert np.array_equal(rgba.reshape(-1, 4)[:, :3], (np.clip(srgb, 0, 1) * 255).astype(np.uint8))


def test_srgb_bytes_against_fp64_transfer():
    """color.h:7-28 + framebuffer.cu:35-40: the oracle's bytes equal the FP64 transfer
    function truncated, except where the FP64 value lies within float rounding of a byte
    boundary (there the float computation decides; both neighbours are allowed)."""
    v = srgb_sweep()
    _, rgba = O.framebuffer_update(np.repeat(v, 3), np.zeros(3 * v.size, np.float32), 1, 0)
    got = rgba.reshape(-1, 4)[:, 0].astype(np.int64)
    x = v.astype(np.float64)
    with np.errstate(invalid="ignore"):
        s = np.where(x <= np.float64(np.float32(0.0031308)), x * 12.92, 1.055 * np.power(x, np.float64(np.float32(1 / 2.2))) - 0.055)
    s = np.where(np.isnan(s), 1.0, s)
    want = np.floor(np.clip(s, 0, 1) * 255).astype(np.int64)
    frac = np.clip(s, 0, 1) * 255 - want
    near = (frac < 1e-4) | (frac > 1 - 1e-4)
    assert np.all((got == want) | (near & (np.abs(got - want) <= 1)))
    assert (got[:200001] == want[:200001]).mean() > 0.9999  # the uniform grid
    # fixed points of the transfer: black, white, saturation above 1, nan as saturated
    # radiance 1 gives byte 254: in float 1.055f - 0.055f rounds to 0.99999994
    assert set(got[v == 0]) == {0} and set(got[v == 1]) == {254} and set(got[v >= 1.001]) == {255}
    assert set(got[np.isnan(v)]) == {255} and set(got[np.isinf(v)]) == {255}
    # monotone on each segment; the hybrid (exponent 1/2.2 instead of 1/2.4) drops from byte
    # 10 to byte 5 where the linear segment hands over to the power segment
    t = np.float32(0.0031308)
    for seg in (np.isfinite(x) & (v <= t), np.isfinite(x) & (v > t)):
        order = np.argsort(x[seg], kind="stable")
        assert np.all(np.diff(got[seg][order]) >= 0)
    assert set(got[v == t]) == {10} and set(got[v == np.nextafter(t, np.float32(1))]) == {5}


# ----------------------------------------------------------------------------- sphere lights
@pytest.mark.parametrize("target", [(0.0, 0.0, 0.0), (0.6, 0.0, 0.5)])
@pytest.mark.parametrize("nee,mode,precision", [(1, 0, 64), (0, 0, 64), (1, 1, 32)])
def test_sphere_light_closed_form(target, nee, mode, precision):
    """sphere.h:30-44 (area sampling over the whole sphere, pdf 1 / (4 pi r^2)),
    sampler.h:20-61 / 89-125 (solid-angle conversion, pdfSceneLights of a BSDF-sampled sphere
    hit) and the integrator at one bounce: the floor's radiance under a sphere source equals
    rho le r^2 h / d^3.  NEE + MIS (light sample + BSDF sample weighted by the balance
    heuristic, both sphere pdfs involved), the naive estimator (BSDF sampling alone) and the
    product's merged schedule all converge to it.  Planted in a copy of the oracle, an area of
    2 pi r^2 moves the NEE estimate by -330 standard errors, and sampling only the z >= 0 half
    of the sphere (with the full-sphere pdf) by +260 at the off-axis target."""
    s, exact = sphere_light_scene(target)
    o = O.Scene(s.export(), 1, 1)
    n = 1 << 22
    tot, sq, _ = o.render_pixels([0], spp=n, max_depth=2, nee=nee, mode=mode, tri_test=mode, ref_env=1 - mode,
                                 precision=precision, seed=3)
    m = tot[0] / n
    se = np.sqrt(np.maximum(sq[0] / n - m * m, 0) / n)
    z = (m - exact) / se
    assert np.all(np.abs(z) < 5), (m, exact, z)
    assert np.all(se < 2e-3 * exact)


# ----------------------------------------------------------------------------- integrator
def _closed_box(albedo, emit, half=1.0):
    """An axis-aligned closed cube with inward-facing walls, all of one emissive Lambertian."""
    h = half
    v = np.array([[x, y, z] for x in (-h, h) for y in (-h, h) for z in (-h, h)], np.float64)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for a, b, c, d in quads:
        for t in ((a, b, c), (a, c, d)):
            p = v[list(t)]
            n = np.cross(p[1] - p[0], p[2] - p[0])
            if np.dot(n, -p.mean(0)) < 0:  # face inward
                p = p[[0, 2, 1]]
            tris.append(p.reshape(-1))
    return _scene_arrays(tris=np.array(tris), mats=[dict(type=0, albedo=albedo, emit=emit)],
                         cam=dict(origin=(0, 0, 0.5), target=(0, 0, 0), up=(0, 1, 0), vfov=1.0))


@pytest.mark.parametrize("nee", [0, 1])
@pytest.mark.parametrize("mode", [0, 1])
def test_white_furnace(nee, mode):
    """Closed box of albedo a and emission E seen through depth d: L = E * sum_{k<d} a^k.
    Both estimators (naive and NEE+MIS) must be unbiased for it."""
    a, E, depth = 0.7, 1.0, 5
    s = O.Scene(_closed_box((a, a, a), (E, E, E)), 8, 8)
    tot, sq, _ = s.render(want_sq=True, spp=64, max_depth=depth, nee=nee, mode=mode, tri_test=1, precision=64)
    spp = 64
    exact = E * sum(a**k for k in range(depth))
    mean = tot / spp
    m = mean.mean()
    var = (sq / spp - mean**2).mean() / (spp * 64)
    assert abs(m - exact) < 4 * math.sqrt(var) + 1e-9, (m, exact)


def test_nee_and_naive_agree_in_expectation():
    """NEE+MIS and the naive estimator converge to the same image (Cornell-like box with a
    small light and a glass and a diffuse sphere)."""
    from scipy.stats import norm  # noqa
    tris = []
    h = 1.0
    walls = [((-h, 0, -h), (h, 0, -h), (h, 0, h), (-h, 0, h)),
             ((-h, 2, -h), (-h, 2, h), (h, 2, h), (h, 2, -h)),
             ((-h, 0, -h), (-h, 2, -h), (h, 2, -h), (h, 0, -h)),
             ((-h, 0, -h), (-h, 0, h), (-h, 2, h), (-h, 2, -h)),
             ((h, 0, -h), (h, 2, -h), (h, 2, h), (h, 0, h))]
    for q in walls:
        q = np.array(q, np.float64)
        tris += [np.concatenate([q[0], q[2], q[1]]), np.concatenate([q[0], q[3], q[2]])]
    L = np.array([(-0.3, 1.98, -0.3), (0.3, 1.98, -0.3), (0.3, 1.98, 0.3), (-0.3, 1.98, 0.3)])
    tris += [np.concatenate([L[0], L[1], L[2]]), np.concatenate([L[0], L[2], L[3]])]
    arr = _scene_arrays(tris=np.array(tris), sph=[[-0.4, 0.3, -0.3, 0.3], [0.4, 0.3, 0.3, 0.3]],
                        mats=[dict(type=0, albedo=(0.75, 0.75, 0.75)), dict(type=0, albedo=(0.5, 0.5, 0.5), emit=(17, 12, 4)),
                              dict(type=2, ior=1.5)],
                        cam=dict(origin=(0, 1, 6.8), target=(0, 1, 0), up=(0, 1, 0), vfov=19.5 / 180 * math.pi))
    arr["tri_mat"] = np.array([0] * 10 + [1, 1], np.int32)
    arr["sph_mat"] = np.array([0, 2], np.int32)
    s = O.Scene(arr, 16, 16)
    spp = 256
    r = {}
    for nee in (0, 1):
        tot, sq, _ = s.render(want_sq=True, spp=spp, max_depth=4, nee=nee, mode=1, tri_test=1, precision=64)
        r[nee] = (tot / spp, (sq / spp - (tot / spp) ** 2) / spp)
    # compare the image means (sum over pixels) with their standard errors
    d = r[1][0].sum(axis=(0, 1)) - r[0][0].sum(axis=(0, 1))
    se = np.sqrt(r[1][1].sum(axis=(0, 1)) + r[0][1].sum(axis=(0, 1)))
    assert np.all(np.abs(d) < 4.5 * se), (d, se)


def test_one_texel_environment_closed_form():
    """The reference's scene 2: a unit-albedo plane lit by a single environment texel (the
    1_pixel_test.exr map, decoded from its PIZ compression).  Every pixel's radiance is
    albedo / pi times the texel's cosine-weighted solid angle times 1e4, for NEE + MIS (the
    alias table's in-texel sampling and pdf) and for the naive estimator (BSDF sampling that
    must land in the texel) alike, within 4 standard errors."""
    scenes = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")
    s, want = one_texel_plane_scene(scenes)
    W = H = 16
    o = O.Scene(s.export(), W, H)
    for nee, spp in ((1, 256), (0, 4096)):
        tot, sq, _ = o.render(want_sq=True, spp=spp, max_depth=3, nee=nee, mode=1, tri_test=1, ref_env=0, precision=64)
        m = tot / spp
        var = (sq / spp - m**2) / spp
        mean = m[..., 0].mean()
        se = np.sqrt(var[..., 0].sum()) / m[..., 0].size
        assert abs(mean - want) < 4 * se + 1e-9, (nee, mean, want, se)
        if nee:
            # NEE samples the one texel directly (the cosine varies by 7% across it): every
            # pixel is within 2% of the closed form
            assert np.all(np.abs(m[..., 0] - want) < 0.02 * want)
