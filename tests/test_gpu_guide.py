"""GPU parity of the guiding network's BF16 MFMA kernels against the numpy oracle."""
import math

import numpy as np
import pytest

import guide_oracle as G
import pathed_capi as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if P.device_count() == 0:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")


def test_mfma_lane_layout():
    """One v_mfma_f32_32x32x16_bf16 with the operand/result lane layout the kernels use."""
    print("case: mfma tile", flush=True)
    rng = np.random.default_rng(0)
    a = G.bf16(rng.normal(size=(32, 16)))
    b = G.bf16(rng.normal(size=(16, 32)))
    c = P.mfma_tile(a, b)
    assert np.allclose(c, a @ b, rtol=1e-5, atol=1e-5)


def test_mfma16_lane_layout():
    """One v_mfma_f32_16x16x32_bf16 with the operand/result lane layout of the inference kernel."""
    print("case: mfma 16x16x32 tile", flush=True)
    rng = np.random.default_rng(3)
    a = G.bf16(rng.normal(size=(16, 32)))
    b = G.bf16(rng.normal(size=(32, 16)))
    c = P.mfma_tile16(a, b)
    assert np.allclose(c, a @ b, rtol=1e-5, atol=1e-5)


def _params(rng):
    p = G.init_params(5)
    last = (G.LAYERS - 1) * G.LAYER_PARAMS
    Wl = p[last:last + 64 * 64].reshape(64, 64)
    Wl[:48] = rng.uniform(-0.3, 0.3, (48, 64))
    return p.astype(np.float32).astype(np.float64)


def test_forward_matches_oracle():
    print("case: guide forward", flush=True)
    rng = np.random.default_rng(1)
    p = _params(rng)
    feats = rng.random((1000, 8)).astype(np.float32)
    g = P.guide_forward(p, feats)
    ref, _ = G.forward(p, feats.astype(np.float64), emulate_bf16=True)
    ref32, _ = G.forward(p, feats.astype(np.float64))
    scale = np.abs(ref32).max()
    print(f"max |gpu - bf16 oracle| / scale {np.abs(g - ref).max() / scale:.2e}, "
          f"vs fp32 oracle {np.abs(g - ref32).max() / scale:.2e}", flush=True)
    # BF16 activations: within the oracle's own BF16 emulation up to accumulation order
    assert np.abs(g - ref).max() <= 3e-3 * scale  # measured 5.6e-4
    assert np.abs(g - ref32).max() <= 5e-2 * scale


def _decode_records(rec):
    """The 64-byte mixture record (16 words per row): octahedral unit means (two unorm16),
    FP16 kappa and lobe weights, the selection probability in lobe 7's weight slot; weights
    renormalised as the guided shade does."""
    rec = np.asarray(rec, np.uint32).reshape(-1, 16)
    oct_w, kw = rec[:, 0::2], rec[:, 1::2]
    u = (oct_w & 0xFFFF).astype(np.float64) * (2 / 65535) - 1
    v = (oct_w >> 16).astype(np.float64) * (2 / 65535) - 1
    n = np.stack([u, v, 1 - np.abs(u) - np.abs(v)], -1)
    t = np.maximum(-n[..., 2], 0)
    n[..., 0] -= np.where(n[..., 0] >= 0, t, -t)
    n[..., 1] -= np.where(n[..., 1] >= 0, t, -t)
    mu = n / np.linalg.norm(n, axis=-1, keepdims=True)
    kappa = (kw & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float64)
    hi = (kw >> 16).astype(np.uint16).view(np.float16).astype(np.float64)
    frac = hi[:, 7].copy()
    pi = hi.copy()
    pi[:, 7] = np.maximum(1 - pi[:, :7].sum(1), 0)
    pi /= pi.sum(1, keepdims=True)
    return mu, kappa, pi, frac


def test_mixture_records_match_the_network_outputs():
    """The inference kernel's mixture (its own epilogue, four lanes per row) and the record
    encoding, against the oracle's mixture of the same kernel's raw outputs: softmax weights,
    unit means, clamped concentrations and selection probability, to the record's FP16 and
    16-bit octahedral precision."""
    print("case: guide mixture records", flush=True)
    rng = np.random.default_rng(3)
    p = _params(rng)
    feats = rng.random((1000, 8)).astype(np.float32)
    raw = P.guide_forward(p, feats)
    rec = P.guide_records(p, feats)
    mu, kappa, pi, frac = _decode_records(rec)
    opi, omu, oml, okappa, _, _, oc = G.mixture(raw)
    ok = oml > 1e-6
    dmu = np.abs(mu - omu).max(-1)[ok].max()
    dk = (np.abs(kappa - okappa) / okappa).max()
    dpi = np.abs(pi - opi).max()
    dc = (np.abs(frac - oc) / oc).max()
    print(f"max |mean| {dmu:.2e}, kappa rel {dk:.2e}, weight {dpi:.2e}, selection rel {dc:.2e}", flush=True)
    assert dmu <= 1e-4 and dk <= 1e-3 and dpi <= 1e-3 and dc <= 1e-3
    # the encoding keeps the lobes in their own slots: the record's dominant lobe is the
    # network's wherever one lobe clearly dominates
    clear = np.sort(opi, 1)[:, -1] > np.sort(opi, 1)[:, -2] + 0.01
    assert (pi.argmax(1) == opi.argmax(1))[clear].all()


def test_gradients_match_oracle():
    print("case: guide gradients", flush=True)
    rng = np.random.default_rng(2)
    p = _params(rng)
    n = 4096
    feats = rng.random((n, 8))
    w = rng.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    pb = rng.uniform(0.05, 0.5, n)
    wt = rng.uniform(0.1, 3.0, n)
    rec = np.zeros((n, 16), np.float32)
    rec[:, :8] = feats
    rec[:, 8:11] = w
    rec[:, 11] = pb
    rec[:, 15] = wt
    loss, g = P.guide_gradients(p, rec)
    lref, gref = G.gradients(p, rec[:, :8].astype(np.float64), rec[:, 8:11].astype(np.float64), rec[:, 11].astype(np.float64),
                             rec[:, 15].astype(np.float64), emulate_bf16=True)
    assert loss == pytest.approx(lref, rel=1e-2)
    # per layer: relative L2 error of the gradient and cosine similarity
    for L in range(G.LAYERS):
        sl = slice(L * G.LAYER_PARAMS, (L + 1) * G.LAYER_PARAMS)
        a, b = g[sl].astype(np.float64), gref[sl]
        rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30)
        print(f"layer {L}: rel L2 {rel:.4f} cos {cos:.6f}", flush=True)
        assert rel < 1e-3 and cos > 0.9999  # measured about 1e-4 and 1.000000
    # and against the exact FP64 gradient
    _, g64 = G.gradients(p, rec[:, :8].astype(np.float64), rec[:, 8:11].astype(np.float64), rec[:, 11].astype(np.float64),
                         rec[:, 15].astype(np.float64))
    cos = g @ g64 / (np.linalg.norm(g) * np.linalg.norm(g64))
    assert cos > 0.995


def test_adam_matches_oracle():
    print("case: guide adam", flush=True)
    rng = np.random.default_rng(3)
    p = _params(rng)
    m = np.zeros(G.PARAMS)
    v = np.zeros(G.PARAMS)
    for step in (1, 2, 3):
        grad = rng.normal(size=G.PARAMS) * 0.1
        last = (G.LAYERS - 1) * G.LAYER_PARAMS
        grad[last + 48 * 64:last + 64 * 64] = 0
        grad[last + 64 * 64 + 48:] = 0
        gp, gm, gv = P.guide_adam(p, m, v, grad, step, 1e-2)
        rp, rm, rv = G.adam(p, m, v, grad.astype(np.float32).astype(np.float64), step, 1e-2)
        live = np.ones(G.PARAMS, bool)
        live[last + 48 * 64:last + 64 * 64] = False
        live[last + 64 * 64 + 48:] = False
        assert np.allclose(gp[live], rp[live], rtol=1e-5, atol=1e-6)
        assert np.allclose(gm[live], rm[live], rtol=1e-5, atol=1e-9)
        p, m, v = gp.astype(np.float64), gm.astype(np.float64), gv.astype(np.float64)


@pytest.mark.parametrize("kappa", [0.05, 2.0, 60.0, 5000.0])
def test_vmf_sampler_matches_density(kappa):
    """Jakob's vMF sampler: cos(theta) to the mean has the CDF
    (exp(kappa (w - 1)) - exp(-2 kappa)) / (1 - exp(-2 kappa)), the azimuth is uniform, and the
    returned density is kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (w - 1))."""
    from scipy import stats

    print(f"case: vmf sampler kappa={kappa}", flush=True)
    rng = np.random.default_rng(int(kappa * 10))
    n = 100000
    mu = rng.normal(size=3)
    mu /= np.linalg.norm(mu)
    d, pdf = P.vmf_sample(np.tile(mu, (n, 1)), np.full(n, kappa), rng.random((n, 2)))
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=2e-6)
    w = np.clip(d.astype(np.float64) @ mu, -1, 1)
    e = math.exp(-2 * kappa)

    def cdf(x):
        return (np.exp(kappa * (np.asarray(x) - 1)) - e) / (1 - e)

    assert stats.kstest(w, cdf).pvalue > 1e-3
    # azimuth about mu in a frame of our own
    a = np.cross(mu, [1.0, 0.0, 0.0] if abs(mu[0]) < 0.9 else [0.0, 1.0, 0.0])
    a /= np.linalg.norm(a)
    b = np.cross(mu, a)
    phi = (np.arctan2(d @ b, d @ a) + np.pi) / (2 * np.pi)
    ok = w < 1 - 1e-6
    assert stats.kstest(phi[ok], "uniform").pvalue > 1e-3
    ref = kappa / (2 * np.pi * (1 - e)) * np.exp(kappa * (w - 1))
    assert np.allclose(pdf, ref, rtol=2e-3 if kappa < 1000 else 3e-2)
