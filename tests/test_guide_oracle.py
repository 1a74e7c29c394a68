"""The guiding-network oracle against central finite differences (FP64) and against closed
forms of the vMF mixture it parameterises."""
import numpy as np
import pytest

import guide_oracle as G


def _batch(rng, n):
    feats = rng.random((n, 8))
    w = rng.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    pb = rng.uniform(0.05, 0.5, n)
    wt = rng.uniform(0.1, 3.0, n)
    return feats, w, pb, wt


def test_one_blob_encoding():
    e = G.encode(np.full((1, 8), 0.5 / 8))
    assert e.shape == (1, 64)
    # value 1/16 sits at the centre of bin 0 of every feature
    assert np.allclose(e[0, ::8], 1.0) and np.allclose(e[0, 1::8], np.exp(-0.5))


def test_vmf_normalised():
    # integral of the vMF density over the sphere = 1 for small and large kappa
    rng = np.random.default_rng(0)
    n = 400000
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for kappa in (0.05, 1.0, 20.0):
        mu = np.tile([[[0.0, 0.0, 1.0]]], (n, 1, 1))
        f, _ = G.vmf(mu, np.full((n, 1), kappa), v)
        assert f.mean() * 4 * np.pi == pytest.approx(1.0, rel=0.02)


def test_mixture_is_a_normalised_density():
    """Softmax weights sum to one, means are unit vectors, kappa and the selection fraction
    respect their clamps, and the weighted vMF mixture integrates to one over the sphere."""
    rng = np.random.default_rng(4)
    out = rng.normal(scale=3.0, size=(1, G.OUT))
    out[0, G.GO_KAPPA:G.GO_KAPPA + G.LOBES] = [-9, -3, 0, 1, 2, 4, 9, 20]
    pi, mu, _, kappa, _, _, c = G.mixture(out)
    assert pi.sum() == pytest.approx(1.0, abs=1e-12)
    assert np.allclose(np.linalg.norm(mu, axis=2), 1.0)
    assert np.allclose(kappa[0], np.exp(np.clip([-9, -3, 0, 1, 2, 4, 9, 20], G.LOGK_MIN, G.LOGK_MAX)))
    assert G.FRAC_MIN <= c[0] <= G.FRAC_MAX
    # stratified uniform directions (equal-area bands in z), the sharp lobes need many
    n = 2_000_000
    k = 4000
    z = 1 - 2 * ((np.arange(k)[:, None] + rng.random((k, n // k))) / k).reshape(-1)
    ph = rng.random(n) * 2 * np.pi
    r = np.sqrt(1 - z * z)
    v = np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)
    q = np.zeros(n)
    for j in range(G.LOBES):
        f, _ = G.vmf(np.tile(mu[:, j:j + 1], (n, 1, 1)), np.full((n, 1), kappa[0, j]), v)
        q += pi[0, j] * f[:, 0]
    assert q.mean() * 4 * np.pi == pytest.approx(1.0, rel=0.03)


def test_analytic_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = G.init_params(3)
    # make the last layer non-trivial so that every output matters
    last = (G.LAYERS - 1) * G.LAYER_PARAMS
    Wl = p[last:last + 64 * 64].reshape(64, 64)
    Wl[:48] = rng.uniform(-0.2, 0.2, (48, 64))
    feats, w, pb, wt = _batch(rng, 16)
    loss, grad = G.gradients(p, feats, w, pb, wt)
    idx = np.concatenate([rng.choice(G.PARAMS - G.LAYER_PARAMS, 40, replace=False), last + rng.choice(48 * 64, 40, replace=False),
                          last + 64 * 64 + np.arange(41)])
    h = 1e-6
    for i in idx:
        pp = p.copy(); pp[i] += h
        pm = p.copy(); pm[i] -= h
        fd = (G.mean_loss(pp, feats, w, pb, wt) - G.mean_loss(pm, feats, w, pb, wt)) / (2 * h)
        assert grad[i] == pytest.approx(fd, rel=2e-4, abs=1e-7), i


def test_adam_matches_definition():
    rng = np.random.default_rng(2)
    p = rng.normal(size=10)
    g = rng.normal(size=10)
    p1, m, v = G.adam(p, np.zeros(10), np.zeros(10), g, 1, 0.01)
    # first step: m_hat = g, v_hat = g^2 -> step of lr * sign(g) (up to eps)
    assert np.allclose(p1, p - 0.01 * np.sign(g), atol=1e-9)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 2**-7, 3.14159], np.float32)
    assert list(G.bf16(x)) == [1.0, 1.0, 1.0 + 2**-7, 3.140625]
