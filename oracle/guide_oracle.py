"""Oracle of the guiding network (numpy, FP64).

TEST INFRASTRUCTURE ONLY: the checker for the product's BF16 MFMA kernels.  The reference
has no learned component, so this restates the product's own specification:

* one-blob encoding of 8 raw features into 8 Gaussian bins each (sigma = one bin);
* 5 dense layers 64 -> 64 (ReLU after the first four; the last layer's 48 outputs, padded to
  64, are raw mixture parameters);
* 8-lobe von Mises-Fisher mixture: softmax logits, normalised means, kappa = exp(clamp(raw)),
  selection probability c = clamp(sigmoid(raw));
* loss per record = -w (log q(w_i) + log(c q(w_i) + (1 - c) p_bsdf(w_i)));
* Adam (0.9, 0.999, 1e-8) on FP32 master weights.

The analytic gradient is pinned by central finite differences in tests/test_guide_oracle.py.
"""
import numpy as np

RAW, BINS, W, LAYERS, OUT, LOBES = 8, 8, 64, 5, 48, 8
LAYER_PARAMS = W * W + W
PARAMS = LAYERS * LAYER_PARAMS
GO_LOGIT, GO_MEAN, GO_KAPPA, GO_FRAC = 0, 8, 32, 40
LOGK_MIN, LOGK_MAX = -3.0, 9.0
FRAC_MIN, FRAC_MAX = 0.02, 0.98


def bf16(x):
    """Round float64/32 values to bfloat16 (round to nearest even) and back."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64).reshape(np.shape(x))


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    Ws, bs = [], []
    for L in range(LAYERS):
        base = L * LAYER_PARAMS
        Ws.append(p[base:base + W * W].reshape(W, W))
        bs.append(p[base + W * W:base + LAYER_PARAMS])
    return Ws, bs


def encode(feats):
    f = np.asarray(feats, dtype=np.float64).reshape(-1, RAW)
    j = np.arange(BINS) + 0.5
    x = f[:, :, None] * BINS - j[None, None, :]
    return np.exp(-0.5 * x * x).reshape(-1, RAW * BINS)


def forward(params, feats, emulate_bf16=False):
    """Raw outputs (n, 48) and the layer inputs (for backward)."""
    Ws, bs = unpack(params)
    rnd = bf16 if emulate_bf16 else (lambda v: v)
    a = rnd(encode(feats))
    acts = [a]
    for L in range(LAYERS):
        z = a @ rnd(Ws[L]).T + bs[L]
        if L < LAYERS - 1:
            a = rnd(np.maximum(z, 0.0))
            acts.append(a)
        else:
            out = z[:, :OUT]
    return out, acts


def mixture(out):
    o = np.asarray(out, dtype=np.float64)
    lg = o[:, GO_LOGIT:GO_LOGIT + LOBES]
    lg = lg - lg.max(1, keepdims=True)
    pi = np.exp(lg)
    pi /= pi.sum(1, keepdims=True)
    m = o[:, GO_MEAN:GO_MEAN + 3 * LOBES].reshape(-1, LOBES, 3)
    ml = np.linalg.norm(m, axis=2)
    mu = m / np.maximum(ml, 1e-300)[:, :, None]
    lk = o[:, GO_KAPPA:GO_KAPPA + LOBES]
    kappa = np.exp(np.clip(lk, LOGK_MIN, LOGK_MAX))
    craw = 1.0 / (1.0 + np.exp(-o[:, GO_FRAC]))
    c = np.clip(craw, FRAC_MIN, FRAC_MAX)
    return pi, mu, ml, kappa, lk, craw, c


def vmf(mu, kappa, w):
    cosw = np.einsum("nkc,nc->nk", mu, w)
    return kappa / (2 * np.pi * (1 - np.exp(-2 * kappa))) * np.exp(kappa * (cosw - 1)), cosw


def loss_and_grad_out(out, w, pb, wt):
    """Per-record loss and d loss / d raw outputs (n, 48)."""
    pi, mu, ml, kappa, lk, craw, c = mixture(out)
    w = np.asarray(w, dtype=np.float64).reshape(-1, 3)
    pb = np.asarray(pb, dtype=np.float64)
    wt = np.asarray(wt, dtype=np.float64)
    f, cosw = vmf(mu, kappa, w)
    q = np.maximum((pi * f).sum(1), 1e-12)
    mix = np.maximum(c * q + (1 - c) * pb, 1e-12)
    loss = -wt * (np.log(q) + np.log(mix))
    dq = -wt * (1 / q + c / mix)
    g = np.zeros((out.shape[0], OUT))
    inside = (craw > FRAC_MIN) & (craw < FRAC_MAX)
    g[:, GO_FRAC] = np.where(inside, -wt * (q - pb) / mix * craw * (1 - craw), 0.0)
    g[:, GO_LOGIT:GO_LOGIT + LOBES] = dq[:, None] * pi * (f - q[:, None])
    df = dq[:, None] * pi * f
    e2 = np.exp(-2 * kappa)
    dlk = 1 / kappa - 2 * e2 / (1 - e2) + (cosw - 1)
    g[:, GO_KAPPA:GO_KAPPA + LOBES] = np.where((lk > LOGK_MIN) & (lk < LOGK_MAX), df * dlk * kappa, 0.0)
    s = df * kappa / np.maximum(ml, 1e-300)
    gm = s[:, :, None] * (w[:, None, :] - cosw[:, :, None] * mu)
    g[:, GO_MEAN:GO_MEAN + 3 * LOBES] = gm.reshape(-1, 3 * LOBES)
    zero = wt <= 0
    g[zero] = 0.0
    loss = np.where(zero, 0.0, loss)
    return loss, np.clip(g, -1e3, 1e3)


def mean_loss(params, feats, w, pb, wt):
    out, _ = forward(params, feats)
    loss, _ = loss_and_grad_out(out, w, pb, wt)
    return loss.mean()


def gradients(params, feats, w, pb, wt, emulate_bf16=False):
    """Gradient of the mean loss with respect to all parameters (PARAMS,)."""
    Ws, bs = unpack(params)
    rnd = bf16 if emulate_bf16 else (lambda v: v)
    out, acts = forward(params, feats, emulate_bf16)
    loss, g = loss_and_grad_out(out, w, pb, wt)
    n = out.shape[0]
    delta = np.zeros((n, W))
    delta[:, :OUT] = rnd(g)
    grad = np.zeros(PARAMS)
    for L in range(LAYERS - 1, -1, -1):
        base = L * LAYER_PARAMS
        grad[base:base + W * W] = (delta.T @ acts[L]).reshape(-1) / n
        grad[base + W * W:base + LAYER_PARAMS] = delta.sum(0) / n
        if L > 0:
            delta = rnd((delta @ rnd(Ws[L])) * (acts[L] > 0))
    # padding rows of the last layer carry no gradient
    last = (LAYERS - 1) * LAYER_PARAMS
    gW = grad[last:last + W * W].reshape(W, W)
    gW[OUT:] = 0.0
    grad[last + W * W + OUT:last + LAYER_PARAMS] = 0.0
    return loss.mean(), grad


def adam(params, m, v, grad, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1**step)
    vh = v / (1 - b2**step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed=0):
    """Same shape of initialisation as the product (values differ: different generator)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAMS)
    for L in range(LAYERS):
        base = L * LAYER_PARAMS
        lim = np.sqrt(6 / W) if L < LAYERS - 1 else 1e-3
        Wm = rng.uniform(-lim, lim, (W, W))
        if L == LAYERS - 1:
            Wm[OUT:] = 0
        p[base:base + W * W] = Wm.reshape(-1)
        if L == LAYERS - 1:
            b = p[base + W * W:base + LAYER_PARAMS]
            s = 1 / np.sqrt(3)
            for k in range(LOBES):
                b[GO_MEAN + 3 * k:GO_MEAN + 3 * k + 3] = [s if k & 1 else -s, s if k & 2 else -s, s if k & 4 else -s]
                b[GO_KAPPA + k] = np.log(2.0)
            b[GO_FRAC] = -1.5
    return p
