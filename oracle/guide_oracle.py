"""Float64 restatement of the learned path-guiding proposal -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py's CPU leg use this module, and only as the
checker.  spica has no learned sampling (its PathIntegrator samples continuations from
the BSDF alone, sources/integrators/path/path.cc:83-94); this file is the specification
the GPU kernels implement:

  inputs     x in [0,1]^7: position in the scene box (3), normal (2), outgoing
             direction (2); a unit vector v maps to ((v_z+1)/2, atan2(v_y,v_x)/2pi + 1/2)
  encoding   one-blob, 32 bins per input: e[32d+b] = exp(-(32 x_d - b - 1/2)^2 / 2)
  network    224 -> 64 -> 64 -> 64 -> 64 -> 40, ReLU between layers
  mixture    8 von Mises-Fisher lobes, per lobe (logit a, mean m (3), log-concentration r):
             pi = softmax(a), mu = m/|m|, kappa = exp(clamp(r, -4, 7))
             v(w) = kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.w - 1))
  loss       per training sample (x, w, F/p): -(F/p) log q(w); the batch loss is the mean
  optimiser  Adam with bias correction

`emulate=True` reproduces the GPU's rounding points: bf16 weights, encodings, hidden
activations and back-propagated deltas, fp32 biases and accumulation.  Its gradient is
pinned by central finite differences of the float64 loss (tests/test_guide_oracle.py);
the vMF density by quadrature over the sphere and the sampler by a histogram test.
"""
import numpy as np

IN, BINS, ENC, HID, LOBES, OUT = 7, 32, 224, 64, 8, 40
SHAPES = [(HID, ENC), (HID, HID), (HID, HID), (HID, HID), (OUT, HID)]
KLO, KHI = -4.0, 7.0


def _offsets():
    offs, o = [], 0
    for n, k in SHAPES:
        offs.append((o, o + n * k))
        o += n * k + n
    return offs, o


OFFSETS, NPARAMS = _offsets()


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64).reshape(a.shape)


def f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def unpack(theta):
    theta = np.asarray(theta, dtype=np.float64)
    layers = []
    for (n, k), (ow, ob) in zip(SHAPES, OFFSETS):
        layers.append((theta[ow:ow + n * k].reshape(n, k), theta[ob:ob + n]))
    return layers


def init_params(seed=1):
    """Xavier-uniform weights, zero biases (any initialisation works for the checks)."""
    rng = np.random.default_rng(seed)
    th = np.zeros(NPARAMS)
    for (n, k), (ow, ob) in zip(SHAPES, OFFSETS):
        s = np.sqrt(6.0 / (n + k))
        th[ow:ow + n * k] = rng.uniform(-s, s, n * k)
    return th


def unit_to_features(v):
    v = np.asarray(v, dtype=np.float64)
    return np.stack([(v[..., 2] + 1.0) * 0.5, np.arctan2(v[..., 1], v[..., 0]) / (2 * np.pi) + 0.5], axis=-1)


def features(p, lo, hi, n, wo):
    """The 7 network inputs for world-space position, normal and outgoing direction."""
    px = np.clip((np.asarray(p) - lo) / (np.asarray(hi) - lo), 0.0, 1.0)
    return np.concatenate([px, unit_to_features(n), unit_to_features(wo)], axis=-1)


def encode(x):
    x = np.asarray(x, dtype=np.float64)
    b = np.arange(BINS) + 0.5
    e = np.exp(-0.5 * (BINS * x[:, :, None] - b[None, None, :]) ** 2)
    return e.reshape(len(x), ENC)


def forward(theta, x, emulate=False):
    """Returns (outputs [N,40], cache) with cache = (inputs of each layer, pre-activations)."""
    layers = unpack(theta)
    a = encode(x)
    if emulate:
        a = bf16(a)
    acts, pre = [a], []
    for li, (W, b) in enumerate(layers):
        Wq = bf16(W) if emulate else W
        bq = f32(b) if emulate else b
        z = a @ Wq.T + bq
        if emulate:
            z = f32(z)
        pre.append(z)
        if li < len(layers) - 1:
            a = np.maximum(z, 0.0)
            if emulate:
                a = bf16(a)
            acts.append(a)
        else:
            a = z
    return a, (acts, pre)


def mixture(o):
    """pi [N,8], mu [N,8,3], kappa [N,8], plus (|m|, r-inside mask) for the gradient."""
    o = np.asarray(o, dtype=np.float64).reshape(len(o), LOBES, 5)
    a = o[:, :, 0]
    a = a - a.max(axis=1, keepdims=True)
    pi = np.exp(a)
    pi /= pi.sum(axis=1, keepdims=True)
    m = o[:, :, 1:4]
    mlen = np.sqrt(np.maximum((m * m).sum(-1), 1e-12))
    mu = m / mlen[..., None]
    r = o[:, :, 4]
    inside = (r > KLO) & (r < KHI)
    kappa = np.exp(np.clip(r, KLO, KHI))
    return pi, mu, kappa, mlen, inside


def log_vmf_norm(kappa):
    return np.log(kappa) - np.log(2 * np.pi) - np.log(-np.expm1(-2.0 * kappa))


def log_pdf(o, w):
    """log q(w) of the mixture for outputs o [N,40] and unit directions w [N,3]."""
    pi, mu, kappa, _, _ = mixture(o)
    w = np.asarray(w, dtype=np.float64)
    lv = log_vmf_norm(kappa) + kappa * ((mu * w[:, None, :]).sum(-1) - 1.0)
    t = np.log(np.maximum(pi, 1e-300)) + lv
    mx = t.max(axis=1, keepdims=True)
    return (mx + np.log(np.exp(t - mx).sum(axis=1, keepdims=True)))[:, 0]


def frame(n):
    """Orthonormal basis around unit n (branch-free construction of Duff et al. 2017)."""
    n = np.asarray(n, dtype=np.float64)
    s = np.where(n[..., 2] >= 0.0, 1.0, -1.0)
    a = -1.0 / (s + n[..., 2])
    b = n[..., 0] * n[..., 1] * a
    t1 = np.stack([1.0 + s * n[..., 0] ** 2 * a, s * b, -s * n[..., 0]], axis=-1)
    t2 = np.stack([b, s + n[..., 1] ** 2 * a, -n[..., 1]], axis=-1)
    return t1, t2


def sample(o, u):
    """Directions from the mixture: u[:,0] picks the lobe by the CDF of pi, u[:,1:3]
    sample the vMF (inverse CDF of cos theta about mu, uniform azimuth)."""
    pi, mu, kappa, _, _ = mixture(o)
    u = np.asarray(u, dtype=np.float64)
    cdf = np.cumsum(pi, axis=1)
    k = np.minimum((u[:, :1] >= cdf).sum(axis=1), LOBES - 1)
    idx = np.arange(len(o))
    m, kk = mu[idx, k], kappa[idx, k]
    v = 1.0 - u[:, 1]
    cz = np.clip(1.0 + np.log1p(v * np.expm1(-2.0 * kk)) / kk, -1.0, 1.0)
    sz = np.sqrt(np.maximum(0.0, 1.0 - cz * cz))
    ph = 2 * np.pi * u[:, 2]
    t1, t2 = frame(m)
    d = cz[:, None] * m + sz[:, None] * (np.cos(ph)[:, None] * t1 + np.sin(ph)[:, None] * t2)
    return d, np.exp(log_pdf(o, d))


def output_grad(o, w, weight):
    """d/do of -weight * log q(w), per sample; returns (loss [N], delta [N,40])."""
    pi, mu, kappa, mlen, inside = mixture(o)
    w = np.asarray(w, dtype=np.float64)
    weight = np.asarray(weight, dtype=np.float64)
    cosk = (mu * w[:, None, :]).sum(-1)
    lv = log_vmf_norm(kappa) + kappa * (cosk - 1.0)
    t = np.log(np.maximum(pi, 1e-300)) + lv
    mx = t.max(axis=1, keepdims=True)
    lq = mx[:, 0] + np.log(np.exp(t - mx).sum(axis=1))
    gam = np.exp(t - lq[:, None])
    W = weight[:, None]
    d = np.zeros((len(o), LOBES, 5))
    d[:, :, 0] = -W * (gam - pi)
    two_k = 2.0 * kappa
    with np.errstate(over="ignore"):  # 2k / expm1(2k) -> 0 for large k
        dlogv_dr = 1.0 - two_k / np.expm1(two_k) + kappa * (cosk - 1.0)
    d[:, :, 4] = np.where(inside, -W * gam * dlogv_dr, 0.0)
    tang = w[:, None, :] - cosk[..., None] * mu
    d[:, :, 1:4] = (-W * gam * kappa / mlen)[..., None] * tang
    return -weight * lq, d.reshape(len(o), OUT)


def loss_grad(theta, x, w, weight, emulate=False):
    """Mean loss over the batch and its gradient w.r.t. the flat parameter vector."""
    layers = unpack(theta)
    o, (acts, pre) = forward(theta, x, emulate)
    if emulate:
        o = f32(o)
    loss, delta = output_grad(o, w, weight)
    N = len(x)
    g = np.zeros(NPARAMS)
    for li in range(len(layers) - 1, -1, -1):
        if emulate:
            delta = bf16(delta)
        (ow, ob) = OFFSETS[li]
        n, k = SHAPES[li]
        g[ow:ow + n * k] = (delta.T @ acts[li]).ravel()
        g[ob:ob + n] = delta.sum(axis=0)
        if li > 0:
            W = bf16(layers[li][0]) if emulate else layers[li][0]
            dh = delta @ W
            if emulate:
                dh = f32(dh)
            delta = dh * (pre[li - 1] > 0.0)
    return loss.mean(), g / N


def loss_value(theta, x, w, weight):
    o, _ = forward(theta, x)
    return float(np.mean(-np.asarray(weight) * log_pdf(o, w)))


def adam(theta, m, v, t, g, lr=1e-2, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t counts from 1); returns (theta, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v
