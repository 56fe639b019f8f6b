"""CPU restatement (numpy, float64) of the proposal network and its training targets.

TEST INFRASTRUCTURE ONLY: the checker for the GPU proposal kernels; the product never imports it.

The proposal has no counterpart in the reference; this file is its specification:
  features  x = [o + d tmin, o + d tmax] (the segment's entry and exit points, 6 values)
  network   h1 = relu(W0 x + b0), h2 = relu(W1 h1 + b1), logits = W2 h2 + b2   (6 -> 32 -> 32 -> 8)
            packed parameters W0 [32,6], b0 [32], W1 [32,32], b1 [32], W2 [8,32], b2 [8]
  density   p = (1 - MIX) softmax(logits) + MIX / 8 over 8 equal bins of [tmin, tmax]; cdf = cumsum(p), cdf[7] = 1
  target    q_b proportional to TV_b + FLOOR * sum(TV) / 8 + 1e-12, where TV_b sums |f(t_{j+1}) - f(t_j)| over the
            consecutive uniform samples (a = 0.5) whose midpoint falls in bin b (uniform q when the field is flat)
  loss      mean over segments of the cross entropy -sum_b q_b log softmax(logits)_b (gradient softmax - q)
Pinned in tests/test_oracle.py by torch autograd of the same network in float64.
"""
import numpy as np

IN, HID, OUT = 6, 32, 8
MIX = 0.25
FLOOR = 0.05
NPARAMS = HID * IN + HID + HID * HID + HID + OUT * HID + OUT


def unpack(params):
    p = np.asarray(params, np.float64)
    o = 0
    W0 = p[o:o + HID * IN].reshape(HID, IN); o += HID * IN
    b0 = p[o:o + HID]; o += HID
    W1 = p[o:o + HID * HID].reshape(HID, HID); o += HID * HID
    b1 = p[o:o + HID]; o += HID
    W2 = p[o:o + OUT * HID].reshape(OUT, HID); o += OUT * HID
    b2 = p[o:o + OUT]
    return W0, b0, W1, b1, W2, b2


def init_params(seed):
    """Xavier-uniform weights, zero biases, float32."""
    rng = np.random.default_rng(seed)
    out = []
    for fi, fo in ((IN, HID), (HID, HID), (HID, OUT)):
        a = np.sqrt(6.0 / (fi + fo))
        out += [rng.uniform(-a, a, fo * fi), np.zeros(fo)]
    return np.concatenate(out).astype(np.float32)


def features(origin, direction, seg_ray, tmin, tmax):
    o = np.asarray(origin, np.float64)[seg_ray]
    d = np.asarray(direction, np.float64)[seg_ray]
    t0 = np.asarray(tmin, np.float64)[:, None]
    t1 = np.asarray(tmax, np.float64)[:, None]
    return np.concatenate([o + d * t0, o + d * t1], 1)


def forward(params, x):
    W0, b0, W1, b1, W2, b2 = unpack(params)
    z1 = x @ W0.T + b0
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W1.T + b1
    h2 = np.maximum(z2, 0)
    logits = h2 @ W2.T + b2
    return logits, (x, z1, h1, z2, h2)


def softmax(logits):
    e = np.exp(logits - logits.max(1, keepdims=True))
    return e / e.sum(1, keepdims=True)


def cdf_from_logits(logits):
    p = (1 - MIX) * softmax(logits) + MIX / OUT
    c = np.cumsum(p, 1)
    c[:, -1] = 1.0
    return c


def backward(params, x, q):
    """Gradient of mean_k CE(q_k, softmax(logits_k)) with respect to the packed parameters."""
    W0, b0, W1, b1, W2, b2 = unpack(params)
    logits, (x, z1, h1, z2, h2) = forward(params, x)
    K = x.shape[0]
    dl = (softmax(logits) - q) / K
    gW2 = dl.T @ h2
    gb2 = dl.sum(0)
    dz2 = (dl @ W2) * (z2 > 0)
    gW1 = dz2.T @ h1
    gb1 = dz2.sum(0)
    dz1 = (dz2 @ W1) * (z1 > 0)
    gW0 = dz1.T @ x
    gb0 = dz1.sum(0)
    return np.concatenate([gW0.ravel(), gb0, gW1.ravel(), gb1, gW2.ravel(), gb2])


def loss(params, x, q):
    logits, _ = forward(params, x)
    ls = logits - logits.max(1, keepdims=True)
    logp = ls - np.log(np.exp(ls).sum(1, keepdims=True))
    return float(-(q * logp).sum(1).mean())


def targets(values, seg_count):
    """Bin targets q [K, 8] from the field's values at each segment's uniform samples (a = 0.5), concatenated
    segment after segment (seg_count samples each)."""
    values = np.asarray(values, np.float64)
    K = len(seg_count)
    q = np.zeros((K, OUT))
    off = 0
    for k, n in enumerate(seg_count):
        f = values[off:off + n]
        off += n
        tv = np.zeros(OUT)
        for j in range(n - 1):
            b = min(OUT - 1, (j + 1) * OUT // n)  # bin of the pair's midpoint, at (j + 1) / n of the segment
            tv[b] += abs(f[j + 1] - f[j])
        tv += FLOOR * tv.sum() / OUT + 1e-12
        q[k] = tv / tv.sum()
    return q
