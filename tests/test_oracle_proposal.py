"""Pins the oracle's proposal parts (no reference counterpart; the build defines them): the placement of samples in
equal-probability cells of a segment's 8-bin density, the proposal network and its cross-entropy gradient, and the
training targets.

Placement checks that fail for plausible mistakes (a cell edge off by one, a weight that is not the cell length, a
CDF read one bin late): a uniform CDF reproduces the uniform samples; the weights of any CDF sum to the segment
length (constant field: kappa x chord); cells of an affine field are integrated exactly by their midpoints whatever
the CDF; a CDF concentrated in one bin puts the samples there. The network is checked against torch autograd in
float64, the targets against hand-computed variations."""
import numpy as np
import pytest
import torch

from oracle import proposal as PR
from test_oracle import F32, _linear_mlp, _rays, cube_chord, exact_interval, full_tiling


def _uniform_cdf(K):
    return np.tile((np.arange(1, 9) / 8.0).astype(F32), (K, 1))


def _random_cdf(K, seed, mix=0.25):
    rng = np.random.default_rng(seed)
    p = rng.dirichlet(np.full(8, 0.5), K) * (1 - mix) + mix / 8
    c = np.cumsum(p, 1).astype(F32)
    c[:, 7] = 1.0
    return c


def test_uniform_cdf_reproduces_uniform_samples(oracle):
    t, ids, prefix = full_tiling(oracle, 2)
    o, d = _rays(120, 21)
    segs = oracle.raysample(t, ids, o, d, 16, emit=False)
    rng = np.random.default_rng(3)
    grid = (rng.normal(size=(64, 8, 5, 5, 5)) * 0.2).astype(F32)
    P = oracle.decoder_num_params(8, 16, 1, 1)
    mlp = dict(params=(rng.normal(size=P) * 0.3).astype(F32), **{"in": 8}, hidden=16, layers=1, out=1, act="silu",
               last_act="softplus", beta=2.0)
    base = oracle.projector_fwd(t, prefix, o, d, segs, grid, mlp, prec=1)
    prop = oracle.projector_fwd(t, prefix, o, d, dict(segs, seg_cdf=_uniform_cdf(len(segs["seg_node"]))), grid, mlp,
                                prec=1)
    assert np.allclose(prop, base, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("jitter", [False, True])
def test_cell_weights_sum_to_segment_length(oracle, jitter):
    t, ids, prefix = full_tiling(oracle, 2)
    o, d = _rays(150, 22)
    segs = oracle.raysample(t, ids, o, d, 32, emit=False)
    K = len(segs["seg_node"])
    segs = dict(segs, seg_cdf=_random_cdf(K, 5), seg_count=np.full(K, 8, np.int32))
    grid = np.zeros((64, 8, 5, 5, 5), F32)
    proj = oracle.projector_fwd(t, prefix, o, d, segs, grid, _linear_mlp(8, 16, kappa=0.5), jitter=jitter, seed=4,
                                prec=1)
    chords = np.array([cube_chord(o[i], d[i]) for i in range(150)])
    assert np.allclose(proj, 0.5 * chords, rtol=1e-9, atol=1e-9)


def test_affine_field_exact_for_any_cdf(oracle):
    depth, G = 2, 5
    t, ids, prefix = full_tiling(oracle, depth)
    a = np.array([0.375, -0.25, 0.875])
    b = 0.25
    grid = np.zeros((len(ids), 8, G, G, G), F32)
    lin = np.linspace(0, 1, G)
    zz, yy, xx = np.meshgrid(lin, lin, lin, indexing="ij")
    for slot, n in enumerate(ids):
        mn, mx = t["pos_min"][n].astype(np.float64), t["pos_max"][n].astype(np.float64)
        grid[slot, 2] = (a[0] * (mn[0] + xx * (mx[0] - mn[0])) + a[1] * (mn[1] + yy * (mx[1] - mn[1])) +
                         a[2] * (mn[2] + zz * (mx[2] - mn[2])) + b)
    o, d = _rays(100, 23)
    segs = oracle.raysample(t, ids, o, d, 16, emit=False)
    K = len(segs["seg_node"])
    segs = dict(segs, seg_cdf=_random_cdf(K, 6), seg_count=np.full(K, 5, np.int32))
    proj = oracle.projector_fwd(t, prefix, o, d, segs, grid, _linear_mlp(8, 16, take_channel=2), prec=1)
    exp = np.zeros(100)
    for k in range(K):
        r, n = segs["seg_ray"][k], segs["seg_node"][k]
        ex = exact_interval(t["pos_min"][n], t["pos_max"][n], o[r], d[r])
        t0, t1 = float(ex[0]), float(ex[1])
        oo, dd = o[r].astype(np.float64), d[r].astype(np.float64)
        exp[r] += (a @ oo + b) * (t1 - t0) + (a @ dd) * (t1 * t1 - t0 * t0) / 2
    assert np.allclose(proj, exp, rtol=1e-9, atol=1e-9)


def test_cells_match_the_placement_rule(oracle):
    # one segment along the x axis through the cube [-1, 1]^3 (depth 0: one node); channel 0 = x exactly (trilinear
    # reproduces it) and the decoder returns softplus(x), so the projection is sum_j w_j softplus(x(t_j)): it pins
    # where the cells are, not only that they tile the segment
    t, ids, prefix = full_tiling(oracle, 0)
    G = 3
    grid = np.zeros((1, 8, G, G, G), F32)
    grid[0, 0] = np.linspace(-1, 1, G)[None, None, :]
    o = np.array([[-2.0, 0.1, 0.2]], F32)
    d = np.array([[1.0, 0.0, 0.0]], F32)
    segs = oracle.raysample(t, ids, o, d, 8, emit=False)
    assert segs["seg_tmin"][0] == 1.0 and segs["seg_tmax"][0] == 3.0
    mlp = dict(_linear_mlp(8, 16, take_channel=0), last_act="softplus")
    for seed, n in ((1, 4), (2, 8), (3, 5)):
        p = np.random.default_rng(seed).dirichlet(np.full(8, 0.3)) * 0.75 + 0.25 / 8
        cdf = np.cumsum(p).astype(F32)[None]
        cdf[0, 7] = 1
        proj = oracle.projector_fwd(t, prefix, o, d, dict(segs, seg_cdf=cdf, seg_count=np.array([n], np.int32)),
                                    grid, mlp, prec=1)
        cells = [PR_cell(cdf[0].astype(np.float64), 1.0, 3.0, n, j) for j in range(n + 1)]
        exp = sum((cells[j + 1] - cells[j]) * np.logaddexp(0, 2 * (-2 + (cells[j] + cells[j + 1]) / 2)) / 2
                  for j in range(n))
        assert abs(proj[0] - exp) < 1e-12, (seed, n, proj[0], exp)


def PR_cell(cdf, tmin, tmax, n, j):
    """The placement rule written out independently (equal-probability cell edges of the piecewise-constant density)."""
    if j <= 0:
        return tmin
    if j >= n:
        return tmax
    u = j / n
    lo = 0.0
    for b in range(8):
        if cdf[b] > u:
            f = (u - lo) / (cdf[b] - lo)
            return tmin + (tmax - tmin) * (b + f) / 8
        lo = cdf[b]
    return tmax


def test_network_and_gradient_match_torch_autograd():
    rng = np.random.default_rng(0)
    params = PR.init_params(1).astype(np.float64) + rng.normal(size=PR.NPARAMS) * 0.05
    x = rng.uniform(-1, 1, (50, 6))
    q = rng.dirichlet(np.ones(8), 50)
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    W0, b0, W1, b1, W2, b2 = torch.split(p, [192, 32, 1024, 32, 256, 8])
    xt = torch.tensor(x)
    h1 = torch.relu(xt @ W0.view(32, 6).T + b0)
    h2 = torch.relu(h1 @ W1.view(32, 32).T + b1)
    logits = h2 @ W2.view(8, 32).T + b2
    lossv = -(torch.tensor(q) * torch.log_softmax(logits, 1)).sum(1).mean()
    lossv.backward()
    lg, _ = PR.forward(params, x)
    assert np.allclose(lg, logits.detach().numpy(), rtol=1e-12, atol=1e-12)
    assert abs(PR.loss(params, x, q) - lossv.item()) < 1e-12
    assert np.allclose(PR.backward(params, x, q), p.grad.numpy(), rtol=1e-10, atol=1e-12)
    c = PR.cdf_from_logits(lg)
    assert np.all(np.diff(c, axis=1) >= PR.MIX / 8 - 1e-12) and np.all(c[:, -1] == 1.0)


def test_targets_follow_the_variation():
    # flat field: uniform target; one jump between samples 10 and 11 of 16: the bin of (11 / 16) gets the variation
    flat = np.full(16, 0.3)
    step = np.where(np.arange(16) > 10, 1.0, 0.0)
    q = PR.targets(np.concatenate([flat, step]), [16, 16])
    assert np.allclose(q[0], 1 / 8)
    b = (11 * 8) // 16
    assert q[1].argmax() == b
    floor = PR.FLOOR * 1.0 / 8
    assert np.isclose(q[1, b], (1.0 + floor) / (1.0 + 8 * floor))
