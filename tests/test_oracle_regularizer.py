"""Pins the CPU restatement of the volume regulariser (oracle/regularizer.py): the total variation and the edge loss
against torch autograd of the reference's own tensor expressions in float64 (TVLoss::forward,
geometry_ex_ex.h:18-61; geometry_ex_ex.cpp:114-134), and the neighbour pairs against closed-form counts and the
geometry of the boxes they join."""
import numpy as np
import torch

from neat_amd import workloads as W
from oracle import regularizer as R


def tree_scene(depth, adaptive=0, seed=0):
    tree = W.build_tree(depth)
    if adaptive:
        ph = W.random_phantom(20, seed + 3)
        active = W.adaptive_active(tree, ph, adaptive, seed)
    else:
        active = (tree["depth"] == depth).astype(np.int32)
    ids, prefix = W.update_active(active)
    return tree, active, ids, prefix


def test_tv_matches_torch_autograd():
    rng = np.random.default_rng(0)
    grid = rng.normal(size=(6, 3, 5, 5, 5))
    slots = np.array([4, 0, 2, 5])
    w = 0.37
    loss, grad = R.tv_loss(grid, slots, w)
    g = torch.tensor(grid, requires_grad=True)
    sub = g[torch.from_numpy(slots)]
    total = None
    for d in range(2, 5):
        n = sub.shape[d]
        l = (sub.narrow(d, 0, n - 1) - sub.narrow(d, 1, n - 1)).abs().mean(dim=(1, 2, 3, 4)).mean()
        total = l if total is None else total + l
    ref = total * sub.shape[1] * w
    ref.backward()
    assert abs(loss - ref.item()) <= 1e-12 * abs(ref.item())
    assert np.allclose(grad, g.grad.numpy(), rtol=0, atol=1e-15)


def test_edge_loss_matches_torch_autograd():
    rng = np.random.default_rng(1)
    f = rng.normal(size=(2 * 37, 8))
    seed, w = 11, 0.25
    loss, grad = R.edge_loss(f, seed, w)
    t = torch.tensor(f, requires_grad=True)
    r = torch.from_numpy(R.pair_factor(seed, np.arange(37)).astype(np.float64))
    e = (t[0::2] - t[1::2]).abs().mean(1) * r
    ref = e.mean() * w * f.shape[1]
    ref.backward()
    assert abs(loss - ref.item()) <= 1e-12 * abs(ref.item())
    assert np.allclose(grad, t.grad.numpy(), rtol=0, atol=1e-15)


def test_pair_factor_is_uniform_in_unit_interval():
    r = R.pair_factor(5, np.arange(1 << 16))
    assert r.min() >= 0 and r.max() < 1 and abs(r.mean() - 0.5) < 0.01


def test_neighbor_pairs_uniform_tree_closed_form():
    # depth d: (2^d)^2 (2^d - 1) interfaces per axis, each seen from both sides, G^2 face points each
    for depth, G in ((1, 3), (2, 5)):
        tree, active, ids, _ = tree_scene(depth)
        coord, node = R.neighbor_pairs(tree, active, ids, G)
        k = 2 ** depth
        assert node.shape[0] == 2 * 3 * 2 * k * k * (k - 1) * G * G
        assert np.array_equal(coord[0::2], coord[1::2])


def test_neighbor_pairs_join_touching_boxes():
    tree, active, ids, _ = tree_scene(3, adaptive=60, seed=2)
    coord, node = R.neighbor_pairs(tree, active, ids, 4)
    assert node.shape[0] > 0
    a, b = node[0::2], node[1::2]
    assert np.all(a != b) and np.all(active[a] > 0) and np.all(active[b] > 0)
    p = coord[0::2]
    for n in (a, b):  # the point lies in both closed boxes
        assert np.all(tree["pos_min"][n] <= p + 1e-6) and np.all(p <= tree["pos_max"][n] + 1e-6)
    # canonical order: the first node of the pairs follows active_node_ids
    order = np.searchsorted(ids, a)
    assert np.all(np.diff(order) >= 0)


def test_node_loss_weight_sums_telescope(oracle):
    """AccumulateSampleLossPerNode's weight sums equal every node's total segment length."""
    from oracle import nodeloss as NL
    tree, active, ids, _ = tree_scene(2)
    o, d = W.cone_beam_rays(2, 16, 16)
    smp = oracle.raysample(tree, ids, o, d, 8, jitter=True, seed=3, emit=True, nthreads=1)
    loss = np.random.default_rng(0).uniform(size=o.shape[0])
    ls, ws = NL.accumulate(smp, loss, tree["pos_min"].shape[0])
    seglen = np.zeros_like(ws)
    np.add.at(seglen, smp["seg_node"], (smp["seg_tmax"].astype(np.float64) - smp["seg_tmin"]))
    assert np.allclose(ws, seglen, rtol=1e-5, atol=1e-6)
    assert ws[active == 0].sum() == 0 and ls.sum() > 0


def test_node_loss_indicator_rays_give_segment_lengths(oracle):
    """Linearity in the per-ray loss: with loss = 1 on one ray and 0 elsewhere, a node's loss sum is that ray's
    segment length inside the node (computed from the segment list, not from the samples' weights)."""
    from oracle import nodeloss as NL
    tree, active, ids, _ = tree_scene(3, adaptive=40, seed=1)
    o, d = W.cone_beam_rays(3, 8, 8)
    smp = oracle.raysample(tree, ids, o, d, 6, jitter=False, seed=0, emit=True, nthreads=1)
    nn = tree["pos_min"].shape[0]
    R_ = o.shape[0]
    seg_ray = smp["seg_ray"].astype(np.int64)
    for r in np.unique(seg_ray)[:: max(1, len(np.unique(seg_ray)) // 12)]:
        loss = np.zeros(R_)
        loss[r] = 1.0
        ls, _ = NL.accumulate(smp, loss, nn)
        want = np.zeros(nn)
        m = seg_ray == r
        np.add.at(want, smp["seg_node"][m], smp["seg_tmax"][m].astype(np.float64) - smp["seg_tmin"][m])
        assert np.allclose(ls, want, rtol=1e-5, atol=1e-6)
    # a constant loss c scales the weight sums
    ls, ws = NL.accumulate(smp, np.full(R_, 2.5), nn)
    assert np.allclose(ls, 2.5 * ws, rtol=1e-12, atol=0)


def test_set_error_semantics():
    """SetErrorForActiveNodes: "override" replaces active nodes' error and keeps inactive ones; "min" keeps the
    smaller of old and new on active nodes whose old error is valid (>= 0), takes the new one where the old is
    negative, and keeps inactive nodes."""
    from oracle import nodeloss as NL
    active = np.array([1, 0, 1, 1, 0, 1], np.int32)
    old = np.array([0.5, 7.0, -1.0, 2.0, -3.0, 0.25], np.float32)
    new = np.array([0.75, 9.0, 4.0, 1.5, 6.0, 0.25], np.float32)
    ov = NL.set_error(old, active, new, "override")
    assert np.array_equal(ov, np.array([0.75, 7.0, 4.0, 1.5, -3.0, 0.25], np.float32))
    mn = NL.set_error(old, active, new, "min")
    assert np.array_equal(mn, np.array([0.5, 7.0, 4.0, 1.5, -3.0, 0.25], np.float32))
