"""CPU restatement of the reference's volume regulariser (GeometryExEx::VolumeRegularizer,
src/geometry/geometry_ex_ex.cpp:86-160) -- TEST INFRASTRUCTURE ONLY: tests/ use it as the checker of the HIP kernels in
neat_amd/csrc/kernels/regularizer.hip; the product never imports it.

  neighbor_pairs  HyperTreeBaseImpl::NodeNeighborSamples (src/structure/HyperTree.cu:830-940) with
                  ActiveNodeIdForGlobalPosition (HyperTree.cu:600-640) and BoxContainsPoint (HyperTree.h:85-95), in
                  the canonical order (active node, side, front, face point); float32 arithmetic as the reference's.
                  Pinned: closed-form pair counts of uniform trees, every pair's point on the face the two boxes share.
  edge_loss       geometry_ex_ex.cpp:114-134 on given features, float64, with the per-pair factor r_k (a splitmix64
                  hash of (seed, k) where the reference draws torch::rand_like). Pinned: torch autograd in float64.
  tv_loss         TVLoss::forward (src/geometry/geometry_ex_ex.h:18-61) times loss_tv, float64. Pinned: torch
                  autograd of the reference's slice expression in float64.
"""
import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z):
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def pair_factor(seed, k):
    """r_k in [0, 1): 24 bits of splitmix64(seed ^ k * golden) (the HIP kernel's pair_factor)."""
    k = np.asarray(k, np.uint64)
    with np.errstate(over="ignore"):
        h = _mix64(np.uint64(seed) ^ (k * _GOLD))
    return (h >> np.uint64(40)).astype(np.float32) * np.float32(1.0 / 16777216.0)


def active_node_at(tree, node_active, q):
    """ActiveNodeIdForGlobalPosition for points q [M, 3] (float32): the active node reached from the root through
    inactive nodes by the first child (in child order) whose closed box contains the point; -1 if none."""
    pmin, pmax = tree["pos_min"], tree["pos_max"]
    num_nodes = pmin.shape[0]
    first_leaf = (num_nodes - 1) // 8
    M = q.shape[0]
    n = np.zeros(M, np.int64)
    res = np.full(M, -2, np.int64)  # -2: unresolved
    for _ in range(12):
        open_ = res == -2
        if not open_.any():
            break
        act = open_ & (node_active[n] > 0)
        res[act] = n[act]
        open_ &= ~act
        leaf = open_ & (n >= first_leaf)
        res[leaf] = -1
        open_ &= ~leaf
        nxt = np.full(M, -1, np.int64)
        for c in range(8):
            ch = 8 * n + 1 + c
            chc = np.clip(ch, 0, num_nodes - 1)
            inside = np.all((pmin[chc] <= q) & (pmax[chc] >= q), axis=1)
            take = open_ & (nxt == -1) & inside
            nxt[take] = ch[take]
        none = open_ & (nxt == -1)
        res[none] = -1
        n = np.where(open_ & (nxt >= 0), nxt, n)
    return res


def neighbor_pairs(tree, node_active, active_ids, grid_size, epsilon=0.001):
    """Neighbour rows in canonical order: coordinates float32 [2P, 3], node ids int64 [2P] (a, neighbour, ...)."""
    G = int(grid_size)
    pmin = tree["pos_min"].astype(np.float32)
    pmax = tree["pos_max"].astype(np.float32)
    eps = np.float32(epsilon)
    sid = np.arange(G * G)
    coords, nodes = [], []
    for a in np.asarray(active_ids, np.int64):
        mn, mx = pmin[a], pmax[a]
        box = (mx - mn).astype(np.float32)
        for side in range(3):
            sa, sb = (side + 1) % 3, (side + 2) % 3
            for front in range(2):
                local = np.zeros((G * G, 3), np.float32)
                local[:, side] = np.float32(1.0 if front else 0.0)
                local[:, sa] = (sid % G).astype(np.float32) / (np.float32(G) - np.float32(1.0))
                local[:, sb] = (sid // G).astype(np.float32) / (np.float32(G) - np.float32(1.0))
                p = (local * box + mn).astype(np.float32)
                q = p.copy()
                q[:, side] = p[:, side] + (eps if front else -eps)
                o = active_node_at(tree, node_active, q)
                keep = (o != -1) & (o != a)
                for i in np.nonzero(keep)[0]:
                    coords += [p[i], p[i]]
                    nodes += [a, o[i]]
    if not nodes:
        return np.zeros((0, 3), np.float32), np.zeros(0, np.int64)
    return np.array(coords, np.float32), np.array(nodes, np.int64)


def edge_loss(features, seed, weight):
    """loss = mean_k(mean_c |f[2k] - f[2k+1]| * r_k) * weight * C and d loss / d features (float64)."""
    f = np.asarray(features, np.float64)
    P, C = f.shape[0] // 2, f.shape[1]
    if P == 0:
        return 0.0, np.zeros_like(f)
    r = pair_factor(seed, np.arange(P)).astype(np.float64)
    d = f[0::2] - f[1::2]
    loss = float(np.mean(np.abs(d).mean(axis=1) * r) * weight * C)
    g = np.zeros_like(f)
    gd = np.sign(d) * (r[:, None] * weight / P)
    g[0::2] = gd
    g[1::2] = -gd
    return loss, g


def tv_loss(grid, slots, weight):
    """TVLoss::forward(grid[slots]) * weight (float64) and its gradient w.r.t. grid (same shape)."""
    g = np.asarray(grid, np.float64)
    sub = g[np.asarray(slots, np.int64)]
    C = g.shape[1]
    loss = 0.0
    grad_sub = np.zeros_like(sub)
    for ax in (2, 3, 4):
        n = sub.shape[ax]
        a = np.take(sub, range(0, n - 1), axis=ax)
        b = np.take(sub, range(1, n), axis=ax)
        d = a - b
        cnt = d[0].size  # the per-batch mean over (C, cells); then the mean over the batch
        loss += np.abs(d).sum() / (cnt * sub.shape[0])
        s = np.sign(d) / (cnt * sub.shape[0])
        ga = np.zeros_like(sub)
        idx_a = [slice(None)] * 5
        idx_b = [slice(None)] * 5
        idx_a[ax] = slice(0, n - 1)
        idx_b[ax] = slice(1, n)
        ga[tuple(idx_a)] += s
        ga[tuple(idx_b)] -= s
        grad_sub += ga
    grad = np.zeros_like(g)
    np.add.at(grad, np.asarray(slots, np.int64), grad_sub * weight * C)
    return loss * weight * C, grad
