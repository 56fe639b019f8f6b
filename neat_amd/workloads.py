"""Synthetic projector workloads: octrees (uniform or phantom-adaptive), cone-beam rays, random-ellipsoid phantoms
with closed-form line integrals, feature grids and decoder parameters, all from a seed.

The configurations follow the projector benchmark suite:
  C1  uniform depth-2 octree (64 leaves), 11^3 x 8 grid, decoder 8->32->32->1, 16 x 128^2 cone-beam rays,
      max 32 samples per node, 20-ellipsoid target phantom.
  C2  depth-5 octree adapted to a 100-ellipsoid phantom (4096 active leaves), 17^3 x 8 grid, decoder 8->64->64->1,
      256 x 512^2 cone-beam rays in batches of 2^18, max 64 samples per node.
Geometry of the cone beam (nikon2neat.cpp:278-291): source 4 (sin(-t), -cos(-t), 0), detector centre on the
opposite side at distance 4, detector up (0, 0, -1); the detector spans +-det_extent world units.
"""
import heapq
import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi


# ---- octree -----------------------------------------------------------------------------------------------------------
def build_tree(max_depth):
    """Node table of the complete octree (HyperTree.cpp:11-132) as numpy arrays."""
    L = _abi.lib()
    n = int(L.neat_tree_num_nodes(max_depth))
    t = dict(parent=np.empty(n, np.int32), children=np.empty((n, 8), np.int32), pos_min=np.empty((n, 3), np.float32),
             pos_max=np.empty((n, 3), np.float32), scale=np.empty(n, np.float32), depth=np.empty(n, np.int32))
    _abi.check(L.neat_tree_build_host(max_depth, *(t[k].ctypes.data for k in
                                                   ("parent", "children", "pos_min", "pos_max", "scale", "depth"))),
               "neat_tree_build_host")
    t["max_depth"] = max_depth
    return t


def update_active(node_active):
    na = np.ascontiguousarray(node_active, np.int32)
    ids = np.empty(na.shape[0], np.int64)
    prefix = np.empty(na.shape[0], np.int64)
    c = _abi.lib().neat_tree_update_active_host(na.ctypes.data, na.shape[0], ids.ctypes.data, prefix.ctypes.data)
    if c < 0:
        raise _abi.NeatError(_abi.NEAT_E_ARG, "neat_tree_update_active_host", _abi.last_error())
    return ids[:c].copy(), prefix


# ---- phantom ------------------------------------------------------------------------------------------------------------
@dataclass
class Phantom:
    center: np.ndarray  # [E,3]
    axes: np.ndarray  # [E,3] semi-axes
    rot: np.ndarray  # [E,3,3] world -> local rotation (rows)
    density: np.ndarray  # [E]

    def density_at(self, p):
        """Sum of densities of the ellipsoids containing the points p [N,3]."""
        out = np.zeros(p.shape[0])
        for c, a, R, rho in zip(self.center, self.axes, self.rot, self.density):
            q = (p - c) @ R.T / a
            out += rho * ((q * q).sum(1) <= 1.0)
        return out

    def density_at_torch(self, p):
        """density_at for a torch tensor p [N,3] on any device (FP64)."""
        import torch
        p = p.to(torch.float64)
        out = torch.zeros(p.shape[0], dtype=torch.float64, device=p.device)
        for c, a, R, rho in zip(self.center, self.axes, self.rot, self.density):
            q = (p - torch.as_tensor(c, device=p.device)) @ torch.as_tensor(R.T, device=p.device)
            q = q / torch.as_tensor(a, device=p.device)
            out += float(rho) * ((q * q).sum(1) <= 1.0).to(torch.float64)
        return out

    def line_integrals_torch(self, origin, direction):
        """line_integrals for torch tensors [R,3] on any device (FP64, the same expressions)."""
        import torch
        dev = origin.device
        o = origin.to(torch.float64)
        d = direction.to(torch.float64)
        total = torch.zeros(o.shape[0], dtype=torch.float64, device=dev)
        dn = torch.linalg.norm(d, dim=1)
        def rot(x, R):
            # x @ R.T as three broadcast products (a 3x3 product is too small for a matrix kernel)
            return (x[:, 0:1] * torch.as_tensor(R[:, 0], device=dev) + x[:, 1:2] * torch.as_tensor(R[:, 1], device=dev)
                    + x[:, 2:3] * torch.as_tensor(R[:, 2], device=dev))

        for c, a, R, rho in zip(self.center, self.axes, self.rot, self.density):
            av = torch.as_tensor(a, device=dev)
            oo = rot(o - torch.as_tensor(c, device=dev), R) / av
            dd = rot(d, R) / av
            A = (dd * dd).sum(1)
            B = 2 * (oo * dd).sum(1)
            C = (oo * oo).sum(1) - 1
            disc = B * B - 4 * A * C
            ok = disc > 0
            sq = torch.sqrt(torch.where(ok, disc, torch.zeros_like(disc)))
            t0 = torch.clamp((-B - sq) / (2 * A), min=0.0)
            t1 = (-B + sq) / (2 * A)
            total += float(rho) * torch.where(ok & (t1 > t0), (t1 - t0) * dn, torch.zeros_like(dn))
        return total

    def line_integrals(self, origin, direction):
        """Closed-form integral of the density along each ray (chord length x density), float64 [R]."""
        o = np.asarray(origin, np.float64)
        d = np.asarray(direction, np.float64)
        total = np.zeros(o.shape[0])
        dn = np.linalg.norm(d, axis=1)
        for c, a, R, rho in zip(self.center, self.axes, self.rot, self.density):
            oo = (o - c) @ R.T / a
            dd = d @ R.T / a
            A = (dd * dd).sum(1)
            B = 2 * (oo * dd).sum(1)
            C = (oo * oo).sum(1) - 1
            disc = B * B - 4 * A * C
            ok = disc > 0
            sq = np.sqrt(np.where(ok, disc, 0))
            t0 = (-B - sq) / (2 * A)
            t1 = (-B + sq) / (2 * A)
            t0 = np.maximum(t0, 0.0)  # rays start at the source
            chord = np.where(ok & (t1 > t0), (t1 - t0) * dn, 0.0)
            total += rho * chord
        return total


def random_phantom(n, seed):
    rng = np.random.default_rng(seed)
    center = rng.uniform(-0.6, 0.6, (n, 3))
    axes = rng.uniform(0.05, 0.35, (n, 3))
    rot = np.empty((n, 3, 3))
    for i in range(n):
        q, r = np.linalg.qr(rng.normal(size=(3, 3)))
        rot[i] = q * np.sign(np.diag(r))
    density = rng.uniform(0.0, 1.0, n)
    return Phantom(center, axes, rot, density)


def adaptive_active(tree, phantom, target_leaves, seed=0, probes=4):
    """Active set refined where the phantom varies: repeatedly split the active node whose density samples
    (probes^3 lattice) have the largest spread times edge length, until target_leaves (or no node can split)."""
    md = tree["max_depth"]
    pmin, pmax, depth = tree["pos_min"], tree["pos_max"], tree["depth"]
    lat = (np.arange(probes) + 0.5) / probes
    grid = np.stack(np.meshgrid(lat, lat, lat, indexing="ij"), -1).reshape(-1, 3)

    def score(n):
        if depth[n] >= md:
            return -1.0
        p = pmin[n] + grid * (pmax[n] - pmin[n])
        v = phantom.density_at(p)
        return float(v.std() * (pmax[n, 0] - pmin[n, 0]) + 1e-9 * (md - depth[n]))

    active = np.zeros(pmin.shape[0], np.int32)
    active[0] = 1
    heap = [(-score(0), 0)]
    leaves = 1
    while heap and leaves + 7 <= target_leaves:
        s, n = heapq.heappop(heap)
        if -s < 0:
            break
        active[n] = 0
        for i in range(8):
            c = 8 * n + 1 + i
            active[c] = 1
            heapq.heappush(heap, (-score(c), c))
        leaves += 7
    return active


# ---- rays ---------------------------------------------------------------------------------------------------------------
def pixel_rays(num_views, det_h, det_w, view, row, col, src_dist=4.0, det_dist=4.0, det_extent=3.0):
    """Rays through detector pixels (view, row, col) (integer arrays of equal length): float32 origin, unit
    direction [N, 3]."""
    view = np.asarray(view)
    th = 2 * np.pi * view / num_views
    s = np.stack([np.sin(-th), -np.cos(-th), np.zeros_like(th)], 1)  # unit source direction
    src = src_dist * s
    cen = -det_dist * s
    eu = np.stack([np.cos(-th), np.sin(-th), np.zeros_like(th)], 1)
    ev = np.array([0.0, 0.0, -1.0])
    u = ((np.asarray(col) + 0.5) / det_w * 2 - 1) * det_extent
    v = ((np.asarray(row) + 0.5) / det_h * 2 - 1) * det_extent
    d = cen + u[:, None] * eu + v[:, None] * ev - src
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    return np.ascontiguousarray(src, np.float32), np.ascontiguousarray(d, np.float32)


def cone_beam_rays(num_views, det_h, det_w, src_dist=4.0, det_dist=4.0, det_extent=3.0, views=None):
    """All rays of the given views in view-major, row-major pixel order."""
    views = np.arange(num_views) if views is None else np.asarray(views)
    v, r, c = np.meshgrid(views, np.arange(det_h), np.arange(det_w), indexing="ij")
    return pixel_rays(num_views, det_h, det_w, v.ravel(), r.ravel(), c.ravel(), src_dist, det_dist, det_extent)


# ---- parameters -----------------------------------------------------------------------------------------------------------
def decoder_params(d_in, hidden, layers, d_out, seed):
    """Packed FCBlock parameters (W [out,in] row-major then b, per layer): Xavier-uniform weights, biases uniform in
    +-1/sqrt(fan_in) (libtorch's Linear default)."""
    rng = np.random.default_rng(seed)
    dims = [d_in] + [hidden] * (layers + 1) + [d_out]
    out = []
    for i in range(len(dims) - 1):
        fi, fo = dims[i], dims[i + 1]
        a = math.sqrt(6.0 / (fi + fo))
        out.append(rng.uniform(-a, a, (fo, fi)).ravel())
        out.append(rng.uniform(-1 / math.sqrt(fi), 1 / math.sqrt(fi), fo))
    return np.concatenate(out).astype(np.float32)


def feature_grid(num_slots, channels, g, seed, std=0.1):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((num_slots, channels, g, g, g)) * std).astype(np.float32)


@dataclass
class Config:
    name: str
    depth: int
    adaptive_leaves: int  # 0 = uniform tree at full depth
    grid: int
    channels: int
    hidden: int
    layers: int
    views: int
    det: int
    max_spn: int
    ellipsoids: int
    batch: int
    act: str = "relu"
    last_act: str = "softplus"
    beta: float = 2.0
    det_extent: float = 3.0
    extra: dict = field(default_factory=dict)


C1 = Config("C1", depth=2, adaptive_leaves=0, grid=11, channels=8, hidden=32, layers=1, views=16, det=128, max_spn=32,
            ellipsoids=20, batch=1 << 18)
C2 = Config("C2", depth=5, adaptive_leaves=4096, grid=17, channels=8, hidden=64, layers=1, views=256, det=512,
            max_spn=64, ellipsoids=100, batch=1 << 18)
# training form: C1 geometry, rays drawn from 256 views; Adam at lr 1e-2 (grid) and 1e-3 (decoder)
C4 = Config("C4", depth=2, adaptive_leaves=0, grid=11, channels=8, hidden=32, layers=1, views=256, det=128,
            max_spn=32, ellipsoids=20, batch=1 << 18, extra=dict(lr_grid=1e-2, lr_decoder=1e-3))


def make_scene(cfg, seed=0, views=None, det=None, phantom_seed=None):
    """Tree, active set, grid and decoder for a configuration (numpy, host)."""
    tree = build_tree(cfg.depth)
    phantom = random_phantom(cfg.ellipsoids, seed + 11 if phantom_seed is None else phantom_seed)
    if cfg.adaptive_leaves:
        active = adaptive_active(tree, phantom, cfg.adaptive_leaves, seed)
    else:
        active = (tree["depth"] == cfg.depth).astype(np.int32)
    ids, prefix = update_active(active)
    slots = max(len(ids), cfg.adaptive_leaves or 0)
    grid = feature_grid(slots, cfg.channels, cfg.grid, seed + 1)
    params = decoder_params(cfg.channels, cfg.hidden, cfg.layers, 1, seed + 2)
    return dict(tree=tree, active=active, active_ids=ids, prefix=prefix, phantom=phantom, grid=grid, params=params,
                mlp=dict(**{"in": cfg.channels}, hidden=cfg.hidden, layers=cfg.layers, out=1, act=cfg.act,
                         last_act=cfg.last_act, beta=cfg.beta))


def batch_rays(cfg, batch, seed=0, det=None):
    """A batch of `batch` rays drawn uniformly (without replacement) over all views and pixels of the
    configuration's detector, sorted by (view, pixel); the whole detector stack when batch covers it."""
    det = det or cfg.det
    total = cfg.views * det * det
    if batch >= total:
        return cone_beam_rays(cfg.views, det, det, det_extent=cfg.det_extent)
    rng = np.random.default_rng(seed + 3)
    idx = np.sort(rng.choice(total, batch, replace=False))
    v, pix = idx // (det * det), idx % (det * det)
    return pixel_rays(cfg.views, det, det, v, pix // det, pix % det, det_extent=cfg.det_extent)


def phantom_field(scene, cfg, floor=0.01, device=None):
    """A grid and decoder whose decoded field is the phantom's density, trilinearly interpolated on every active
    leaf's grid vertices: channel 0 holds softplus_beta^-1(density) (density clamped to >= floor), the other channels
    0, and the decoder passes channel 0 through (relu(x) - relu(-x) in both hidden layers), so that
    field = softplus_beta(x0). Used as a reconstructed-looking field (piecewise smooth, sharp at ellipsoid surfaces)
    for the proposal experiments. Needs relu hidden units."""
    assert cfg.act == "relu"
    tree, ids, prefix = scene["tree"], scene["active_ids"], scene["prefix"]
    G, C, H, beta = cfg.grid, cfg.channels, cfg.hidden, cfg.beta
    grid = np.zeros_like(scene["grid"])
    lin = np.arange(G) / (G - 1)
    zz, yy, xx = np.meshgrid(lin, lin, lin, indexing="ij")
    if device is not None:
        # all leaves at once on the device (torch FP64)
        import torch
        mn = torch.as_tensor(tree["pos_min"][ids].astype(np.float64), device=device)[:, None, :]
        mx = torch.as_tensor(tree["pos_max"][ids].astype(np.float64), device=device)[:, None, :]
        unit = torch.as_tensor(np.stack([xx, yy, zz], -1).reshape(1, -1, 3), device=device)
        pts = (mn + unit * (mx - mn)).reshape(-1, 3)
        rho = torch.clamp(scene["phantom"].density_at_torch(pts), min=floor)
        val = (torch.log(torch.expm1(beta * rho)) / beta).reshape(len(ids), G, G, G).cpu().numpy()
        grid[prefix[ids], 0] = val
        del pts, rho
    else:
        for node in ids:
            mn, mx = tree["pos_min"][node].astype(np.float64), tree["pos_max"][node].astype(np.float64)
            pts = np.stack([mn[0] + xx * (mx[0] - mn[0]), mn[1] + yy * (mx[1] - mn[1]),
                            mn[2] + zz * (mx[2] - mn[2])], -1)
            rho = np.maximum(scene["phantom"].density_at(pts.reshape(-1, 3)), floor)
            grid[prefix[node], 0] = (np.log(np.expm1(beta * rho)) / beta).reshape(G, G, G)
    P = []
    W0 = np.zeros((H, C)); W0[0, 0], W0[1, 0] = 1, -1
    W1 = np.zeros((H, H)); W1[0, 0], W1[1, 1] = 1, 1
    W2 = np.zeros((1, H)); W2[0, 0], W2[0, 1] = 1, -1
    for W in (W0, np.zeros(H), W1, np.zeros(H), W2, np.zeros(1)):
        P.append(np.asarray(W).ravel())
    return grid.astype(np.float32), np.concatenate(P).astype(np.float32)
