"""Writes projector_small.npz beside this script: a small case of the projector path (depth-2 uniform octree, 4^3 x 8
grids, decoder 8 -> 32 -> 32 -> 1 relu / softplus, 64 rays including axis-parallel rays, rays along node faces,
edges and corners, rays starting inside the cube and rays that miss it) with the oracle's results: the canonical
segment list, the FP32 and FP64 projections with jitter, and the gradients of sum(g * projection) for a fixed g.

The reference ships no test vectors, so these are the oracle's own outputs, frozen: a later change to the oracle or
to the HIP path shows up as a difference against them. Regenerate (only on purpose) from the repository root with
    python tests/golden/make_golden.py [small] [c2]

projector_c2.npz holds the same kind of results for 4096 rays of the full-size configuration (make_c2).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from neat_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402

MAX_SPN = 8
JITTER_SEED = 5


def special_rays():
    o, d = [], []
    # axis-parallel through node faces (x = 0, y = 0.5), edges and corners, and through cell interiors
    for a in range(3):
        for off in [(0.0, 0.0), (0.5, 0.5), (0.25, -0.75), (1.0, 0.0), (-0.5, 0.3)]:
            p = np.zeros(3)
            p[a] = -3.0
            q = [k for k in range(3) if k != a]
            p[q[0]], p[q[1]] = off
            dd = np.zeros(3)
            dd[a] = 1.0 if len(o) % 2 == 0 else -1.0
            if dd[a] < 0:
                p[a] = 3.0
            o.append(p)
            d.append(dd)
    # diagonals through the cube's corners and centre, origin inside the cube, grazing and missing rays
    o += [np.array([-2.0, -2.0, -2.0]), np.array([2.0, -2.0, 2.0]), np.array([0.1, 0.2, 0.3]),
          np.array([0.0, 0.0, 0.0]), np.array([-3.0, 1.0, 0.0]), np.array([-3.0, 1.5, 0.0]),
          np.array([-3.0, 0.999, 0.999])]
    d += [np.array([1.0, 1.0, 1.0]), np.array([-1.0, 1.0, -1.0]), np.array([0.3, -0.4, 0.5]),
          np.array([1.0, 0.0, 0.0]), np.array([1.0, 0.0, 0.0]), np.array([1.0, 0.0, 0.0]),
          np.array([1.0, 0.0, 0.0])]
    return o, d


def make():
    o, d = special_rays()
    rng = np.random.default_rng(17)
    while len(o) < 64:
        p = rng.uniform(-3, 3, 3)
        t = rng.uniform(-0.9, 0.9, 3)
        o.append(p)
        d.append(t - p)
    o = np.array(o, np.float32)
    d = np.array(d, np.float64)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)

    tree = W.build_tree(2)
    active = (tree["depth"] == 2).astype(np.int32)
    ids, prefix = W.update_active(active)
    grid = W.feature_grid(len(ids), 8, 4, seed=23, std=0.3)
    params = W.decoder_params(8, 32, 1, 1, seed=29)
    mlp = {"in": 8, "hidden": 32, "layers": 1, "out": 1, "act": "relu", "last_act": "softplus", "beta": 2.0,
           "params": params}
    segs = O.raysample(tree, ids, o, d, MAX_SPN, emit=False, nthreads=1)
    gproj = rng.normal(size=o.shape[0]).astype(np.float32)
    out = dict(origin=o, direction=d, node_active=active, grid=grid, params=params, grad_proj=gproj,
               max_spn=np.int32(MAX_SPN), jitter_seed=np.int64(JITTER_SEED))
    for k in ("ray_seg_offset", "ray_sample_offset", "seg_node", "seg_ray", "seg_count", "seg_tmin", "seg_tmax"):
        out[k] = segs[k]
    for prec, tag in ((0, "fp32"), (1, "fp64")):
        p = O.projector_fwd(tree, prefix, o, d, segs, grid, mlp, jitter=True, seed=JITTER_SEED, prec=prec, nthreads=1)
        gg, gp = O.projector_bwd(tree, prefix, o, d, segs, grid, mlp, gproj.astype(np.float64), jitter=True,
                                 seed=JITTER_SEED, prec=prec, nthreads=1)
        if prec == 0:  # the FP32 restatement's results are float32 values
            p, gg, gp = p.astype(np.float32), gg.astype(np.float32), gp.astype(np.float32)
        out[f"proj_{tag}"], out[f"grad_grid_{tag}"], out[f"grad_params_{tag}"] = p, gg, gp
    return out


C2_RAYS = 4096
C2_SEED = 31
C2_GRID_ENTRIES = 4096


def make_c2():
    """The full-size configuration (W.C2: depth-5 tree adapted to the phantom, 4096 leaves, 17^3 x 8 grids, decoder
    8 -> 64 -> 64 -> 1) rebuilt from its seeds, on 4096 of its rays with jitter: the canonical segment list, and from
    the FP32 and the FP64 oracle the projections, the decoder gradient and, of the 644 MB grid gradient, the per-slot
    L2 norms and 4096 sampled entries. The scene itself is not stored: W.make_scene(W.C2) regenerates it."""
    cfg = W.C2
    sc = W.make_scene(cfg)
    o, d = W.batch_rays(cfg, C2_RAYS, seed=C2_SEED)
    segs = O.raysample(sc["tree"], sc["active_ids"], o, d, cfg.max_spn, emit=False)
    rng = np.random.default_rng(C2_SEED)
    gproj = rng.normal(size=o.shape[0]).astype(np.float32)
    pick = rng.choice(sc["grid"].size, C2_GRID_ENTRIES, replace=False).astype(np.int64)
    mlp = dict(sc["mlp"], params=sc["params"])
    out = dict(origin=o, direction=d, grad_proj=gproj, grid_pick=pick, jitter_seed=np.int64(C2_SEED))
    for k in ("ray_seg_offset", "ray_sample_offset", "seg_node", "seg_ray", "seg_count", "seg_tmin", "seg_tmax"):
        out[k] = segs[k]
    for prec, tag in ((0, "fp32"), (1, "fp64")):
        out[f"proj_{tag}"] = O.projector_fwd(sc["tree"], sc["prefix"], o, d, segs, sc["grid"], mlp, jitter=True,
                                             seed=C2_SEED, prec=prec).astype(np.float32 if prec == 0 else np.float64)
    for prec, tag in ((0, "fp32"), (1, "fp64")):
        gg, gp = O.projector_bwd(sc["tree"], sc["prefix"], o, d, segs, sc["grid"], mlp, gproj.astype(np.float64),
                                 jitter=True, seed=C2_SEED, prec=prec, nthreads=min(8, O.max_threads()))
        dt = np.float32 if prec == 0 else np.float64
        gg = gg.astype(dt)
        out[f"grad_params_{tag}"] = gp.astype(dt)
        out[f"grad_grid_slot_norm_{tag}"] = np.sqrt((gg.reshape(gg.shape[0], -1).astype(np.float64) ** 2).sum(1))
        out[f"grad_grid_pick_{tag}"] = gg.reshape(-1)[pick]
        del gg
    return out


if __name__ == "__main__":
    which = sys.argv[1:] or ["small", "c2"]
    if "small" in which:
        out = make()
        path = os.path.join(HERE, "projector_small.npz")
        np.savez_compressed(path, **out)
        print(path, {k: getattr(v, "shape", ()) for k, v in out.items()})
    if "c2" in which:
        out = make_c2()
        path = os.path.join(HERE, "projector_c2.npz")
        np.savez_compressed(path, **out)
        print(path, {k: getattr(v, "shape", ()) for k, v in out.items()})
