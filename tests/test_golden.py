"""Golden fixtures (tests/golden/projector_small.npz, written by tests/golden/make_golden.py from the oracle): a
depth-2 octree, 4^3 x 8 grids, a 8 -> 32 -> 32 -> 1 relu decoder and 64 rays including axis-parallel rays along node
faces, edges and corners, rays starting inside the cube and rays that miss it.

CPU: the oracle reproduces the fixture exactly (segment list, FP32 and FP64 projections and gradients).
GPU: the HIP sampler reproduces the segment list bit for bit; the fused FP32 projector and its gradients are within
1e-5 (relative, norm-wise) of the FP32 oracle's stored values; the BF16 projection is within 1e-3 of the FP64 one."""
import os

import numpy as np
import pytest

from neat_amd import workloads as W

Z = dict(np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "projector_small.npz")))
SEG_KEYS = ("ray_seg_offset", "ray_sample_offset", "seg_node", "seg_ray", "seg_count", "seg_tmin", "seg_tmax")
MLP = {"in": 8, "hidden": 32, "layers": 1, "out": 1, "act": "relu", "last_act": "softplus", "beta": 2.0}


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def host_scene():
    tree = W.build_tree(2)
    ids, prefix = W.update_active(Z["node_active"])
    return tree, ids, prefix


def test_fixture_covers_the_special_cases():
    counts = np.diff(Z["ray_seg_offset"])
    assert (counts == 0).any() and (counts >= 4).any()  # misses and rays crossing several leaves
    # segment entry t is non-decreasing along every ray (canonical order)
    for r in range(len(counts)):
        a, b = Z["ray_seg_offset"][r], Z["ray_seg_offset"][r + 1]
        assert np.all(np.diff(Z["seg_tmin"][a:b]) >= 0)
    assert np.all(Z["seg_count"] >= 1) and np.all(Z["seg_count"] <= int(Z["max_spn"]) + 1)


def test_oracle_reproduces_golden(oracle):
    tree, ids, prefix = host_scene()
    o, d = Z["origin"], Z["direction"]
    segs = oracle.raysample(tree, ids, o, d, int(Z["max_spn"]), emit=False, nthreads=1)
    for k in SEG_KEYS:
        assert np.array_equal(segs[k], Z[k]), k
    mlp = dict(MLP, params=Z["params"])
    for prec, tag in ((0, "fp32"), (1, "fp64")):
        p = oracle.projector_fwd(tree, prefix, o, d, segs, Z["grid"], mlp, jitter=True, seed=int(Z["jitter_seed"]),
                                 prec=prec, nthreads=1)
        gg, gp = oracle.projector_bwd(tree, prefix, o, d, segs, Z["grid"], mlp, Z["grad_proj"].astype(np.float64),
                                      jitter=True, seed=int(Z["jitter_seed"]), prec=prec, nthreads=1)
        dt = np.float32 if prec == 0 else np.float64
        assert np.array_equal(p.astype(dt), Z[f"proj_{tag}"])
        assert np.array_equal(gg.astype(dt), Z[f"grad_grid_{tag}"])
        assert np.array_equal(gp.astype(dt), Z[f"grad_params_{tag}"])


@pytest.mark.gpu
def test_hip_path_matches_golden():
    import torch

    from neat_amd import ops, pipeline

    print("case: golden fixture projector_small (64 rays, depth-2 tree)", flush=True)
    tree, ids, prefix = host_scene()
    cfg = W.Config("golden", depth=2, adaptive_leaves=0, grid=4, channels=8, hidden=32, layers=1, views=1, det=1,
                   max_spn=int(Z["max_spn"]), ellipsoids=0, batch=64)
    scene = dict(tree=tree, active=Z["node_active"], active_ids=ids, prefix=prefix, grid=Z["grid"],
                 params=Z["params"], mlp=dict(MLP))
    ds = pipeline.to_device(scene, cfg)
    to = torch.from_numpy(Z["origin"]).cuda()
    td = torch.from_numpy(Z["direction"]).cuda()
    seed = int(Z["jitter_seed"])
    segs = pipeline.sample(ds, to, td, jitter=True, seed=seed, emit=False)
    for k in SEG_KEYS:
        got = segs[k].cpu().numpy()
        assert np.array_equal(got, Z[k]), k
    g = torch.from_numpy(Z["grad_proj"]).cuda()
    for precision in ("fp32", "bf16"):
        pj = ops.Projector(ds.pos_min, ds.pos_max, ds.prefix, to, td, segs, ds.grid, ds.params, ds.mlp, True, seed,
                           precision=precision)
        proj = pj.forward().cpu().numpy()
        gg, gp = pj.backward(g)
        gg, gp = gg.cpu().numpy(), gp.cpu().numpy()
        if precision == "fp32":
            e = [rel(proj, Z["proj_fp32"]), rel(gg, Z["grad_grid_fp32"]), rel(gp, Z["grad_params_fp32"])]
            print("fp32 vs golden fp32 oracle (proj, grid, params)", e)
            assert max(e) <= 1e-5
        else:
            e = rel(proj, Z["proj_fp64"])
            print("bf16 projection vs golden fp64 oracle", e)
            assert e <= 1e-3


C2_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "projector_c2.npz")


@pytest.mark.gpu
def test_hip_path_matches_golden_full_size():
    """tests/golden/projector_c2.npz: 4096 rays of the full-size configuration (regenerated from its seeds); the HIP
    sampler reproduces the segment list bit for bit, the fused FP32 projector and its gradients are within 1e-5 of
    the FP32 oracle's stored values, the BF16 projection within 1e-3 of the FP64 one."""
    import torch

    from neat_amd import ops, pipeline

    print("case: golden fixture projector_c2 (4096 rays, full-size configuration)", flush=True)
    Zc = dict(np.load(C2_PATH))
    cfg = W.C2
    sc = W.make_scene(cfg)
    ds = pipeline.to_device(sc, cfg)
    to = torch.from_numpy(Zc["origin"]).cuda()
    td = torch.from_numpy(Zc["direction"]).cuda()
    seed = int(Zc["jitter_seed"])
    segs = pipeline.sample(ds, to, td, jitter=True, seed=seed, emit=False)
    for k in SEG_KEYS:
        assert np.array_equal(segs[k].cpu().numpy(), Zc[k]), k
    g = torch.from_numpy(Zc["grad_proj"]).cuda()
    pj = ops.Projector(ds.pos_min, ds.pos_max, ds.prefix, to, td, segs, ds.grid, ds.params, ds.mlp, True, seed)
    proj = pj.forward().cpu().numpy()
    gg, gp = pj.backward(g)
    gnorm = gg.reshape(gg.shape[0], -1).double().pow(2).sum(1).sqrt().cpu().numpy()
    gpick = gg.contiguous().reshape(-1)[torch.from_numpy(Zc["grid_pick"]).cuda()].cpu().numpy()
    gp = gp.cpu().numpy()
    e = [rel(proj, Zc["proj_fp32"]), rel(gnorm, Zc["grad_grid_slot_norm_fp32"]), rel(gpick, Zc["grad_grid_pick_fp32"])]
    print("fp32 vs golden fp32 oracle (proj, grid slot norms, grid entries)", e)
    assert max(e) <= 1e-5
    # the decoder gradient sums about 0.6 M samples' terms: both FP32 sums (GPU per-wave partials, the oracle's
    # per-thread partials) carry their own rounding, so each is held to the FP64 oracle
    e64 = [rel(gp, Zc["grad_params_fp64"]), rel(Zc["grad_params_fp32"], Zc["grad_params_fp64"]),
           rel(proj, Zc["proj_fp64"]), rel(gpick, Zc["grad_grid_pick_fp64"])]
    print("vs fp64 oracle: gpu params, fp32-oracle params, gpu proj, gpu grid entries", e64)
    # sampled grid entries sum many signed terms: FP32 sits about 4e-3 from FP64 there (the FP32 oracle too)
    assert e64[0] <= 1e-4 and e64[2] <= 1e-5 and e64[3] <= 1e-2
    pb = ops.Projector(ds.pos_min, ds.pos_max, ds.prefix, to, td, segs, ds.grid, ds.params, ds.mlp, True, seed,
                       precision="bf16")
    eb = rel(pb.forward().cpu().numpy(), Zc["proj_fp64"])
    print("bf16 projection vs golden fp64 oracle", eb)
    assert eb <= 1e-3
