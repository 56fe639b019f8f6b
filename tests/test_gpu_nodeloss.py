"""GPU checks of the evaluation pass's per-node loss (neat_node_loss_accumulate and the module layer's
AccumulateSampleLossPerNode / SetErrorForActiveNodes) against oracle/nodeloss.py: per-node loss and weight sums
within 1e-6 (relative, norm-wise) of the float64 oracle on the same canonical samples; node_error updates equal to
the reference's expressions."""
import numpy as np
import pytest
import torch

from neat_amd import ops, pipeline, workloads as W
from oracle import nodeloss as NL

pytestmark = pytest.mark.gpu
DEV = "cuda"


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("cfg_name,nrays,jitter", [("C1", 4000, True), ("C2", 600, False)])
def test_node_loss_vs_oracle(oracle, cfg_name, nrays, jitter):
    print("case: node loss", cfg_name, nrays, jitter, flush=True)
    cfg = getattr(W, cfg_name)
    sc = W.make_scene(cfg, seed=2)
    ds = pipeline.to_device(sc, cfg)
    o, d = W.batch_rays(cfg, nrays, seed=5)
    to, td = torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV)
    segs = pipeline.sample(ds, to, td, jitter=jitter, seed=4, emit=False)
    loss = np.random.default_rng(1).uniform(size=(1, nrays)).astype(np.float32)
    nn = sc["tree"]["pos_min"].shape[0]
    ls, ws = ops.node_loss_accumulate(segs, torch.from_numpy(loss[0]).to(DEV), nn, jitter=jitter, seed=4)
    smp = oracle.raysample(sc["tree"], sc["active_ids"], o, d, cfg.max_spn, jitter=jitter, seed=4, emit=True)
    rls, rws = NL.accumulate(smp, loss, nn)
    e = (rel(ls.cpu().numpy(), rls), rel(ws.cpu().numpy(), rws))
    print("node loss / weight sums vs oracle", e)
    assert max(e) <= 1e-6


def test_module_layer_node_error():
    import neat_amd.host as host
    H = host.load()
    print("case: AccumulateSampleLossPerNode + SetErrorForActiveNodes", flush=True)
    cfg = W.C1
    sc = W.make_scene(cfg, seed=3)
    tree = H.HyperTreeBase(3, cfg.depth)
    tree.node_active.copy_(torch.from_numpy(sc["active"]))
    tree.UpdateActive()
    tree.to(DEV)
    p = H.CombinedParams()
    p.net_params.grid_size = cfg.grid
    p.net_params.grid_features = cfg.channels
    p.octree_params.tree_optimizer_params.max_active_nodes = sc["grid"].shape[0]
    geo = H.GeometryExEx(1, 3, tree, p)
    geo.to(DEV)
    o, d = W.batch_rays(cfg, 3000, seed=9)
    rays = H.RayList(torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV))
    samples = tree.CreateSamplesForRays(rays, cfg.max_spn, False)
    per_ray = torch.from_numpy(np.random.default_rng(2).uniform(size=(1, 3000)).astype(np.float32)).to(DEV)
    ls, ws = geo.AccumulateSampleLossPerNode(samples, per_ray)
    ray_index = samples.ray_index.cpu().numpy()
    smp = dict(weight=samples.weight.cpu().numpy(), node_id=samples.node_id.cpu().numpy(), ray_index=ray_index)
    nn = tree.NumNodes()
    rls, rws = NL.accumulate(smp, per_ray.cpu().numpy(), nn)
    e = (rel(ls.cpu().numpy(), rls), rel(ws.cpu().numpy(), rws))
    print("module layer node loss vs oracle on the emitted samples", e)
    assert max(e) <= 1e-6
    active = tree.node_active.cpu().numpy()
    before = tree.node_error.cpu().numpy().copy()
    tree.SetErrorForActiveNodes(ls, "override")
    after1 = tree.node_error.cpu().numpy().copy()
    assert np.array_equal(after1, NL.set_error(before, active, ls.cpu().numpy(), "override"))
    half = (ls * 0.5).contiguous()
    tree.SetErrorForActiveNodes(half, "min")
    assert np.array_equal(tree.node_error.cpu().numpy(), NL.set_error(after1, active, half.cpu().numpy(), "min"))
