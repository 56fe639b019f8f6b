"""Checkpoints (reconstruct.cpp:36-62): SaveCheckpoint / LoadCheckpoint round trip of the geometry and the tree
through torch::save / torch::load, under the reference's parameter and buffer names; the restored modules equal
the saved ones tensor for tensor and project the same image."""
import numpy as np
import pytest
import torch

from neat_amd import workloads as W

pytestmark = pytest.mark.gpu
DEV = "cuda"

# HyperTree.cpp:43-57 and ImplicitNet.h:109 of the reference
TREE_BUFFERS = {"node_parent", "node_children", "node_position_min", "node_position_max", "node_scale", "node_error",
                "node_max_density", "node_active", "node_culled", "node_depth", "node_active_id", "active_node_ids"}


@pytest.fixture(scope="module")
def H():
    import neat_amd.host as host
    return host.load()


def test_checkpoint_round_trip(H, tmp_path):
    from test_gpu_host import make_geometry
    print("case: checkpoint round trip C1", flush=True)
    cfg = W.C1
    sc, tree, _, geo = make_geometry(H, cfg, seed=1)
    with torch.no_grad():
        tree.node_error.copy_(torch.rand_like(tree.node_error))
    assert TREE_BUFFERS <= {k for k, _ in tree.named_buffers()}
    assert "explicit_grid_generator.grid_data" in {k for k, _ in geo.named_parameters()}
    prefix = str(tmp_path / "scene_")
    H.SaveCheckpoint(geo, tree, prefix)
    sc2, tree2, _, geo2 = make_geometry(H, cfg, seed=9)
    H.LoadCheckpoint(geo2, tree2, prefix)
    for (k1, a), (k2, b) in zip(geo.named_parameters(), geo2.named_parameters()):
        assert k1 == k2 and torch.equal(a, b), k1
    for (k1, a), (k2, b) in zip(tree.named_buffers(), tree2.named_buffers()):
        assert k1 == k2 and a.dtype == b.dtype and torch.equal(a.cpu(), b.cpu()), k1
    o, d = W.batch_rays(cfg, 2048, seed=3)
    rays = H.RayList(torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV))
    with torch.no_grad():
        i1 = geo.ComputeImage(tree.CreateSamplesForRays(rays, cfg.max_spn, False), 1, 2048)
        i2 = geo2.ComputeImage(tree2.CreateSamplesForRays(rays, cfg.max_spn, False), 1, 2048)
    assert torch.equal(i1, i2)


def test_checkpoint_with_int32_active_prefix_loads(H, tmp_path):
    """The reference registers node_active_id as int32 (HyperTree.cpp:41-53): a tree file holding it that way loads, and
    the restored tree's derived state is rebuilt in the dtypes the kernels use."""
    from test_gpu_host import make_geometry
    print("case: checkpoint with an int32 node_active_id", flush=True)
    cfg = W.C1
    sc, tree, _, geo = make_geometry(H, cfg, seed=1)
    want = {k: v.clone() for k, v in tree.named_buffers()}
    buf = dict(tree.named_buffers())["node_active_id"]
    buf.data = buf.data.to(torch.int32)
    prefix = str(tmp_path / "ref_")
    H.SaveCheckpoint(geo, tree, prefix)
    sc2, tree2, _, geo2 = make_geometry(H, cfg, seed=9)
    H.LoadCheckpoint(geo2, tree2, prefix)
    got = dict(tree2.named_buffers())
    assert got["node_active_id"].dtype == torch.int64
    for k, v in want.items():
        assert torch.equal(got[k].cpu(), v.cpu()), k
    o, d = W.batch_rays(cfg, 1024, seed=4)
    rays = H.RayList(torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV))
    with torch.no_grad():
        i1 = geo2.ComputeImage(tree2.CreateSamplesForRays(rays, cfg.max_spn, False), 1, 1024)
    assert bool(torch.isfinite(i1).all())
