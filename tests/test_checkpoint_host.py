"""Checkpoint loading on the CPU (reconstruct.cpp:36-62): a file is checked completely before anything changes, and
loading copies values into the live tensors, so their layouts stay (the reference saves its grid as plain NCDHW,
ImplicitNet.h:79-119, while this build keeps it channels-last-3d for the fused projector)."""
import pytest
import torch

from neat_amd import workloads as W


@pytest.fixture(scope="module")
def H():
    import neat_amd.host as host
    return host.load()


def make(H, depth, grid=11, active_depth=None, seed=0):
    torch.manual_seed(seed)
    t = H.HyperTreeBase(3, depth)
    t.SetActive(depth if active_depth is None else active_depth)
    p = H.CombinedParams()
    p.net_params.grid_size = grid
    p.net_params.grid_features = 8
    p.net_params.decoder_hidden_features = 32
    p.net_params.decoder_hidden_layers = 1
    p.net_params.decoder_activation = "relu"
    p.octree_params.tree_optimizer_params.max_active_nodes = 8 ** depth
    p.train_params.grid_init = "zero"
    geo = H.GeometryExEx(1, 3, t, p)
    with torch.no_grad():
        geo.grid.copy_(torch.randn(geo.grid.shape))
    return t, geo


def test_ncdhw_grid_file_loads_channels_last(H, tmp_path):
    t, geo = make(H, 2, seed=1)
    want = geo.grid.detach().clone()
    # a file holding the grid as the reference stores it: contiguous [N, C, Z, Y, X]
    geo.grid.data = geo.grid.data.contiguous()
    assert geo.grid.is_contiguous()
    prefix = str(tmp_path / "ref_")
    H.SaveCheckpoint(geo, t, prefix)
    t2, geo2 = make(H, 2, seed=2)
    strides = geo2.grid.stride()
    H.LoadCheckpoint(geo2, t2, prefix)
    assert geo2.grid.stride() == strides and geo2.grid.is_contiguous(memory_format=torch.channels_last_3d)
    assert torch.equal(geo2.grid.detach(), want)
    assert geo2.FusedUnsupportedReason() == ""
    for (k1, a), (k2, b) in zip(geo.named_parameters(), geo2.named_parameters()):
        assert k1 == k2 and torch.equal(a.detach(), b.detach()), k1


def test_tree_of_another_depth_is_refused_and_nothing_changes(H, tmp_path):
    t1, geo1 = make(H, 1)
    prefix = str(tmp_path / "d1_")
    H.SaveCheckpoint(geo1, t1, prefix)
    t2, geo2 = make(H, 2, active_depth=1)
    before = {k: v.clone() for k, v in t2.named_buffers()}
    flags = t2.node_flags.clone()
    grid = geo2.grid.detach().clone()
    with pytest.raises(RuntimeError, match="has shape"):
        H.LoadCheckpoint(geo2, t2, prefix)
    assert t2.NumNodes() == 73 and t2.MaxDepth() == 2
    for k, v in t2.named_buffers():
        assert torch.equal(v, before[k]), k
    assert torch.equal(t2.node_flags, flags) and torch.equal(geo2.grid.detach(), grid)


def test_geometry_of_another_width_is_refused_and_nothing_changes(H, tmp_path):
    """A geometry file whose decoder is narrower than the live one fails its shape check before the tree or the grid
    are written."""
    t, geo = make(H, 1)
    prefix = str(tmp_path / "m_")
    H.SaveCheckpoint(geo, t, prefix)
    t2, _ = make(H, 1, seed=3)
    p = H.CombinedParams()
    p.net_params.grid_size = 11
    p.net_params.grid_features = 8
    p.net_params.decoder_hidden_features = 64
    p.net_params.decoder_hidden_layers = 1
    p.net_params.decoder_activation = "relu"
    p.octree_params.tree_optimizer_params.max_active_nodes = 8
    p.train_params.grid_init = "zero"
    geo_wide = H.GeometryExEx(1, 3, t2, p)
    grid = geo_wide.grid.detach().clone()
    with torch.no_grad():
        t2.node_error.fill_(7.0)
    with pytest.raises(RuntimeError, match="has shape"):
        H.LoadCheckpoint(geo_wide, t2, prefix)
    assert torch.equal(geo_wide.grid.detach(), grid)
    assert bool((t2.node_error == 7.0).all())
