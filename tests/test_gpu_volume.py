"""GPU checks of the volume export (HyperTreeBaseImpl::UniformPhantomSamplesGPU and
HierarchicalNeuralGeometry::UniformSampledVolume): voxel coordinates bit for bit against the reference's float
expression, node ids and validity bit for bit against the oracle's table walk (oracle/regularizer.py), and the
decoded values within 1e-5 of the FP64 oracle (local coordinates, grid sample, decoder)."""
import numpy as np
import pytest
import torch

from neat_amd import workloads as W
from oracle import regularizer as R

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.fixture(scope="module")
def H():
    import neat_amd.host as host
    return host.load()


def lattice(shape, swap_xy):
    n = int(np.prod(shape))
    c = np.zeros((n, 3), np.int64)
    tmp = np.arange(n)
    dims = range(2, -1, -1) if swap_xy else range(3)
    for d in dims:
        c[:, d] = tmp % shape[d]
        tmp = tmp // shape[d]
    pos = c.astype(np.float32) / (np.asarray(shape, np.float32) - np.float32(1))
    return pos * np.float32(2) - np.float32(1)


@pytest.mark.parametrize("cfg_name,swap_xy", [("C1", False), ("C1", True), ("C2", False)])
def test_uniform_samples_vs_oracle(H, cfg_name, swap_xy):
    from test_gpu_host import make_geometry
    print("case: uniform volume samples", cfg_name, swap_xy, flush=True)
    cfg = getattr(W, cfg_name)
    sc, tree, _, geo = make_geometry(H, cfg, seed=4)
    shape = [13, 9, 6]
    smp = tree.UniformPhantomSamplesGPU(shape, swap_xy)
    coord = lattice(shape, swap_xy)
    assert np.array_equal(smp.global_coordinate.cpu().numpy().view(np.int32), coord.view(np.int32))
    node = R.active_node_at(sc["tree"], sc["active"], coord)
    want = np.where(node < 0, sc["active_ids"][0], node)
    assert np.array_equal(smp.node_id.cpu().numpy(), want)
    assert np.array_equal(smp.weight.cpu().numpy(), (node >= 0).astype(np.float32))
    assert np.array_equal(smp.ray_index.cpu().numpy(), np.arange(coord.shape[0]))


def test_uniform_sampled_volume_vs_oracle(H, oracle):
    from test_gpu_host import make_geometry
    print("case: UniformSampledVolume C1", flush=True)
    cfg = W.C1
    sc, tree, _, geo = make_geometry(H, cfg, seed=6)
    shape = [17, 11, 8]
    vol, vnode, valid = geo.UniformSampledVolume(shape, 1)
    assert list(vol.shape) == [1] + shape and list(vnode.shape) == shape and list(valid.shape) == shape
    coord = lattice(shape, False)
    node = R.active_node_at(sc["tree"], sc["active"], coord)
    ok = node >= 0
    assert np.array_equal(valid.numpy().reshape(-1), ok.astype(np.int64))
    # outside every active node: the first active node, as the reference writes it (geometry.cpp:228-233)
    assert np.array_equal(vnode.numpy().reshape(-1), np.where(ok, node, sc["active_ids"][0]))
    uv = oracle.local_samples(sc["tree"], coord[ok], node[ok])
    f = oracle.grid_sample_fwd(sc["grid"].astype(np.float64), sc["prefix"][node[ok]], uv.astype(np.float64))
    y = oracle.decoder_fwd(f, sc["params"], cfg.channels, cfg.hidden, cfg.layers, 1, act=cfg.act,
                           last_act="softplus", beta=cfg.beta)
    got = vol.numpy().reshape(-1)[ok]
    e = float(np.linalg.norm(got - y.reshape(-1)) / np.linalg.norm(y))
    print("volume vs oracle rel", e, "valid", int(ok.sum()), "of", ok.size)
    assert e <= 1e-5 and np.all(vol.numpy().reshape(-1)[~ok] == 0)
