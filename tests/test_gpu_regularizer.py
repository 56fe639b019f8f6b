"""GPU checks of the volume regulariser (neat_neighbor_pairs, neat_edge_loss_*, neat_tv_loss_*) against the CPU
restatement in oracle/regularizer.py: the neighbour rows bit for bit (canonical order), the losses and gradients
within 1e-6 (relative) of the float64 oracle, and the composed regulariser (pairs -> local coordinates -> grid sample
-> edge loss, plus total variation) within 1e-5, both as that chain of steps and as the fused one-pass edge loss
(neat_edge_loss_direct), which must also agree with the chain to FP64 summation order (loss) and float atomic order
(gradient)."""
import numpy as np
import pytest
import torch

from neat_amd import ops, workloads as W
from oracle import oracle as O, regularizer as R

pytestmark = pytest.mark.gpu
DEV = "cuda"


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def scene(depth, adaptive, G, seed=0):
    tree = W.build_tree(depth)
    if adaptive:
        active = W.adaptive_active(tree, W.random_phantom(20, seed + 3), adaptive, seed)
    else:
        active = (tree["depth"] == depth).astype(np.int32)
    ids, prefix = W.update_active(active)
    grid = W.feature_grid(len(ids) + 3, 8, G, seed + 1, std=0.3)
    dev = dict(pos_min=torch.from_numpy(tree["pos_min"]).to(DEV), pos_max=torch.from_numpy(tree["pos_max"]).to(DEV),
               flags=ops.tree_flags(active, depth).to(DEV), prefix=torch.from_numpy(prefix).to(DEV),
               ids=torch.from_numpy(ids).to(DEV),
               grid=torch.from_numpy(grid).to(DEV).contiguous(memory_format=torch.channels_last_3d))
    return tree, active, ids, prefix, grid, dev


@pytest.mark.parametrize("depth,adaptive,G", [(2, 0, 5), (3, 60, 4), (5, 400, 6)])
def test_neighbor_rows_bit_exact(depth, adaptive, G):
    print("case: neighbour rows", depth, adaptive, G, flush=True)
    tree, active, ids, prefix, grid, d = scene(depth, adaptive, G)
    smp = ops.neighbor_samples(d["pos_min"], d["pos_max"], d["flags"], d["ids"], G)
    coord, node = R.neighbor_pairs(tree, active, ids, G)
    assert smp["num_pairs"] * 2 == node.shape[0]
    assert np.array_equal(smp["node_id"].cpu().numpy(), node)
    assert np.array_equal(smp["global_coordinate"].cpu().numpy(), coord)
    n = node.shape[0]
    assert np.array_equal(smp["ray_index"].cpu().numpy(), np.arange(n))
    assert np.array_equal(smp["local_index_in_ray"].cpu().numpy(), np.arange(n) % 2)
    assert np.all(smp["weight"].cpu().numpy() == 1.0)


def test_tv_and_edge_kernels_match_oracle():
    print("case: tv and edge kernels", flush=True)
    tree, active, ids, prefix, grid, d = scene(3, 60, 5, seed=4)
    slots = prefix[ids]
    w = 1e-4
    loss = ops.tv_loss_fwd(d["grid"], torch.from_numpy(slots).to(DEV), w).item()
    gg = torch.zeros_like(d["grid"])
    ops.tv_loss_bwd(d["grid"], torch.from_numpy(slots).to(DEV), w, torch.ones(1, device=DEV), gg)
    rl, rg = R.tv_loss(grid, slots, w)
    print("tv loss", loss, rl, "grad rel", rel(gg.cpu().numpy(), rg))
    assert abs(loss - rl) <= 1e-6 * abs(rl) and rel(gg.cpu().numpy(), rg) <= 1e-6
    f = np.random.default_rng(3).normal(size=(2 * 1000, 8)).astype(np.float32)
    ft = torch.from_numpy(f).to(DEV)
    el = ops.edge_loss_fwd(ft, 7, 1e-3).item()
    eg = ops.edge_loss_bwd(ft, 7, 1e-3, torch.ones(1, device=DEV)).cpu().numpy()
    rl, rg = R.edge_loss(f, 7, 1e-3)
    print("edge loss", el, rl, "grad rel", rel(eg, rg))
    assert abs(el - rl) <= 1e-6 * abs(rl) and rel(eg, rg) <= 1e-6


@pytest.mark.parametrize("fused,channels_last", [(True, True), (False, True), (True, False)])
def test_volume_regularizer_matches_oracle_composition(fused, channels_last):
    print("case: composed regulariser fused", fused, "channels_last", channels_last, flush=True)
    tree, active, ids, prefix, grid, d = scene(3, 60, 5, seed=6)
    if not channels_last:
        d["grid"] = d["grid"].contiguous()
    wt, we, seed = 1e-4, 1e-3, 9
    fn = ops.volume_regularizer if fused else ops.volume_regularizer_unfused
    loss, gg = fn(d["pos_min"], d["pos_max"], d["flags"], d["prefix"], d["ids"], d["grid"], loss_tv=wt, loss_edge=we,
                  seed=seed)
    coord, node = R.neighbor_pairs(tree, active, ids, grid.shape[2])
    uv = O.local_samples(tree, coord, node)
    slot = prefix[node]
    f = O.grid_sample_fwd(grid.astype(np.float64), slot, uv.astype(np.float64))
    el, ef = R.edge_loss(f, seed, we)
    eg, _ = O.grid_sample_bwd(grid.astype(np.float64), slot, uv.astype(np.float64), ef)
    tl, tg = R.tv_loss(grid, prefix[ids], wt)
    e_loss = abs(loss.item() - (el + tl)) / abs(el + tl)
    e_grad = rel(gg.cpu().numpy(), eg + tg)
    print("regulariser loss", loss.item(), el + tl, "rel", e_loss, "grad rel", e_grad)
    assert e_loss <= 1e-5 and e_grad <= 1e-5


@pytest.mark.parametrize("depth,adaptive,G", [(3, 60, 5), (5, 400, 6), (2, 0, 4)])
def test_fused_edge_loss_equals_chain(depth, adaptive, G):
    """neat_edge_loss_direct against the unfused chain on the same inputs: loss to 1e-7 (the chain sums a pair's
    channels in FP32 before its FP64 sum, the fused pass adds every channel in FP64), gradient to 1e-6 (float atomics
    in another order), and the loss-only and gradient-only modes equal the combined call."""
    print("case: fused edge loss vs chain", depth, adaptive, G, flush=True)
    tree, active, ids, prefix, grid, d = scene(depth, adaptive, G, seed=2)
    args = (d["pos_min"], d["pos_max"], d["flags"], d["prefix"], d["ids"], d["grid"])
    l1, g1 = ops.volume_regularizer(*args, loss_tv=0.0, loss_edge=1e-3, seed=5)
    l0, g0 = ops.volume_regularizer_unfused(*args, loss_tv=0.0, loss_edge=1e-3, seed=5)
    e_loss = abs(l1.item() - l0.item()) / abs(l0.item())
    e_grad = rel(g1.cpu().numpy(), g0.cpu().numpy())
    print("fused vs chain: loss", l1.item(), l0.item(), "rel", e_loss, "grad rel", e_grad)
    assert e_loss <= 1e-7 and e_grad <= 1e-6
    lo = ops.edge_loss_direct(*args, seed=5, weight=1e-3)
    assert lo.item() == l1.item()
    gg = torch.zeros_like(d["grid"])
    none = ops.edge_loss_direct(*args, seed=5, weight=1e-3, want_loss=False, grad_loss=torch.ones(1, device=DEV),
                                grad_grid=gg)
    assert none is None and rel(gg.cpu().numpy(), g1.cpu().numpy()) <= 1e-6


@pytest.mark.parametrize("channels_last", [True, False])
def test_tv_fwd_bwd_equals_separate_passes(channels_last):
    """neat_tv_loss_fwd_bwd against neat_tv_loss_fwd + neat_tv_loss_bwd: the same loss and gradient bit for bit (same
    per-entry arithmetic and summation order); overwrite leaves slots outside the list untouched, accumulate adds."""
    print("case: tv fwd+bwd in one pass, channels_last", channels_last, flush=True)
    tree, active, ids, prefix, grid, d = scene(3, 60, 5, seed=8)
    g = d["grid"] if channels_last else d["grid"].contiguous()
    slots = torch.from_numpy(prefix[ids]).to(DEV)
    one = torch.full((1,), 0.75, device=DEV)
    l_sep = ops.tv_loss_fwd(g, slots, 1e-4)
    g_sep = torch.zeros_like(g)
    ops.tv_loss_bwd(g, slots, 1e-4, one, g_sep)
    g_ow = torch.full_like(g, 7.0)
    l_ow = ops.tv_loss_fwd_bwd(g, slots, 1e-4, one, g_ow, overwrite=True)
    g_acc = torch.full_like(g, 1.5)
    l_acc = ops.tv_loss_fwd_bwd(g, slots, 1e-4, one, g_acc, overwrite=False)
    assert l_ow.item() == l_sep.item() and l_acc.item() == l_sep.item()
    s = slots.cpu().numpy()
    other = np.setdiff1d(np.arange(g.shape[0]), s)
    assert torch.equal(g_ow[slots], g_sep[slots]) and bool((g_ow[torch.from_numpy(other).to(DEV)] == 7.0).all())
    assert torch.equal(g_acc, g_sep + 1.5)
