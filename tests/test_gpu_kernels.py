"""GPU parity of the individual kernels, called through the C ABI, against the CPU oracle.

Bar: intersection lists, node ids, sample counts, offsets and sample positions bit-exact; FP32 values within
1e-5 relative (norm-wise, and element-wise against a scale-aware bound)."""
import numpy as np
import pytest
import torch

from neat_amd import ops, workloads as W

pytestmark = pytest.mark.gpu
DEV = "cuda"
RTOL = 1e-5


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def assert_close(a, b, rtol=RTOL):
    """Norm-wise relative error <= rtol (the bar), and element-wise |a - b| <= rtol (|b| + max|b|): sums with
    cancellation (gradients accumulated from many signed terms) carry absolute, not relative, rounding error."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    assert rel_err(a, b) <= rtol, rel_err(a, b)
    scale = np.abs(b).max() if b.size else 1.0
    assert np.all(np.abs(a - b) <= rtol * (np.abs(b) + scale)), np.abs(a - b).max()


def special_rays():
    """Axis-aligned rays, rays on cell faces / edges / corners, zero components, origin inside, misses."""
    o = [[-3, -0.6, 0.3], [-3, -0.5, 0.0], [-3, 0.0, 0.0], [-3, 0.5, 0.5], [0.1, -4, 0.2], [0.0, 0.0, 0.0],
         [0.3, 0.2, -0.1], [-3, 2.0, 0.0], [-3, -0.25, 0.75], [3, 3, 3], [0.25, 0.25, -4], [-1, -1, -1]]
    d = [[1, 0, 0], [1, 0, 0], [1, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 1], [0.3, -0.5, 0.8], [1, 0, 0],
         [1, 0.0, 0], [-1, -1, -1], [0, 0, 1], [1, 1, 1]]
    o = np.array(o, np.float32)
    d = np.array(d, np.float64)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return o, d


def run_sampler(scene, o, d, max_spn, jitter=False, seed=0, single_walk=True):
    t = scene["tree"]
    pmin = torch.from_numpy(t["pos_min"]).to(DEV)
    pmax = torch.from_numpy(t["pos_max"]).to(DEV)
    flags = ops.tree_flags(scene["active"], t["max_depth"]).to(DEV)
    r = ops.raysample(pmin, pmax, flags, t["max_depth"], torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV),
                      max_spn, jitter=jitter, seed=seed, single_walk=single_walk)
    torch.cuda.synchronize()
    return {k: (v.cpu().numpy() if torch.is_tensor(v) else v) for k, v in r.items()}


def compare_sampler(gpu, ref):
    for k in ("ray_seg_offset", "ray_sample_offset", "seg_node", "seg_ray", "seg_count", "seg_tmin", "seg_tmax",
              "seg_sample_offset", "t", "weight", "ray_index", "local_index", "node_id"):
        g, r = np.asarray(gpu[k]), np.asarray(ref[k])
        assert g.shape == r.shape, (k, g.shape, r.shape)
        if g.dtype.kind == "f":
            assert np.array_equal(g.view(np.int32), r.view(np.int32)), (k, np.nonzero(g != r)[0][:5])
        else:
            assert np.array_equal(g, r), (k, np.nonzero(g != r)[0][:5])
    assert np.array_equal(gpu["coord"].view(np.int32), ref["coord"].view(np.int32))
    counts = np.diff(ref["ray_sample_offset"])
    assert gpu["max_samples_per_ray"] == (counts.max() if counts.size else 0)


@pytest.mark.parametrize("jitter", [False, True])
def test_sampler_bitexact_uniform_tree(oracle, jitter):
    print("case: sampler C1 geometry, jitter", jitter, flush=True)
    sc = W.make_scene(W.C1)
    o, d = W.batch_rays(W.C1, 1 << 14, seed=1)
    so, sd = special_rays()
    o = np.concatenate([so, o])
    d = np.concatenate([sd, d])
    gpu = run_sampler(sc, o, d, 32, jitter, seed=5)
    ref = oracle.raysample(sc["tree"], sc["active_ids"], o, d, 32, jitter=jitter, seed=5)
    compare_sampler(gpu, ref)


@pytest.mark.parametrize("single_walk", [True, False])
def test_sampler_bitexact_adaptive_tree(oracle, single_walk):
    print("case: sampler C2 geometry (4096 adaptive leaves), 4096 rays, single walk", single_walk, flush=True)
    sc = W.make_scene(W.C2)
    o, d = W.batch_rays(W.C2, 4096, seed=2)
    gpu = run_sampler(sc, o, d, 64, single_walk=single_walk)
    ref = oracle.raysample(sc["tree"], sc["active_ids"], o, d, 64)
    compare_sampler(gpu, ref)
    assert gpu["seg_node"].size > 10000


def test_sampler_bitexact_full_batch_scale(oracle):
    """A quarter of a full 2^18-ray batch of the full-size configuration (65,536 rays, about 0.49 M segments and
    13 M samples): every segment, count, t and weight bit for bit against the oracle's brute force over all active
    leaves, with jitter."""
    print("case: sampler C2 geometry, 65536 rays, jitter", flush=True)
    sc = W.make_scene(W.C2)
    o, d = W.batch_rays(W.C2, 1 << 16, seed=12)
    gpu = run_sampler(sc, o, d, 64, jitter=True, seed=9)
    ref = oracle.raysample(sc["tree"], sc["active_ids"], o, d, 64, jitter=True, seed=9)
    compare_sampler(gpu, ref)
    assert gpu["seg_node"].size > 400000


def test_sampler_single_walk_rewalks_rays_beyond_the_scratch(oracle):
    """Every node of a depth-3 tree active (nested): rays cross more segments than the single walk's per-ray scratch
    holds (3 * 2^3 + 8 = 32), and those rays are walked again by the compaction; the list stays bit-exact."""
    print("case: sampler, all nodes of a depth-3 tree active", flush=True)
    t = W.build_tree(3)
    act = np.ones(t["depth"].shape[0], np.int32)
    ids, prefix = W.update_active(act)
    sc = dict(tree=t, active=act, active_ids=ids)
    o, d = W.batch_rays(W.C1, 4096, seed=7)
    # near-diagonal rays cross up to 22 + 10 + 4 + 1 nested boxes
    rng = np.random.default_rng(3)
    do = (np.array([-2.0, -2.0, -2.0]) + rng.uniform(-0.05, 0.05, (512, 3))).astype(np.float32)
    dd = np.array([1.0, 1.0, 1.0]) + rng.uniform(-0.08, 0.08, (512, 3))
    dd = (dd / np.linalg.norm(dd, axis=1, keepdims=True)).astype(np.float32)
    o, d = np.concatenate([o, do]), np.concatenate([d, dd])
    gpu = run_sampler(sc, o, d, 8)
    assert np.diff(gpu["ray_seg_offset"]).max() > 32
    compare_sampler(gpu, oracle.raysample(t, ids, o, d, 8))


def test_sampler_nested_active_and_empty(oracle):
    print("case: sampler with nested active nodes, no rays, no active nodes", flush=True)
    t = W.build_tree(3)
    rng = np.random.default_rng(3)
    act = (rng.random(t["depth"].shape[0]) < 0.15).astype(np.int32)  # arbitrary, overlapping active set
    ids, prefix = W.update_active(act)
    sc = dict(tree=t, active=act, active_ids=ids)
    o, d = W.batch_rays(W.C1, 2048, seed=4)
    compare_sampler(run_sampler(sc, o, d, 16), oracle.raysample(t, ids, o, d, 16))
    e = np.zeros((0, 3), np.float32)
    g = run_sampler(sc, e, e, 16)
    assert g["ray_seg_offset"].tolist() == [0] and g["seg_node"].size == 0
    none = dict(tree=t, active=np.zeros_like(act), active_ids=np.zeros(0, np.int64))
    g = run_sampler(none, o[:64], d[:64], 16)
    assert g["ray_sample_offset"][-1] == 0


def _grid_t(grid, channels_last):
    g = torch.from_numpy(grid).to(DEV)
    return g.contiguous(memory_format=torch.channels_last_3d) if channels_last else g


@pytest.mark.parametrize("channels_last", [False, True])
def test_grid_sample(oracle, channels_last):
    print("case: grid sample, channels_last", channels_last, flush=True)
    rng = np.random.default_rng(5)
    grid = (rng.normal(size=(20, 8, 11, 11, 11)) * 0.1).astype(np.float32)
    S = 50000
    idx = rng.integers(0, 20, S)
    uv = rng.uniform(-1.05, 1.05, (S, 3)).astype(np.float32)
    uv[:100] = rng.choice([-1.0, 1.0, 0.0], (100, 3))
    idx[:3] = [-1, 20, 7]  # out-of-range slots sample zeros
    g = _grid_t(grid, channels_last)
    out = ops.gs3d_fwd(g, torch.from_numpy(idx).to(DEV), torch.from_numpy(uv).to(DEV)).cpu().numpy()
    ok = (idx >= 0) & (idx < 20)
    ref = oracle.grid_sample_fwd(grid, np.where(ok, idx, 0), uv)
    assert np.array_equal(out[ok], ref[ok])  # same operations in the same order: bit-exact
    assert np.all(out[~ok] == 0)
    go = rng.normal(size=out.shape).astype(np.float32)
    gg, guv = ops.gs3d_bwd(g, torch.from_numpy(idx).to(DEV), torch.from_numpy(uv).to(DEV),
                           torch.from_numpy(go).to(DEV))
    rgg, rguv = oracle.grid_sample_bwd(grid, idx[ok], uv[ok], go[ok])
    assert_close(gg.cpu().numpy(), rgg)
    assert_close(guv.cpu().numpy()[ok], rguv)


@pytest.mark.parametrize("hidden,act", [(32, "relu"), (64, "relu"), (64, "silu"), (16, "id")])
@pytest.mark.parametrize("layers", [0, 1])
def test_decoder(oracle, hidden, act, layers):
    print("case: decoder", hidden, act, layers, flush=True)
    rng = np.random.default_rng(6)
    S = 20000 + 37
    params = W.decoder_params(8, hidden, layers, 1, 7)
    x = (rng.normal(size=(S, 8)) * 0.3).astype(np.float32)
    tx = torch.from_numpy(x).to(DEV)
    tp = torch.from_numpy(params).to(DEV)
    y = ops.decoder_fwd(tx, tp, 8, hidden, layers, 1, act, "softplus", 2.0).cpu().numpy()
    ry = oracle.decoder_fwd(x, params, 8, hidden, layers, 1, act, "softplus", 2.0, f64=True)
    assert_close(y, ry)
    gy = rng.normal(size=(S, 1)).astype(np.float32)
    gp, gx = ops.decoder_bwd(tx, tp, torch.from_numpy(gy).to(DEV), 8, hidden, layers, 1, act, "softplus", 2.0)
    rgp, rgx = oracle.decoder_bwd(x, params, gy, 8, hidden, layers, 1, act, "softplus", 2.0, f64=True)
    assert_close(gp.cpu().numpy(), rgp)
    assert_close(gx.cpu().numpy(), rgx)


def test_integrate_and_mse(oracle):
    print("case: integrate + mse", flush=True)
    rng = np.random.default_rng(8)
    R = 3000
    counts = rng.integers(0, 40, R)
    off = np.concatenate([[0], np.cumsum(counts)])
    S = int(off[-1])
    ri = np.repeat(np.arange(R), counts)
    v = rng.normal(size=(S, 1)).astype(np.float32)
    w = rng.uniform(0, 0.1, S).astype(np.float32)
    ref = oracle.integrate_fwd(v, w, ri, R)
    tv, tw, tri = (torch.from_numpy(a).to(DEV) for a in (v, w, ri))
    seg = ops.integrate_fwd(tv, tw, tri, R, torch.from_numpy(off).to(DEV)).cpu().numpy()
    assert np.array_equal(seg, ref)  # in-order sum: identical to index_add_ on the CPU
    atom = ops.integrate_fwd(tv, tw, tri, R).cpu().numpy()
    assert_close(atom, ref)
    g = rng.normal(size=(1, R)).astype(np.float32)
    gv = ops.integrate_bwd(torch.from_numpy(g).to(DEV), tw, tri).cpu().numpy()
    assert np.array_equal(gv[:, 0], g[0, ri] * w)
    p = rng.normal(size=R).astype(np.float32)
    tgt = rng.normal(size=R).astype(np.float32)
    loss, grad = ops.mse_fp64(torch.from_numpy(p).to(DEV), torch.from_numpy(tgt).to(DEV))
    dd = p.astype(np.float64) - tgt
    assert abs(loss.item() - (dd * dd).mean()) <= 1e-12 * (dd * dd).mean()
    assert np.allclose(grad.cpu().numpy(), (2 * dd / R).astype(np.float32), rtol=1e-6)
