"""Training form (Adam reconstruction step over the C1 geometry, rays from 256 views): the Adam kernel against
torch.optim.Adam, the fused backward's grid gradient against central finite differences of the FP64 oracle's loss
(eps 1e-3, 64 random touched grid entries), and a reconstruction run through the libtorch module layer
(GeometryExEx.ComputeImage -> loss -> backward -> OptimizerStep) whose loss must fall."""
import numpy as np
import pytest
import torch

from neat_amd import host, ops, pipeline, workloads as W

pytestmark = pytest.mark.gpu
DEV = "cuda"


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("n", [4099, 1 << 20])
def test_adam_step_matches_torch_adam(n):
    print("case: adam vs torch.optim.Adam, n =", n, flush=True)
    g = torch.Generator(device=DEV).manual_seed(n)
    p = torch.randn(n, device=DEV, generator=g)
    ref = p.clone().requires_grad_(True)
    opt = torch.optim.Adam([ref], lr=1e-2, foreach=False)
    m, v = torch.zeros_like(p), torch.zeros_like(p)
    for step in range(1, 4):
        grad = torch.randn(n, device=DEV, generator=g) * (10.0 ** -step)
        ref.grad = grad.clone()
        opt.step()
        ops.adam_step(p, grad, m, v, lr=1e-2, step=step)
    torch.cuda.synchronize()
    e = rel(p.cpu().numpy(), ref.detach().cpu().numpy())
    print("adam: param rel err %.3e" % e, flush=True)
    assert e <= 1e-6


def test_grid_gradient_central_differences(oracle):
    print("case: fused grid gradient vs central FD of the FP64 oracle loss", flush=True)
    cfg = W.C4
    sc = W.make_scene(cfg)
    ds = pipeline.to_device(sc, cfg)
    o, d = W.batch_rays(cfg, 2048, seed=5)
    to, td = torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV)
    tgt = sc["phantom"].line_integrals(o, d)
    segs = pipeline.sample(ds, to, td, jitter=True, seed=12, emit=False)
    pj = ops.Projector(ds.pos_min, ds.pos_max, ds.prefix, to, td, segs, ds.grid, ds.params, ds.mlp, True, 12)
    proj = pj.forward()
    loss, gproj = ops.mse_fp64(proj, torch.from_numpy(tgt.astype(np.float32)).to(DEV))
    gg, _ = pj.backward(gproj)
    gg = gg.cpu().numpy()
    hseg = {k: v.cpu().numpy() for k, v in segs.items() if torch.is_tensor(v)}
    mlp = dict(sc["mlp"], params=sc["params"])

    def L(grid):
        p = oracle.projector_fwd(sc["tree"], sc["prefix"], o, d, hseg, grid, mlp, jitter=True, seed=12, prec=1)
        return float(np.mean((p - tgt) ** 2))

    rng = np.random.default_rng(7)
    touched = np.flatnonzero(np.abs(gg) > 1e-3 * np.abs(gg).max())
    picks = rng.choice(touched, 64, replace=False)
    eps = 1e-3
    fd, an = [], []
    for f in picks:
        idx = np.unravel_index(f, gg.shape)
        gp_, gm_ = sc["grid"].copy(), sc["grid"].copy()
        gp_[idx] += eps
        gm_[idx] -= eps
        h = float(gp_[idx]) - float(gm_[idx])  # the perturbation float32 represents
        fd.append((L(gp_) - L(gm_)) / h)
        an.append(gg[idx])
    e = rel(an, fd)
    print("grid gradient vs central differences on 64 entries: rel err %.3e" % e, flush=True)
    assert e <= 2e-3


def make_geo(H, cfg, precision):
    sc = W.make_scene(cfg)
    tree = H.HyperTreeBase(3, cfg.depth)
    tree.node_active.copy_(torch.from_numpy(sc["active"]))
    tree.UpdateActive()
    tree.to(DEV)
    p = H.CombinedParams()
    p.net_params.grid_size = cfg.grid
    p.net_params.grid_features = cfg.channels
    p.net_params.decoder_hidden_features = cfg.hidden
    p.net_params.decoder_hidden_layers = cfg.layers
    p.net_params.decoder_activation = cfg.act
    p.net_params.softplus_beta = cfg.beta
    p.net_params.decoder_lr = cfg.extra["lr_decoder"]
    p.octree_params.max_samples_per_node = cfg.max_spn
    p.octree_params.tree_optimizer_params.max_active_nodes = sc["grid"].shape[0]
    p.train_params.grid_init = "zero"
    p.train_params.lr_exex_grid_adam = cfg.extra["lr_grid"]
    p.train_params.projector_precision = precision
    geo = H.GeometryExEx(1, 3, tree, p)
    geo.to(DEV)
    with torch.no_grad():
        geo.grid.copy_(torch.from_numpy(sc["grid"]).to(DEV))
    geo.decoder.load_packed(torch.from_numpy(sc["params"]).to(DEV))
    geo.CreateGeometryOptimizer()
    return sc, tree, geo


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_reconstruction_steps_reduce_loss(precision):
    print("case: Adam reconstruction through the module layer,", precision, flush=True)
    H = host.load()
    cfg = W.C4
    sc, tree, geo = make_geo(H, cfg, precision)
    losses = []
    for step in range(20):
        o, d = W.batch_rays(cfg, 1 << 16, seed=100 + step)
        tgt = torch.from_numpy(sc["phantom"].line_integrals(o, d).astype(np.float32)).to(DEV)
        rays = H.RayList(torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV))
        smp = tree.CreateSamplesForRays(rays, cfg.max_spn, True)
        img = geo.ComputeImage(smp, 1, o.shape[0])
        loss = ((img - tgt.view(1, -1)) ** 2).mean()
        if step == 0:
            g0 = geo.grid.detach().clone()
        loss.backward()
        if step == 0:
            # the first Adam step moves every touched entry by lr * g / (|g| + eps) (bias-corrected moments = g, g^2)
            gr = geo.grid.grad.detach().clone()
        geo.OptimizerStep(step)
        if step == 0:
            expect = g0 - cfg.extra["lr_grid"] * gr / (gr.abs() + 1e-8)
            e = rel(geo.grid.detach().cpu().numpy(), expect.cpu().numpy())
            print("first Adam step vs closed form: rel %.3e" % e, flush=True)
            assert e <= 1e-6
            assert float(geo.grid.grad.abs().max()) == 0.0  # OptimizerStep zeroes the gradients
        losses.append(loss.item())
    print(precision, "losses", ["%.4g" % v for v in losses], flush=True)
    assert losses[-1] < 0.8 * losses[0]


def _train(cfg, sc, precision, steps, batch):
    """Adam (lr 1e-2 grid, 1e-3 decoder) on the fused projector, same init, rays and jitter seeds for any precision;
    the BF16 form refreshes its BF16 grid copy after every step. Returns the per-step FP64 losses."""
    ds = pipeline.to_device(sc, cfg)
    mg, vg = torch.zeros_like(ds.grid), torch.zeros_like(ds.grid)
    mp, vp = torch.zeros_like(ds.params), torch.zeros_like(ds.params)
    g16 = ops.grid_to_bf16(ds.grid) if precision == "bf16" else None
    losses = []
    for step in range(1, steps + 1):
        o, d = W.batch_rays(cfg, batch, seed=1000 + step)
        to, td = torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV)
        tgt = torch.from_numpy(sc["phantom"].line_integrals(o, d).astype(np.float32)).to(DEV)
        segs = pipeline.sample(ds, to, td, jitter=True, seed=step, emit=False)
        pj = ops.Projector(ds.pos_min, ds.pos_max, ds.prefix, to, td, segs, ds.grid, ds.params, ds.mlp, True, step,
                           precision=precision, grid_bf16=g16)
        loss, gproj = ops.mse_fp64(pj.forward(), tgt)
        gg, gp = pj.backward(gproj)
        ops.adam_step(ds.grid, gg, mg, vg, cfg.extra["lr_grid"], step)
        ops.adam_step(ds.params, gp, mp, vp, cfg.extra["lr_decoder"], step)
        if g16 is not None:
            ops.grid_to_bf16(ds.grid, out=g16)
        losses.append(loss.item())
    return np.array(losses)


def test_bf16_training_tracks_fp32():
    """200 Adam steps of the training form (C1 geometry, rays from 256 views, 2^16 rays per step) in BF16 and in
    FP32 from the same start on the same rays: the BF16 loss curve stays within 5% of the FP32 one over the last 20
    steps, and both fall well below their start."""
    print("case: 200 Adam steps, bf16 vs fp32", flush=True)
    cfg = W.C4
    sc = W.make_scene(cfg)
    l32 = _train(cfg, sc, "fp32", 200, 1 << 16)
    l16 = _train(cfg, sc, "bf16", 200, 1 << 16)
    a32, a16 = l32[-20:].mean(), l16[-20:].mean()
    print("loss start %.4g; last-20 mean fp32 %.4g, bf16 %.4g (rel %.3e)" % (l32[0], a32, a16, abs(a16 - a32) / a32),
          flush=True)
    print("fp32 every 20th:", ["%.4g" % v for v in l32[::20]], flush=True)
    print("bf16 every 20th:", ["%.4g" % v for v in l16[::20]], flush=True)
    assert a32 < 0.5 * l32[0] and a16 < 0.5 * l16[0]
    assert abs(a16 - a32) <= 0.05 * a32
