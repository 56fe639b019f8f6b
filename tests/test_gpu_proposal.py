"""The learned proposal on the GPU: network forward / backward and training targets against the numpy oracle
(oracle/proposal.py), the projector's proposal-guided samples against the C++ oracle, and the gain of the trained
network over the same sample budget without it.

Bars: proposal logits and gradients (BF16 on the matrix cores, FP32 accumulation) within 2e-2 / 3e-2 relative of the
float64 oracle; CDF rows, counts and targets exact to float rounding; FP32 projector with proposal samples within
1e-5 of the FP32 oracle (projection and gradients)."""
import numpy as np
import pytest
import torch

from neat_amd import ops, pipeline, proposal, workloads as W
from oracle import proposal as PR

pytestmark = pytest.mark.gpu
DEV = "cuda"


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def c2_segments(nrays, seed, jitter=False):
    cfg = W.C2
    sc = W.make_scene(cfg)
    ds = pipeline.to_device(sc, cfg)
    o, d = W.batch_rays(cfg, nrays, seed=seed)
    to, td = torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV)
    segs = pipeline.sample(ds, to, td, jitter=jitter, seed=seed, emit=False)
    return cfg, sc, ds, o, d, to, td, segs


def host_segs(segs):
    return {k: v.cpu().numpy() for k, v in segs.items() if torch.is_tensor(v)}


def test_proposal_forward_matches_oracle():
    print("case: proposal forward vs oracle", flush=True)
    cfg, sc, ds, o, d, to, td, segs = c2_segments(3000, 1)
    rng = np.random.default_rng(0)
    params = (PR.init_params(3) + rng.normal(size=PR.NPARAMS) * 0.05).astype(np.float32)
    cdf, cnt, lg = ops.proposal_fwd(to, td, segs, torch.from_numpy(params).to(DEV), 8, want_logits=True)
    h = host_segs(segs)
    x = PR.features(o, d, h["seg_ray"], h["seg_tmin"], h["seg_tmax"])
    ref, _ = PR.forward(params, x)
    lg = lg.cpu().numpy()
    e = rel(lg, ref)
    print("proposal logits (bf16) vs fp64 oracle: rel %.3e" % e, flush=True)
    assert e <= 2e-2
    cdf, cnt = cdf.cpu().numpy(), cnt.cpu().numpy()
    use = h["seg_count"] > 8
    assert np.array_equal(cnt, np.where(use, 8, h["seg_count"]))
    assert np.all(cdf[~use] == 0)
    assert np.allclose(cdf[use], PR.cdf_from_logits(lg[use].astype(np.float64)), rtol=0, atol=2e-6)
    assert use.any() and (~use).any()


def test_proposal_backward_matches_oracle():
    print("case: proposal backward vs oracle", flush=True)
    cfg, sc, ds, o, d, to, td, segs = c2_segments(2000, 2)
    rng = np.random.default_rng(1)
    params = (PR.init_params(4) + rng.normal(size=PR.NPARAMS) * 0.05).astype(np.float32)
    K = segs["seg_ray"].shape[0]
    q = rng.dirichlet(np.ones(8), K).astype(np.float32)
    g = ops.proposal_bwd(to, td, segs, torch.from_numpy(params).to(DEV), torch.from_numpy(q).to(DEV)).cpu().numpy()
    h = host_segs(segs)
    x = PR.features(o, d, h["seg_ray"], h["seg_tmin"], h["seg_tmax"])
    ref = PR.backward(params, x, q.astype(np.float64))
    e = rel(g, ref)
    print("proposal gradient (bf16) vs fp64 oracle: rel %.3e" % e, flush=True)
    assert e <= 3e-2


def test_sample_values_and_targets_match_oracle(oracle):
    print("case: field values at uniform samples and proposal targets vs oracle", flush=True)
    cfg, sc, ds, o, d, to, td, segs = c2_segments(600, 3)
    pj = ops.Projector(ds.pos_min, ds.pos_max, ds.prefix, to, td, segs, ds.grid, ds.params, ds.mlp, False, 0)
    values, off = pj.sample_values()
    values, off = values.cpu().numpy(), off.cpu().numpy()
    h = host_segs(segs)
    # the oracle's uniform samples of the same segment list, decoded in float32
    e = oracle.raysample(sc["tree"], sc["active_ids"], o, d, cfg.max_spn, jitter=False)
    uv = oracle.local_samples(sc["tree"], e["coord"], e["node_id"])
    feat = oracle.grid_sample_fwd(sc["grid"], sc["prefix"][e["node_id"]], uv)
    m = sc["mlp"]
    y = oracle.decoder_fwd(feat, sc["params"], m["in"], m["hidden"], m["layers"], m["out"], m["act"], m["last_act"],
                           m["beta"], f64=False)[:, 0]
    seg_start = np.concatenate([[0], np.cumsum(h["seg_count"])])[:-1]
    gpu = np.concatenate([values[off[k]:off[k] + h["seg_count"][k]] for k in range(len(off))])
    ref = np.concatenate([y[seg_start[k]:seg_start[k] + h["seg_count"][k]] for k in range(len(off))])
    ev = rel(gpu, ref)
    print("sample values vs fp32 oracle: rel %.3e" % ev, flush=True)
    assert ev <= 1e-5
    q = ops.proposal_targets(torch.from_numpy(values).to(DEV), torch.from_numpy(off).to(DEV), segs["seg_count"])
    qref = PR.targets(gpu, h["seg_count"])
    assert np.allclose(q.cpu().numpy(), qref, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("jitter", [False, True])
def test_projector_with_proposal_samples_matches_oracle(oracle, jitter):
    print("case: projector with proposal samples vs oracle, jitter", jitter, flush=True)
    cfg, sc, ds, o, d, to, td, segs = c2_segments(700, 4, jitter=jitter)
    ps = proposal.ProposalSampler(DEV, num_samples=8, seed=5)
    psegs = ps.apply(to, td, segs)
    gproj = np.random.default_rng(6).normal(size=o.shape[0]).astype(np.float32)
    pj = ops.Projector(ds.pos_min, ds.pos_max, ds.prefix, to, td, psegs, ds.grid, ds.params, ds.mlp, jitter, 4)
    proj = pj.forward().cpu().numpy()
    gg, gp = pj.backward(torch.from_numpy(gproj).to(DEV))
    h = host_segs(psegs)
    mlp = dict(sc["mlp"], params=sc["params"])
    ref = oracle.projector_fwd(sc["tree"], sc["prefix"], o, d, h, sc["grid"], mlp, jitter=jitter, seed=4)
    rgg, rgp = oracle.projector_bwd(sc["tree"], sc["prefix"], o, d, h, sc["grid"], mlp, gproj, jitter=jitter, seed=4)
    e = (rel(proj, ref), rel(gg.cpu().numpy(), rgg), rel(gp.cpu().numpy(), rgp))
    print("proposal-sampled projector vs fp32 oracle (proj, grid, params):", e, flush=True)
    assert max(e) <= 1e-5


def test_trained_proposal_beats_uniform_at_equal_budget():
    print("case: trained proposal vs uniform at 8 samples per segment, phantom field", flush=True)
    cfg = W.C2
    sc = W.make_scene(cfg)
    sc["grid"], sc["params"] = W.phantom_field(sc, cfg, device=DEV)
    ds = pipeline.to_device(sc, cfg)
    ps = proposal.ProposalSampler(DEV, num_samples=8, lr=3e-3, seed=7)
    losses = []
    for step in range(300):
        o, d = W.batch_rays(cfg, 4096, seed=1000 + step)
        to, td = torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV)
        segs = pipeline.sample(ds, to, td, jitter=False, emit=False)
        q = ps.train_step(ds, to, td, segs, precision="fp32")
        if step % 50 == 0 or step == 299:
            losses.append(ps.loss(to, td, segs, q))
    print("proposal cross entropy during training:", ["%.4f" % v for v in losses], flush=True)
    o, d = W.batch_rays(cfg, 8192, seed=99)
    to, td = torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV)
    segs = pipeline.sample(ds, to, td, jitter=False, emit=False)

    def project(s):
        pj = ops.Projector(ds.pos_min, ds.pos_max, ds.prefix, to, td, s, ds.grid, ds.params, ds.mlp, False, 0)
        return pj.forward().cpu().numpy().astype(np.float64)

    judge = project(segs)
    psegs = ps.apply(to, td, segs)
    usegs = dict(psegs, seg_cdf=torch.zeros_like(psegs["seg_cdf"]))  # same counts, uniform cells
    e_prop, e_unif = rel(project(psegs), judge), rel(project(usegs), judge)
    s_full, s_prop = int(segs["ray_sample_offset"][-1].item()), psegs["num_samples"]
    print("samples: uniform %d, capped %d (%.2fx fewer); error vs the uniform judge: proposal %.3e, uniform %.3e "
          "(gain %.2fx)" % (s_full, s_prop, s_full / s_prop, e_prop, e_unif, e_unif / e_prop), flush=True)
    assert losses[-1] < losses[0]
    assert e_prop < e_unif


def test_roughness_table_and_fallback_gate():
    print("case: per-slot roughness table and uniform fallback gate", flush=True)
    cfg, sc, ds, o, d, to, td, segs = c2_segments(500, 8)
    pj = ops.Projector(ds.pos_min, ds.pos_max, ds.prefix, to, td, segs, ds.grid, ds.params, ds.mlp, False, 0)
    values, off = pj.sample_values()
    N = ds.grid.shape[0]
    rough = torch.full((N,), 2.0, device=DEV)
    ops.proposal_roughness_update(values, off, segs, ds.prefix, N, rough, 0.5)
    h = host_segs(segs)
    v, of = values.cpu().numpy(), off.cpu().numpy()
    slot = sc["prefix"][h["seg_node"]]
    ref = np.full(N, 1.0)  # 2.0 decayed by 0.5
    for k in range(len(of)):
        tv = np.abs(np.diff(v[of[k]:of[k] + h["seg_count"][k]].astype(np.float64))).sum()
        ref[slot[k]] = max(ref[slot[k]], tv)
    assert np.allclose(rough.cpu().numpy(), ref, rtol=1e-5, atol=1e-6)
    tau = float(np.median(ref))
    ps = proposal.ProposalSampler(DEV, num_samples=8, seed=3, num_slots=N, tau=tau)
    ps.roughness.copy_(rough)
    p = ps.apply(to, td, segs, ds.prefix)
    cnt, cdf = p["seg_count"].cpu().numpy(), p["seg_cdf"].cpu().numpy()
    rough_seg = ref[slot] > tau
    assert np.array_equal(cnt[rough_seg], h["seg_count"][rough_seg]) and np.all(cdf[rough_seg] == 0)
    smooth = ~rough_seg & (h["seg_count"] > 8)
    assert smooth.any() and np.all(cnt[smooth] == 8) and np.all(cdf[smooth, 7] == 1)
    # three tiers: r <= tau_lo -> 8 proposal samples, tau_lo < r <= tau_u -> min(uniform count, 16) proposal samples
    # for segments over 8, r > tau_u -> uniform
    # (a table with slot % 3 as every slot's roughness puts a third of the slots in each tier)
    r = (slot % 3).astype(np.float64)
    n = h["seg_count"]
    tau_lo, tau_u = 0.5, 1.5
    ps3 = proposal.ProposalSampler(DEV, num_samples=8, seed=3, num_slots=N, tau=tau_lo, tau_uniform=tau_u,
                                   mid_samples=16)
    ps3.roughness.copy_(torch.arange(N, device=DEV).remainder(3).to(torch.float32))
    p3 = ps3.apply(to, td, segs, ds.prefix)
    cnt3, cdf3 = p3["seg_count"].cpu().numpy(), p3["seg_cdf"].cpu().numpy()
    lo, mid, hi = r <= tau_lo, (r > tau_lo) & (r <= tau_u), r > tau_u
    assert lo.any() and mid.any() and hi.any()
    assert np.array_equal(cnt3[lo], np.minimum(n[lo], 8))
    assert np.all(cdf3[lo & (n > 8), 7] == 1) and np.all(cdf3[lo & (n <= 8)] == 0)
    mp = mid & (n > 8)
    assert (mid & (n > 16)).any()
    assert np.array_equal(cnt3[mp], np.minimum(n[mp], 16)) and np.all(cdf3[mp, 7] == 1)
    assert np.array_equal(cnt3[mid & ~mp], n[mid & ~mp]) and np.all(cdf3[mid & ~mp] == 0)
    assert np.array_equal(cnt3[hi], n[hi]) and np.all(cdf3[hi] == 0)


@pytest.mark.parametrize("tau_uniform,mid", [(None, 8), (2.0, 32)])
def test_proposal_with_fallback_meets_the_error_bar(tau_uniform, mid):
    """Two tiers (8 proposal samples or uniform, tau 0.3) and three tiers (up to 32 proposal samples in slots
    rougher than 0.3 and at most 2.0 rough): each within 1e-3 of the FP64 uniform judge, with fewer samples; the
    share of segments that fall back to uniform sampling is printed."""
    print("case: trained proposal + roughness fallback vs the FP64 uniform judge, phantom field", tau_uniform, mid,
          flush=True)
    from oracle import oracle as O
    cfg = W.C2
    sc = W.make_scene(cfg)
    sc["grid"], sc["params"] = W.phantom_field(sc, cfg, device=DEV)
    ds = pipeline.to_device(sc, cfg)
    ps = proposal.ProposalSampler(DEV, num_samples=8, lr=3e-3, seed=7, num_slots=ds.grid.shape[0], tau=0.3,
                                  tau_uniform=tau_uniform, mid_samples=mid)
    for step in range(200):
        o, d = W.batch_rays(cfg, 4096, seed=2000 + step)
        to, td = torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV)
        ps.train_step(ds, to, td, pipeline.sample(ds, to, td, jitter=False, emit=False), precision="fp32")
    o, d = W.batch_rays(cfg, 2048, seed=98)
    to, td = torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV)
    segs = pipeline.sample(ds, to, td, jitter=False, emit=False)
    h = host_segs(segs)
    judge = O.projector_fwd(sc["tree"], sc["prefix"], o, d, h, sc["grid"], dict(sc["mlp"], params=sc["params"]),
                            prec=1)
    psegs = ps.apply(to, td, segs, ds.prefix)
    pj = ops.Projector(ds.pos_min, ds.pos_max, ds.prefix, to, td, psegs, ds.grid, ds.params, ds.mlp, False, 0)
    e = rel(pj.forward().cpu().numpy(), judge)
    s_full, s_prop = int(segs["ray_sample_offset"][-1].item()), psegs["num_samples"]
    fr = ps.tier_fractions(segs, ds.prefix)
    print("proposal + fallback (tau 0.3, tau_uniform %s, mid %d): error vs FP64 uniform judge %.3e, samples %d of %d "
          "(%.2fx fewer); tiers (8, mid, uniform) of segments over 8 samples: %.3f %.3f %.3f" % (
              tau_uniform, mid, e, s_prop, s_full, s_full / s_prop, *fr), flush=True)
    assert e <= 1e-3
    assert s_prop < s_full


def test_trained_field_three_tier_gate_at_bench_settings():
    """The bench's proposal form on a field this project trained (the phantom fitted on the leaf grids, then Adam
    steps of the fused BF16 projector against the phantom's projections, as pipeline.train_field does for the bench,
    here 100 steps on 2^16 rays): the three-tier gate at the bench's settings (tau 0.3, tau_uniform 2.0, at most 24
    proposal samples), the network trained online for 200 steps. Against the FP64 oracle with uniform samples on the
    same segments: the FP32 projection with proposal samples within 1e-3 (the sampling error alone; the BF16 form's
    is printed), fewer samples, and fewer than 20% of the long segments sent back to uniform sampling."""
    print("case: trained field, three-tier gate (0.3, 2.0, 24) vs the FP64 uniform judge", flush=True)
    from oracle import oracle as O
    cfg = W.C2
    sc = W.make_scene(cfg)
    sc["grid"], sc["params"] = W.phantom_field(sc, cfg, device=DEV)
    ds = pipeline.to_device(sc, cfg)
    losses, g16 = pipeline.train_field(ds, sc["phantom"], cfg, 100, 1 << 16, seed=31)
    print("field training loss: first %.3e, last %.3e" % (losses[0], losses[-1]), flush=True)
    assert losses[-1] < losses[0]
    sc["grid"], sc["params"] = ds.grid.cpu().contiguous().numpy(), ds.params.cpu().numpy()
    ps = proposal.ProposalSampler(DEV, num_samples=8, lr=3e-3, seed=7, num_slots=ds.grid.shape[0], tau=0.3,
                                  tau_uniform=2.0, mid_samples=24)
    for step in range(200):
        o, d = W.batch_rays(cfg, 4096, seed=3000 + step)
        to, td = torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV)
        ps.train_step(ds, to, td, pipeline.sample(ds, to, td, jitter=False, emit=False), precision="bf16",
                      grid_bf16=g16)
    o, d = W.batch_rays(cfg, 2048, seed=97)
    to, td = torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV)
    segs = pipeline.sample(ds, to, td, jitter=False, emit=False)
    judge = O.projector_fwd(sc["tree"], sc["prefix"], o, d, host_segs(segs), sc["grid"],
                            dict(sc["mlp"], params=sc["params"]), prec=1)
    psegs = ps.apply(to, td, segs, ds.prefix)

    def project(precision):
        pj = ops.Projector(ds.pos_min, ds.pos_max, ds.prefix, to, td, psegs, ds.grid, ds.params, ds.mlp, False, 0,
                           precision=precision, grid_bf16=g16 if precision == "bf16" else None)
        return pj.forward().cpu().numpy()

    e32, e16 = rel(project("fp32"), judge), rel(project("bf16"), judge)
    s_full, s_prop = int(segs["ray_sample_offset"][-1].item()), psegs["num_samples"]
    fr = ps.tier_fractions(segs, ds.prefix)
    print("error vs FP64 uniform judge: FP32 %.3e, BF16 %.3e; samples %d of %d (%.2fx fewer); tiers (8, 24, uniform) "
          "of segments over 8 samples: %.3f %.3f %.3f" % (e32, e16, s_prop, s_full, s_full / s_prop, *fr), flush=True)
    assert e32 <= 1e-3
    assert s_prop < s_full
    assert fr[2] < 0.2
