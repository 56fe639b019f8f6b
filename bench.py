#!/usr/bin/env python
"""Projector benchmark: fwd+bwd rays/s on one MI355X.

One step = one fwd+bwd pass of the projector over one batch of 2^18 synthetic cone-beam rays of the full-size
configuration (depth-5 octree adapted to a 100-ellipsoid phantom, 4096 leaves, 17^3 x 8 grids, decoder
8->64->64->1, 256 views x 512^2 detector, max 64 samples per node): ray sampling, gather, decoder, line integral,
FP64 MSE against the phantom's closed-form projections, and the backward to grid and decoder gradients.
Inputs (rays, targets, grid, decoder) are resident in HBM before the timed region.

Prints one JSON line (rank 0). Besides the throughput it carries the roofline of the dominant kernel (timed with
HIP events on the stream it runs on) and a CPU baseline: the oracle (the reference's CPU path restated in C++)
timed on a bounded sample of the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--path", default="fused", choices=["fused", "unfused"])
    ap.add_argument("--precision", default="bf16", choices=["bf16", "fp32", "fp32_valu"],
                    help="fused path arithmetic: BF16 on the matrix cores (default), FP32 on the matrix cores, or "
                         "FP32 on the vector ALUs")
    ap.add_argument("--batches", type=int, default=4, help="distinct resident ray batches cycled by the steps")
    ap.add_argument("--cpu-rays", type=int, default=12288, help="rays in the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--sampling", default="auto", choices=["auto", "proposal", "uniform"],
                    help="proposal (the default for the fused BF16 path): BF16 + learned proposal form on the field "
                         "of --field; segments with more than 8 uniform samples get 8 proposal-guided ones where their "
                         "slot's observed variation is at most --proposal-tau, up to --proposal-mid-samples where it "
                         "is at most --proposal-tau-uniform, and their uniform samples elsewhere; the "
                         "proposal network takes one online training step every --proposal-train-every steps on 1/32 "
                         "of the batch's rays; the projection error is reported against the FP64 oracle with uniform "
                         "samples. uniform (the default otherwise): the reference's uniform in-node sampling over a "
                         "random N(0, 0.1) grid.")
    ap.add_argument("--field", default="trained", choices=["trained", "phantom"],
                    help="field of the proposal form: trained (default): the phantom fitted on the leaf grids, then "
                         "--field-steps Adam steps of the fused BF16 training loop against the phantom's projections "
                         "(untimed set-up); phantom: the fitted phantom alone")
    ap.add_argument("--field-init", default="phantom", choices=["phantom", "random"],
                    help="start of the trained field: the fitted phantom, or the random N(0, 0.1) grid and Xavier "
                         "decoder of the configuration")
    ap.add_argument("--field-steps", type=int, default=300, help="training steps of the trained field (2^18 rays)")
    ap.add_argument("--proposal-tau", type=float, default=0.3,
                    help="slots whose observed field variation is at most this take 8 proposal samples per segment")
    ap.add_argument("--proposal-tau-uniform", type=float, default=2.0,
                    help="fallback threshold: segments in slots whose observed field variation exceeds it keep their "
                         "uniform samples (equal to --proposal-tau: two tiers)")
    ap.add_argument("--proposal-mid-samples", type=int, default=24,
                    help="proposal samples per segment in slots between the two thresholds")
    ap.add_argument("--proposal-train-every", type=int, default=4,
                    help="the proposal network's online training step runs every this many steps")
    ap.add_argument("--proposal-pretrain", type=int, default=200,
                    help="untimed proposal training steps (4096 rays each) before the timed region")
    ap.add_argument("--regularizer", action="store_true",
                    help="add the reference training step's volume regulariser (total variation of the active grids "
                         "at loss_tv 1e-4 and the edge loss along neighbouring leaves' faces at loss_edge 1e-3, with "
                         "its grid gradient) to every step, timed as its own stage")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the in-process comparison forms (FP32 on the matrix cores, the first FP32 versions) and "
                         "the timed training step")
    ap.add_argument("--compare-steps", type=int, default=3, help="timed steps of each comparison form")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (the reference kernel for scaling HBM "
                         "byte counters) and exit")
    return ap.parse_args()


def copy_calibration(mib):
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4}), flush=True)
    return 0


class Runner:
    """One projector form over the resident scene: (path, precision, proposal sampler or None). step(i, timed) runs
    one fwd+bwd pass over batch i % batches with per-stage HIP events on the launch stream when timed; with
    `train`, an Adam step on the grid (lr 1e-2) and the decoder (lr 1e-3) and the BF16 grid refresh follow
    (geometry_ex_ex.cpp:40-49, geometry.h:66-88)."""

    def __init__(self, a, cfg, ds, batches, path, precision, prop=None, grid_bf16=None, train=False,
                 active_ids=None, ray_grad=False):
        import torch
        self.a, self.cfg, self.ds, self.batches = a, cfg, ds, batches
        self.path, self.precision, self.prop, self.grid_bf16, self.train = path, precision, prop, grid_bf16, train
        self.active_ids = active_ids
        names = (["sample", "plan", "proj_fwd", "loss", "bwd_decoder", "bwd_params", "bwd_grid"]
                 if path == "fused" else ["sample", "gs_fwd", "dec_fwd", "integrate", "dec_bwd", "gs_bwd"])
        if prop is not None:
            names = ["sample", "proposal", "proposal_train"] + names[1:]
        if a.regularizer and path == "fused":
            names = names + ["regularizer"]
        if ray_grad:
            names = names + ["ray_grad"]
        if train:
            names = names + ["optimizer"]
        self.stage_names = names
        self.ray_grad = ray_grad
        self.events = {k: [] for k in names}
        self.gg_buf = torch.empty_like(ds.grid)
        self.gp_buf = torch.empty_like(ds.params)
        if train:
            self.m_grid, self.v_grid = torch.zeros_like(ds.grid), torch.zeros_like(ds.grid)
            self.m_par, self.v_par = torch.zeros_like(ds.params), torch.zeros_like(ds.params)
            self.opt_steps = 0

    def step(self, i, timed):
        import torch
        from neat_amd import ops, pipeline
        a, ds = self.a, self.ds
        o, d, tgt = self.batches[i % len(self.batches)]
        ev = {k: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for k in self.stage_names}
        rec = (lambda k, j: ev[k][j].record()) if timed else (lambda k, j: None)
        if self.path == "fused":
            rec("sample", 0)
            segs = pipeline.sample(ds, o, d, jitter=True, seed=i, emit=False)
            rec("sample", 1)
            if self.prop is not None:
                rec("proposal", 0)
                segs = self.prop.apply(o, d, segs, ds.prefix)
                rec("proposal", 1)
                rec("proposal_train", 0)
                if i % a.proposal_train_every == 0:
                    so, sd = o[::32].contiguous(), d[::32].contiguous()
                    self.prop.train_step(ds, so, sd, pipeline.sample(ds, so, sd, jitter=False, emit=False),
                                         precision="bf16", grid_bf16=self.grid_bf16)
                rec("proposal_train", 1)
            rec("plan", 0)
            pj = ops.Projector(ds.pos_min, ds.pos_max, ds.prefix, o, d, segs, ds.grid, ds.params, ds.mlp, True, i,
                               precision=self.precision, grid_bf16=self.grid_bf16)
            rec("plan", 1)
            rec("proj_fwd", 0)
            proj = pj.forward()
            rec("proj_fwd", 1)
            rec("loss", 0)
            loss, gproj = ops.mse_fp64(proj, tgt)
            rec("loss", 1)
            # the backward's three stages (decoder, parameter gradient, grid gradient), timed one by one
            for st, name in ((1, "bwd_decoder"), (2, "bwd_params"), (3, "bwd_grid")):
                rec(name, 0)
                pj.backward_stage(st, gproj, self.gg_buf, self.gp_buf)
                rec(name, 1)
            if self.ray_grad:
                rec("ray_grad", 0)
                pj.ray_grad()
                rec("ray_grad", 1)
            if a.regularizer:
                rec("regularizer", 0)
                ops.volume_regularizer(ds.pos_min, ds.pos_max, ds.flags, ds.prefix, self.active_ids, ds.grid, seed=i)
                rec("regularizer", 1)
            nsamp = pj.S
        else:
            R = o.shape[0]
            rec("sample", 0)
            smp = pipeline.sample(ds, o, d, jitter=True, seed=i)
            rec("sample", 1)
            uv = ops.local_samples(smp["coord"], smp["node_id"], ds.pos_min, ds.pos_max)
            slot = ops.node_to_slot(smp["node_id"], ds.prefix)
            rec("gs_fwd", 0)
            feat = ops.gs3d_fwd(ds.grid, slot, uv)
            rec("gs_fwd", 1)
            rec("dec_fwd", 0)
            y = ops.decoder_fwd(feat, ds.params, *pipeline._mlp_args(ds))
            rec("dec_fwd", 1)
            rec("integrate", 0)
            proj = ops.integrate_fwd(y, smp["weight"], smp["ray_index"], R, smp["ray_sample_offset"])
            loss, gproj = ops.mse_fp64(proj.reshape(-1), tgt)
            gy = ops.integrate_bwd(gproj.reshape(1, R), smp["weight"], smp["ray_index"])
            rec("integrate", 1)
            rec("dec_bwd", 0)
            gparams, gfeat = ops.decoder_bwd(feat, ds.params, gy, *pipeline._mlp_args(ds))
            rec("dec_bwd", 1)
            rec("gs_bwd", 0)
            ggrid, _ = ops.gs3d_bwd(ds.grid, slot, uv, gfeat, want_uv=False)
            rec("gs_bwd", 1)
            self.gg_buf, self.gp_buf = ggrid, gparams
            nsamp = int(smp["ray_sample_offset"][R].item())
        if self.train:
            rec("optimizer", 0)
            self.opt_steps += 1
            ops.adam_step(ds.grid, self.gg_buf, self.m_grid, self.v_grid, 1e-2, self.opt_steps)
            ops.adam_step(ds.params, self.gp_buf, self.m_par, self.v_par, 1e-3, self.opt_steps)
            if self.grid_bf16 is not None:
                ops.grid_to_bf16(ds.grid, out=self.grid_bf16)
            rec("optimizer", 1)
        if timed:
            for k in self.stage_names:
                self.events[k].append(ev[k])
        return nsamp

    def run(self, steps, warmup):
        """warmup untimed steps, then `steps` timed between two device synchronisations: (ms per step, mean samples
        per batch, mean ms per stage)."""
        import torch
        for i in range(warmup):
            self.step(i, False)
        torch.cuda.synchronize()
        samples = []
        t0 = time.perf_counter()
        for i in range(steps):
            samples.append(self.step(warmup + i, True))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        stage_ms = {k: float(np.mean([s.elapsed_time(e) for s, e in self.events[k]])) for k in self.stage_names}
        return (t1 - t0) * 1e3 / steps, float(np.mean(samples)), stage_ms


def main():
    a = parse()
    if a.sampling == "auto":
        a.sampling = "proposal" if (a.path == "fused" and a.precision == "bf16") else "uniform"
    a.proposal = a.sampling == "proposal"
    if a.copy_calibration:
        return copy_calibration(a.copy_calibration)
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it", "n_gpus": a.gpus}))
        return 2
    import torch

    from neat_amd import _abi, host, ops, pipeline, workloads as W

    cfg = W.C2
    dev = "cuda"
    scene = W.make_scene(cfg, seed=a.seed)
    field_info = "random grid N(0, 0.1)"
    if a.proposal:
        if a.path != "fused" or a.precision != "bf16":
            raise SystemExit("--sampling proposal needs the fused BF16 path")
        if a.field == "phantom" or a.field_init == "phantom":
            scene["grid"], scene["params"] = W.phantom_field(scene, cfg, device=dev)
            field_info = "phantom fitted on the leaf grids"
    ds = pipeline.to_device(scene, cfg, dev)
    if a.proposal and a.field == "trained":
        t0 = time.time()
        losses, _ = pipeline.train_field(ds, scene["phantom"], cfg, a.field_steps, cfg.batch, seed=a.seed + 31)
        # the oracle (projection error) reads the trained field
        scene["grid"] = ds.grid.cpu().contiguous().numpy()
        scene["params"] = ds.params.cpu().numpy()
        field_info = {"field": f"trained: {field_info if a.field_init == 'phantom' else 'random init'}, then "
                               f"{a.field_steps} Adam steps (lr 1e-2 grid, 1e-3 decoder) of the fused BF16 path on "
                               "2^18 rays against the phantom's projections",
                      "loss_first": float(losses[0]), "loss_last10_mean": float(losses[-10:].mean()),
                      "setup_s": round(time.time() - t0, 1)}
    batches = []
    for b in range(a.batches):
        o, d = W.batch_rays(cfg, cfg.batch, seed=a.seed + 100 + b)
        to, td = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
        batches.append((to, td, scene["phantom"].line_integrals_torch(to, td).to(torch.float32)))
    torch.cuda.synchronize()

    # the BF16 grid copy is refreshed once per optimizer step in training; here the grid is fixed, so once
    grid_bf16 = ops.grid_to_bf16(ds.grid) if (a.path == "fused" and a.precision == "bf16") else None
    prop = None
    if a.proposal:
        from neat_amd import proposal as PP
        prop = PP.ProposalSampler(dev, num_samples=8, lr=3e-3, seed=a.seed + 7, num_slots=ds.grid.shape[0],
                                  tau=a.proposal_tau, tau_uniform=a.proposal_tau_uniform,
                                  mid_samples=a.proposal_mid_samples)
        for k in range(a.proposal_pretrain):  # untimed set-up: the network starts trained
            o, d = W.batch_rays(cfg, 4096, seed=a.seed + 5000 + k)
            to, td = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
            prop.train_step(ds, to, td, pipeline.sample(ds, to, td, jitter=False, emit=False), precision="bf16",
                            grid_bf16=grid_bf16)
        torch.cuda.synchronize()
    active_ids = torch.from_numpy(scene["active_ids"]).to(dev) if a.regularizer else None
    head = Runner(a, cfg, ds, batches, a.path, a.precision, prop, grid_bf16, active_ids=active_ids)
    ms_per_step, mean_samples, stage_ms = head.run(a.steps, a.warmup)
    rays_per_s = cfg.batch / (ms_per_step / 1e3)

    dominant = max(stage_ms, key=stage_ms.get)
    C, H = cfg.channels, cfg.hidden
    fwd_flop = 2 * (C * H + H * H + H * 1)  # decoder forward per sample, 2 FLOP per MAC
    mm_peak = BF16_PEAK_TFLOPS if (a.path == "fused" and a.precision == "bf16") else FP32_PEAK_TFLOPS
    mm = "fp32_valu" != a.precision
    prec_tag = {"bf16": "PrecBF16", "fp32": "PrecF32"}.get(a.precision)
    act = 1 if cfg.act == "relu" else 2
    grid_bytes = ds.grid.numel() * 4
    if not (mm and prec_tag):
        bwd_kernel = f"k_proj_bw1<{C},{H}>"
    elif prec_tag == "PrecBF16" and cfg.act == "relu":
        bwd_kernel = f"k_mfma_bwd_relu<{H}>"
    else:
        bwd_kernel = f"k_mfma_bwd<{prec_tag},{H},{act}>"
    fwd_kernel = (f"k_mfma_fwd_relu<{H}>" if prec_tag == "PrecBF16" and cfg.act == "relu" else
                  f"k_mfma_fwd<{prec_tag},{H},{act}>" if mm and prec_tag else f"k_proj_fwd<{C},{H},{act}>")
    # algorithmic work per launch of each stage's dominant kernel (per-unit figure x the units of one launch):
    # decoder FLOPs per sample (backward: 2x the forward, no recomputation counted); the scatter reads 16 B of
    # dL/dfeature per sample and writes the gradient grid once
    work = {
        "proj_fwd": ("mfma", fwd_flop * mean_samples, "TFLOP/s", mm_peak, fwd_kernel),
        "bwd_decoder": ("mfma", 2 * fwd_flop * mean_samples, "TFLOP/s", mm_peak, bwd_kernel),
        "bwd_grid": ("hbm", 48 * mean_samples + grid_bytes, "GB/s", HBM_PEAK_GBS, f"k_proj_bw2<{C}>"),
        "bwd_params": ("hbm", 4.0 * 1024 * ds.params.numel(), "GB/s", HBM_PEAK_GBS, "k_dec_reduce_partials"),
        "plan": ("hbm", 16 * mean_samples / 27, "GB/s", HBM_PEAK_GBS, "plan"),
        "proposal": ("mfma", 2 * (6 * 32 + 32 * 32 + 32 * 8) * mean_samples / 8, "TFLOP/s", BF16_PEAK_TFLOPS,
                     "k_prop_fwd"),
        "proposal_train": ("mfma", 3 * 2 * (6 * 32 + 32 * 32 + 32 * 8) * mean_samples / 128, "TFLOP/s",
                           BF16_PEAK_TFLOPS, "k_prop_bwd"),
        "loss": ("hbm", 12 * cfg.batch, "GB/s", HBM_PEAK_GBS, "k_mse"),
        "regularizer": ("hbm", 3 * grid_bytes, "GB/s", HBM_PEAK_GBS, "k_tv_bwd"),
        "dec_fwd": ("mfma", fwd_flop * mean_samples, "TFLOP/s", FP32_PEAK_TFLOPS, "k_dec_fwd"),
        "dec_bwd": ("mfma", 2 * fwd_flop * mean_samples, "TFLOP/s", FP32_PEAK_TFLOPS, "k_dec_bwd"),
        "gs_fwd": ("hbm", (12 + 8 + 4 * C) * mean_samples, "GB/s", HBM_PEAK_GBS, "k_gs3d_fwd"),
        "gs_bwd": ("hbm", (12 + 8 + 4 * C) * mean_samples + 2 * grid_bytes, "GB/s", HBM_PEAK_GBS, "k_gs3d_bwd"),
        "sample": ("hbm", 44 * mean_samples / 27 + 32 * cfg.batch, "GB/s", HBM_PEAK_GBS, "k_ray_segments"),
        "integrate": ("hbm", (4 + 4 + 8) * mean_samples * 2 + 16 * cfg.batch, "GB/s", HBM_PEAK_GBS, "k_integrate"),
    }
    bound, amount, unit, peak, kname = work[dominant]
    secs = stage_ms[dominant] / 1e3
    achieved = amount / secs / (1e12 if unit == "TFLOP/s" else 1e9)
    roofline = {"bound": bound, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
                "frac": round(achieved / peak, 5),
                "traffic": measured_traffic(workload_name(a), kname), "kernel": kname,
                "stage": dominant, "kernel_ms": round(stage_ms[dominant], 4)}

    err = projection_error(scene, cfg, ds, prop, a, grid_bf16)
    tiers = None
    if prop is not None:
        o0, d0 = batches[0][0], batches[0][1]
        fr = prop.tier_fractions(pipeline.sample(ds, o0, d0, jitter=False, emit=False), ds.prefix)
        tiers = {"proposal_8": round(fr[0], 4), f"proposal_{a.proposal_mid_samples}": round(fr[1], 4),
                 "uniform": round(fr[2], 4), "of": "segments with more than 8 uniform samples (one batch)"}

    # comparison forms on the same field and rays, after the timed region: the FP32 forms sample uniformly (the
    # proposal is part of the BF16 form); the first plain-FP32 GPU version (the unfused kernel chain) is the
    # denominator the north-star target is stated against
    baselines, vs_baseline, train = None, None, None
    if not a.no_compare and a.path == "fused" and a.precision == "bf16" and a.proposal and not a.regularizer:
        baselines = {}
        for name, path, prec in (("fp32_mfma", "fused", "fp32"), ("fp32_v1_fused_valu", "fused", "fp32_valu"),
                                 ("fp32_unfused_chain", "unfused", "fp32")):
            r = Runner(a, cfg, ds, batches, path, prec)
            ms, ns, _ = r.run(a.compare_steps, 1)
            baselines[name] = {"rays_per_s": round(cfg.batch / (ms / 1e3), 1), "ms_per_step": round(ms, 3),
                               "samples_per_batch": round(ns), "ratio": round(rays_per_s / (cfg.batch / (ms / 1e3)), 2)}
            del r
            torch.cuda.empty_cache()
        # pose refinement: the headline form with the positional gradient onto the rays (fused, as the rays of
        # GetRays with pose gradients take it)
        r = Runner(a, cfg, ds, batches, "fused", "bf16", prop, grid_bf16, ray_grad=True)
        ms, ns, sm = r.run(a.compare_steps, 1)
        baselines["bf16_proposal_with_ray_grad"] = {"rays_per_s": round(cfg.batch / (ms / 1e3), 1),
                                                    "ms_per_step": round(ms, 3), "ray_grad_ms": round(sm["ray_grad"], 3)}
        del r
        vs_baseline = baselines["fp32_v1_fused_valu"]["ratio"]
        baselines["denominator"] = ("fp32_v1_fused_valu (the first plain-FP32 GPU projector: fused, vector ALUs, "
                                    "kept unchanged)")
        # the training step: the headline form followed by Adam on the grid and the decoder and the BF16 refresh,
        # on copies of the field (the headline's tensors stay as they were)
        tds = pipeline.DeviceScene(ds.max_depth, ds.pos_min, ds.pos_max, ds.flags, ds.prefix, ds.grid.clone(),
                                   ds.params.clone(), ds.mlp, ds.max_spn)
        tr = Runner(a, cfg, tds, batches, "fused", "bf16", prop, ops.grid_to_bf16(tds.grid), train=True)
        ms, ns, sm = tr.run(a.steps, a.warmup)
        train = {"rays_per_s": round(cfg.batch / (ms / 1e3), 1), "ms_per_step": round(ms, 3),
                 "optimizer_ms": round(sm["optimizer"], 3),
                 "what": "headline fwd+bwd, then Adam on the grid (lr 1e-2) and decoder (lr 1e-3) and the BF16 grid "
                         "refresh"}
        del tr, tds
        torch.cuda.empty_cache()

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(scene, cfg, a.cpu_rays, a.seed)

    prec_name = workload_name(a)
    config = {"workload": f"projector fwd+bwd ({prec_name}): depth-5 octree, 4096 adaptive leaves, 17^3x8 grids, "
                          "decoder 8->64->64->1, 256x512^2 cone beam, batch 2^18 rays, max 64 samples/node",
              "path": a.path, "precision": prec_name, "batch_rays": cfg.batch,
              "samples_per_batch": round(mean_samples),
              "samples_per_s": round(mean_samples / (ms_per_step / 1e3), 1),
              "field": field_info,
              "proposal_tiers": tiers,
              "projection_error_vs_fp64": err,
              "stage_ms": {k: round(v, 3) for k, v in stage_ms.items()},
              "build_id": {"libneat_amd": _abi.build_id(), "neat_host": host.load().build_id()}}
    if baselines is not None:
        config["baselines"] = baselines
        config["train_step"] = train
    out = {
        "metric": "fwd+bwd rays/s at 1 GPU, BF16+proposal vs plain FP32, at <=1e-3 rel err vs FP64",
        "value": round(rays_per_s, 1),
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": vs_baseline,
        "dtype": prec_name.split("_")[0],
        "data": "synthetic",
        "config": config,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def projection_error(scene, cfg, ds, prop, a, grid_bf16, nrays=2048):
    """Relative projection error of the benchmarked form against the FP64 oracle with the reference's uniform samples,
    on a fixed sub-batch (no jitter), with the error bar of the metric."""
    if a.path != "fused":
        return None
    import torch
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from neat_amd import ops, pipeline, workloads as W
    from oracle import oracle as O
    o, d = W.batch_rays(cfg, nrays, seed=a.seed + 777)
    to, td = torch.from_numpy(o).cuda(), torch.from_numpy(d).cuda()
    segs = pipeline.sample(ds, to, td, jitter=False, emit=False)
    hseg = {k: v.cpu().numpy() for k, v in segs.items() if torch.is_tensor(v)}
    ref = O.projector_fwd(scene["tree"], scene["prefix"], o, d, hseg, scene["grid"], dict(scene["mlp"],
                          params=scene["params"]), prec=1)
    if prop is not None:
        segs = prop.apply(to, td, segs, ds.prefix)
    pj = ops.Projector(ds.pos_min, ds.pos_max, ds.prefix, to, td, segs, ds.grid, ds.params, ds.mlp, False, 0,
                       precision=a.precision, grid_bf16=grid_bf16)
    p = pj.forward().cpu().numpy().astype(np.float64)
    e = float(np.linalg.norm(p - ref) / np.linalg.norm(ref))
    return {"rel": e, "bar": 1e-3, "meets_bar": e <= 1e-3, "rays": nrays,
            "judge": "FP64 oracle, uniform samples (max 64 per node), same segments"}


def workload_name(a):
    """Name of the measured form: bf16+proposal, bf16, fp32, fp32_valu or fp32_unfused."""
    if a.path != "fused":
        return "fp32_unfused"
    return a.precision + ("+proposal" if a.proposal else "")


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` in `workload` from the committed rocprofv3 counter summary (FETCH_SIZE,
    WRITE_SIZE in KiB per dispatch, each scaled by the reading of a plain copy of known size), or None when that
    workload was not profiled."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("workloads", {}).get(workload, {}).get("kernels", {}).get(kernel)
    if not k:
        return None
    return round(1024 * (k["fetch_kib"] * t["fetch_scale"] + k["write_kib"] * t["write_scale"]))


def host_cores():
    """Host threads this process may use: its CPU affinity, capped by OMP_NUM_THREADS when that is set (the GPU box
    gives each GPU a share of the host)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def _cpu_pass(scene, cfg, o, d, seed, nthreads):
    """The oracle's FP32 restatement of the reference path on rays (o, d): sampling by brute force over the active
    leaves, then the composed forward and backward; returns the elapsed seconds."""
    from oracle import oracle as O
    tgt = scene["phantom"].line_integrals(o, d)
    t0 = time.perf_counter()
    segs = O.raysample(scene["tree"], scene["active_ids"], o, d, cfg.max_spn, emit=False, nthreads=nthreads)
    proj = O.projector_fwd(scene["tree"], scene["prefix"], o, d, segs, scene["grid"], dict(scene["mlp"],
                           params=scene["params"]), jitter=True, seed=seed, prec=0, nthreads=nthreads)
    gproj = 2 * (proj - tgt) / o.shape[0]
    O.projector_bwd(scene["tree"], scene["prefix"], o, d, segs, scene["grid"], dict(scene["mlp"],
                    params=scene["params"]), gproj, jitter=True, seed=seed, prec=0, nthreads=nthreads)
    return time.perf_counter() - t0


def cpu_baseline(scene, cfg, num_rays, seed):
    """The reference's CPU path, restated (the oracle), in the two variants the baseline plan names: (i) one thread,
    like the reference's CPU loops, on a bounded sample of the benchmarked configuration; (ii) ray-parallel forward
    and thread-privatised backward over every host core this process may use, on a larger sample of the same
    configuration and on one full 2^18-ray batch of the small configuration (depth 2, 64 leaves)."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from neat_amd import workloads as W
    cores = host_cores()
    o, d = W.batch_rays(cfg, num_rays, seed=seed + 999)
    dt1 = _cpu_pass(scene, cfg, o, d, seed, 1)
    n_all = num_rays * max(1, min(cores, 16)) // 2
    o, d = W.batch_rays(cfg, n_all, seed=seed + 998)
    dtn = _cpu_pass(scene, cfg, o, d, seed, cores)
    c1 = W.C1
    sc1 = W.make_scene(c1, seed=seed)
    o, d = W.batch_rays(c1, c1.batch, seed=seed + 997)
    dtc1 = _cpu_pass(sc1, c1, o, d, seed, cores)
    return {"value": round(n_all / dtn, 2), "unit": "rays/s", "cores": cores, "kind": "port",
            "sample": f"{n_all} rays of the same workload (sampling + fwd + bwd), oracle over {cores} threads "
                      f"(ray-parallel forward, thread-privatised backward), {dtn:.1f} s",
            "single_thread": {"value": round(num_rays / dt1, 2), "unit": "rays/s", "cores": 1,
                              "sample": f"{num_rays} rays of the same workload, one thread, {dt1:.1f} s"},
            "c1_full_batch": {"value": round(c1.batch / dtc1, 2), "unit": "rays/s", "cores": cores,
                              "seconds": round(dtc1, 2),
                              "sample": f"one full batch of {c1.batch} rays of the small configuration (depth-2 "
                                        f"octree, 64 leaves, 11^3x8 grids, decoder 8->32->32->1), {cores} threads"}}


if __name__ == "__main__":
    sys.exit(main())
