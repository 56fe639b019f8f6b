"""Device-resident projector scenes and the unfused FP32 projector step, composed from the C-ABI kernels.

The unfused step is the reference's own decomposition of ComputeImage (geometry.cpp:347-424) run on HIP kernels:
CreateSamplesForRays -> ComputeLocalSamples -> GlobalNodeIdToLocalActiveId -> IndirectGridSample3D -> decoder +
softplus -> IntegrateSamplesXRay, then the MSE and the hand-written backward of every stage.
"""
from dataclasses import dataclass

import numpy as np
import torch

from . import ops


@dataclass
class DeviceScene:
    max_depth: int
    pos_min: torch.Tensor
    pos_max: torch.Tensor
    flags: torch.Tensor
    prefix: torch.Tensor
    grid: torch.Tensor  # [slots, C, G, G, G] float32, channels-last-3d memory format
    params: torch.Tensor  # packed decoder parameters
    mlp: dict
    max_spn: int

    @property
    def num_nodes(self):
        return self.pos_min.shape[0]


def to_device(scene, cfg, device="cuda", channels_last=True):
    t = scene["tree"]
    g = torch.from_numpy(scene["grid"]).to(device)
    if channels_last:
        g = g.contiguous(memory_format=torch.channels_last_3d)
    return DeviceScene(
        max_depth=t["max_depth"],
        pos_min=torch.from_numpy(t["pos_min"]).to(device),
        pos_max=torch.from_numpy(t["pos_max"]).to(device),
        flags=ops.tree_flags(scene["active"], t["max_depth"]).to(device),
        prefix=torch.from_numpy(scene["prefix"]).to(device),
        grid=g,
        params=torch.from_numpy(scene["params"]).to(device),
        mlp=dict(scene["mlp"]),
        max_spn=cfg.max_spn,
    )


def sample(ds, origin, direction, jitter=False, seed=0, emit=True):
    return ops.raysample(ds.pos_min, ds.pos_max, ds.flags, ds.max_depth, origin, direction, ds.max_spn,
                         jitter=jitter, seed=seed, emit=emit)


def _mlp_args(ds):
    m = ds.mlp
    return (m["in"], m["hidden"], m["layers"], m["out"], m["act"], m["last_act"], m["beta"])


def unfused_forward(ds, smp, num_rays):
    uv = ops.local_samples(smp["coord"], smp["node_id"], ds.pos_min, ds.pos_max)
    slot = ops.node_to_slot(smp["node_id"], ds.prefix)
    feat = ops.gs3d_fwd(ds.grid, slot, uv)
    y = ops.decoder_fwd(feat, ds.params, *_mlp_args(ds))
    proj = ops.integrate_fwd(y, smp["weight"], smp["ray_index"], num_rays, smp["ray_sample_offset"])
    return proj, dict(uv=uv, slot=slot, feat=feat, y=y)


def unfused_backward(ds, smp, inter, grad_proj):
    gy = ops.integrate_bwd(grad_proj, smp["weight"], smp["ray_index"])
    gparams, gfeat = ops.decoder_bwd(inter["feat"], ds.params, gy, *_mlp_args(ds))
    ggrid, _ = ops.gs3d_bwd(ds.grid, inter["slot"], inter["uv"], gfeat, want_uv=False)
    return ggrid, gparams


def unfused_step(ds, origin, direction, target, jitter=False, seed=0):
    """One fwd+bwd pass over a batch: returns (loss [1] float64, proj [1,R], grad_grid, grad_params, num_samples)."""
    R = origin.shape[0]
    smp = sample(ds, origin, direction, jitter, seed)
    proj, inter = unfused_forward(ds, smp, R)
    loss, gproj = ops.mse_fp64(proj.reshape(-1), target.reshape(-1))
    ggrid, gparams = unfused_backward(ds, smp, inter, gproj.reshape(1, R))
    return loss, proj, ggrid, gparams, int(smp["ray_sample_offset"][R].item())


def train_field(ds, phantom, cfg, steps, batch, precision="bf16", lr_grid=1e-2, lr_decoder=1e-3, seed=0):
    """Fits the scene's field to the phantom's projections in place: `steps` Adam steps (the reference trainer's
    learning rates, geometry_ex_ex.cpp:40-49) of the fused projector on `batch` rays per step drawn uniformly over all
    views and pixels, the targets the phantom's exact line integrals (computed on the device). The BF16 form
    refreshes its grid copy after every step. Returns (per-step FP64 losses, the BF16 grid copy or None)."""
    from . import workloads as W
    dev = ds.grid.device
    mg, vg = torch.zeros_like(ds.grid), torch.zeros_like(ds.grid)
    mp, vp = torch.zeros_like(ds.params), torch.zeros_like(ds.params)
    g16 = ops.grid_to_bf16(ds.grid) if precision == "bf16" else None
    losses = []
    for step in range(1, steps + 1):
        o, d = W.batch_rays(cfg, batch, seed=seed + 7919 * step)
        to, td = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
        tgt = phantom.line_integrals_torch(to, td).to(torch.float32)
        segs = sample(ds, to, td, jitter=True, seed=seed + step, emit=False)
        pj = ops.Projector(ds.pos_min, ds.pos_max, ds.prefix, to, td, segs, ds.grid, ds.params, ds.mlp, True,
                           seed + step, precision=precision, grid_bf16=g16)
        loss, gproj = ops.mse_fp64(pj.forward(), tgt)
        gg, gp = pj.backward(gproj)
        ops.adam_step(ds.grid, gg, mg, vg, lr_grid, step)
        ops.adam_step(ds.params, gp, mp, vp, lr_decoder, step)
        if g16 is not None:
            ops.grid_to_bf16(ds.grid, out=g16)
        losses.append(loss)
    return np.array([float(v) for v in losses]), g16
