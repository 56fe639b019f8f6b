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
