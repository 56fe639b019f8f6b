"""Online-trained proposal for in-node sampling (no reference counterpart).

A per-segment network (neat_proposal_fwd: 6 -> 32 -> 32 -> 8, BF16 on the matrix cores) proposes a piecewise-constant
density over 8 bins of each segment; segments with more than `num_samples` uniform samples then get `num_samples`
samples in equal-probability cells of it (neat_projector_args::seg_cdf), the others keep their uniform samples (the
fallback). Training is online on the field's own values: a subset of rays is sampled uniformly (a = 0.5), the field
is decoded at those samples (neat_projector_sample_values), the per-bin variation gives the target
(neat_proposal_targets), and the network takes an Adam step on the cross entropy (neat_proposal_bwd,
neat_adam_step). All of it runs in this package's kernels.
"""
import numpy as np
import torch

from . import ops


def init_params(seed=0):
    """Xavier-uniform weights and zero biases in the packed layout W0 [32,6] b0 W1 [32,32] b1 W2 [8,32] b2."""
    rng = np.random.default_rng(seed)
    out = []
    for fi, fo in ((6, 32), (32, 32), (32, 8)):
        a = np.sqrt(6.0 / (fi + fo))
        out += [rng.uniform(-a, a, fo * fi), np.zeros(fo)]
    p = np.concatenate(out).astype(np.float32)
    assert p.size == ops.proposal_num_params()
    return p


class ProposalSampler:
    """The network, its Adam state, and the fallback's per-slot roughness table (the largest variation of the field
    along a segment seen in the slot, decayed every training step): segments in slots rougher than `tau` get
    `mid_samples` proposal samples if their slot is at most `tau_uniform` rough, and keep their uniform samples
    otherwise (tau None: no fallback; tau_uniform None: tau, i.e. proposal or uniform)."""

    def __init__(self, device="cuda", num_samples=8, lr=1e-3, seed=0, num_slots=0, tau=None, decay=0.98,
                 tau_uniform=None, mid_samples=32):
        self.params = torch.from_numpy(init_params(seed)).to(device)
        self.exp_avg = torch.zeros_like(self.params)
        self.exp_avg_sq = torch.zeros_like(self.params)
        self.num_samples = num_samples
        self.lr = lr
        self.steps = 0
        self.tau = tau
        self.tau_uniform = tau_uniform
        self.mid_samples = mid_samples
        self.decay = decay
        self.roughness = torch.zeros(max(num_slots, 1), dtype=torch.float32, device=device)
        self.num_slots = num_slots

    def apply(self, origin, direction, segs, prefix=None):
        """The segment list with proposal counts and CDF rows (a new dict; the input is not modified). With a
        fallback threshold, `prefix` (node -> slot) is needed."""
        cdf, cnt, _ = ops.proposal_fwd(origin, direction, segs, self.params, self.num_samples)
        if self.tau is not None:
            ops.proposal_gate(segs, cnt, cdf, prefix, self.num_slots, self.roughness, self.tau, self.tau_uniform,
                              self.mid_samples)
        out = dict(segs)
        out["seg_count"] = cnt
        out["seg_cdf"] = cdf
        out["num_samples"] = int(cnt.sum().item()) if cnt.numel() else 0
        return out

    def tier_fractions(self, segs, prefix):
        """Shares of the segments with more than num_samples uniform samples that the gate sends to each tier:
        (num_samples proposal samples, up to mid_samples proposal samples, uniform)."""
        n = segs["seg_count"]
        big = n > self.num_samples
        nb = max(int(big.sum().item()), 1)
        slot = prefix[segs["seg_node"].long()]
        ok = (slot >= 0) & (slot < self.roughness.shape[0])
        r = torch.where(ok, self.roughness[slot.clamp(0, self.roughness.shape[0] - 1)],
                        torch.full_like(n, float("inf"), dtype=torch.float32))
        tu = self.tau if self.tau_uniform is None else self.tau_uniform
        lo = big & (r <= self.tau)
        mid = big & (r > self.tau) & (r <= tu)
        return (float(lo.sum().item()) / nb, float(mid.sum().item()) / nb, 1.0 - float((lo | mid).sum().item()) / nb)

    def targets(self, ds, origin, direction, segs_uniform, precision="bf16", grid_bf16=None):
        """Bin targets of a uniformly sampled (a = 0.5) segment list from the field's values."""
        pj = ops.Projector(ds.pos_min, ds.pos_max, ds.prefix, origin, direction, segs_uniform, ds.grid, ds.params,
                           ds.mlp, False, 0, precision=precision, grid_bf16=grid_bf16)
        values, off = pj.sample_values()
        if self.tau is not None:
            ops.proposal_roughness_update(values, off, segs_uniform, ds.prefix, self.num_slots, self.roughness,
                                          self.decay)
        return ops.proposal_targets(values, off, segs_uniform["seg_count"])

    def train_step(self, ds, origin, direction, segs_uniform, precision="bf16", grid_bf16=None):
        """One Adam step of the network on a uniformly sampled (a = 0.5) segment list; returns the targets."""
        q = self.targets(ds, origin, direction, segs_uniform, precision, grid_bf16)
        g = ops.proposal_bwd(origin, direction, segs_uniform, self.params, q)
        self.steps += 1
        ops.adam_step(self.params, g, self.exp_avg, self.exp_avg_sq, self.lr, self.steps)
        return q

    def loss(self, origin, direction, segs, q):
        """Mean cross entropy of the current network against targets q (for monitoring)."""
        _, _, lg = ops.proposal_fwd(origin, direction, segs, self.params, self.num_samples, want_logits=True)
        return float(-(q * torch.log_softmax(lg, 1)).sum(1).mean().item())
