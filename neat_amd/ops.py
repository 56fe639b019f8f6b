"""Thin torch-tensor wrappers over the C ABI (allocation and stream plumbing only; all compute is in the HIP library).

These are what the parity tests drive directly; the libtorch module layer (_neat_host) calls the same entry points.
"""
import ctypes

import numpy as np
import torch

from . import _abi
from ._abi import call, ptr, stream_ptr

_I64 = torch.int64
_I32 = torch.int32
_F32 = torch.float32


def _dev(t):
    assert t.is_cuda, "neat_amd runs on the GPU only"
    return t


def tree_flags(node_active, max_depth):
    """uint8 flags per node (bit 0 active, bit 1 active descendant), computed on the host."""
    na = np.ascontiguousarray(np.asarray(node_active.cpu() if torch.is_tensor(node_active) else node_active),
                              dtype=np.int32)
    out = np.empty(na.shape[0], np.uint8)
    _abi.check(_abi.lib().neat_tree_flags_host(na.ctypes.data, na.shape[0], max_depth, out.ctypes.data),
               "neat_tree_flags_host")
    return torch.from_numpy(out)


def tree_validate(pos_min, pos_max, children, max_depth):
    pmin = np.ascontiguousarray(pos_min, np.float32)
    pmax = np.ascontiguousarray(pos_max, np.float32)
    ch = np.ascontiguousarray(children, np.int32)
    return _abi.lib().neat_tree_validate_host(pmin.ctypes.data, pmax.ctypes.data, ch.ctypes.data, pmin.shape[0],
                                              max_depth)


def raysample(pos_min, pos_max, flags, max_depth, origin, direction, max_spn, jitter=False, seed=0, emit=True,
              stream=None, single_walk=True):
    """Canonical segment list (+ SampleList columns when emit) for rays [R,3] on the GPU: one walk per ray into a
    per-ray scratch and a compaction (single_walk), or the count walk and a second walk that writes the list."""
    dev = origin.device
    s = stream_ptr(stream)
    R = origin.shape[0]
    num_nodes = pos_min.shape[0]
    o = _dev(origin.contiguous())
    d = _dev(direction.contiguous())
    seg_off = torch.empty(R + 1, dtype=_I64, device=dev)
    samp_off = torch.empty(R + 1, dtype=_I64, device=dev)
    maxr = torch.empty(1, dtype=_I32, device=dev)
    ws_bytes = _abi.lib().neat_raysample_workspace_bytes(R)
    ws = torch.empty(max(ws_bytes, 8), dtype=torch.uint8, device=dev)
    if single_walk:
        sc_bytes = _abi.lib().neat_raysample_scratch_bytes(R, max_depth)
        scratch = torch.empty(max(sc_bytes, 16), dtype=torch.uint8, device=dev)
        call("neat_raysample_walk", s, ptr(pos_min), ptr(pos_max), ptr(flags), num_nodes, max_depth, ptr(o), ptr(d),
             R, max_spn, ptr(seg_off), ptr(samp_off), ptr(maxr), ptr(scratch), sc_bytes, ptr(ws), ws_bytes)
    else:
        call("neat_raysample_count", s, ptr(pos_min), ptr(pos_max), ptr(flags), num_nodes, max_depth, ptr(o), ptr(d),
             R, max_spn, ptr(seg_off), ptr(samp_off), ptr(maxr), ptr(ws), ws_bytes)
    tot = torch.stack([seg_off[R], samp_off[R], maxr[0].to(_I64)]).cpu()  # one device -> host read
    K, S, maxs = int(tot[0]), int(tot[1]), int(tot[2])
    r = dict(ray_seg_offset=seg_off, ray_sample_offset=samp_off, max_samples_per_ray=maxs,
             seg_node=torch.empty(K, dtype=_I32, device=dev), seg_ray=torch.empty(K, dtype=_I32, device=dev),
             seg_count=torch.empty(K, dtype=_I32, device=dev), seg_tmin=torch.empty(K, dtype=_F32, device=dev),
             seg_tmax=torch.empty(K, dtype=_F32, device=dev), seg_sample_offset=torch.empty(K, dtype=_I64, device=dev))
    cols = (ptr(r["seg_node"]), ptr(r["seg_ray"]), ptr(r["seg_count"]), ptr(r["seg_tmin"]), ptr(r["seg_tmax"]),
            ptr(r["seg_sample_offset"]))
    if single_walk:
        call("neat_raysample_compact", s, ptr(pos_min), ptr(pos_max), ptr(flags), num_nodes, max_depth, ptr(o), ptr(d),
             R, max_spn, ptr(seg_off), ptr(samp_off), ptr(scratch), sc_bytes, K, *cols)
    else:
        call("neat_raysample_segments", s, ptr(pos_min), ptr(pos_max), ptr(flags), num_nodes, max_depth, ptr(o),
             ptr(d), R, max_spn, ptr(seg_off), ptr(samp_off), K, *cols)
    if emit:
        r.update(t=torch.empty(S, dtype=_F32, device=dev), weight=torch.empty(S, dtype=_F32, device=dev),
                 coord=torch.empty(S, 3, dtype=_F32, device=dev), ray_index=torch.empty(S, dtype=_I64, device=dev),
                 local_index=torch.empty(S, dtype=_I64, device=dev), node_id=torch.empty(S, dtype=_I64, device=dev))
        call("neat_raysample_emit", s, ptr(o), ptr(d), R, ptr(seg_off), ptr(samp_off), K, ptr(r["seg_node"]),
             ptr(r["seg_ray"]), ptr(r["seg_count"]), ptr(r["seg_tmin"]), ptr(r["seg_tmax"]),
             ptr(r["seg_sample_offset"]), S, int(bool(jitter)), seed, ptr(r["t"]), ptr(r["weight"]), ptr(r["coord"]),
             ptr(r["ray_index"]), ptr(r["local_index"]), ptr(r["node_id"]))
    return r


def local_samples(coord, node_id, pos_min, pos_max, stream=None):
    out = torch.empty_like(coord)
    call("neat_local_samples", stream_ptr(stream), ptr(coord.contiguous()), ptr(node_id.contiguous()),
         node_id.shape[0], ptr(pos_min), ptr(pos_max), pos_min.shape[0], ptr(out))
    return out


def node_to_slot(node_id, prefix, stream=None):
    out = torch.empty_like(node_id)
    call("neat_node_to_slot", stream_ptr(stream), ptr(node_id.contiguous()), node_id.shape[0], ptr(prefix),
         prefix.shape[0], ptr(out))
    return out


def _strides(grid):
    return np.ascontiguousarray(grid.stride(), dtype=np.int64)


def gs3d_fwd(grid, index, uv, stream=None):
    N, C, Z, Y, X = grid.shape
    st = _strides(grid)
    out = torch.empty(index.shape[0], C, dtype=_F32, device=grid.device)
    call("neat_gs3d_fwd", stream_ptr(stream), ptr(grid), N, C, Z, Y, X, st.ctypes.data, ptr(index.contiguous()),
         ptr(uv.contiguous()), index.shape[0], ptr(out))
    return out


def gs3d_bwd(grid, index, uv, grad_out, want_uv=True, stream=None):
    N, C, Z, Y, X = grid.shape
    st = _strides(grid)
    gg = torch.zeros_like(grid)  # preserves the grid's memory format, so the strides match
    assert gg.stride() == grid.stride()
    guv = torch.empty(index.shape[0], 3, dtype=_F32, device=grid.device) if want_uv else None
    call("neat_gs3d_bwd", stream_ptr(stream), ptr(grid), N, C, Z, Y, X, st.ctypes.data, ptr(index.contiguous()),
         ptr(uv.contiguous()), index.shape[0], ptr(grad_out.contiguous()), ptr(gg), ptr(guv))
    return gg, guv


def decoder_fwd(x, params, d_in, hidden, layers, d_out, act="relu", last_act="softplus", beta=2.0, stream=None):
    y = torch.empty(x.shape[0], d_out, dtype=_F32, device=x.device)
    call("neat_decoder_fwd", stream_ptr(stream), ptr(x.contiguous()), x.shape[0], d_in, hidden, layers, d_out,
         ptr(params.contiguous()), _abi.ACT[act], _abi.ACT[last_act], beta, ptr(y))
    return y


def decoder_bwd(x, params, grad_y, d_in, hidden, layers, d_out, act="relu", last_act="softplus", beta=2.0,
                stream=None):
    S = x.shape[0]
    gp = torch.empty_like(params)
    gx = torch.empty_like(x)
    wsb = _abi.lib().neat_decoder_workspace_bytes(S, d_in, hidden, layers, d_out)
    ws = torch.empty(max(wsb, 8), dtype=torch.uint8, device=x.device)
    call("neat_decoder_bwd", stream_ptr(stream), ptr(x.contiguous()), S, d_in, hidden, layers, d_out,
         ptr(params.contiguous()), _abi.ACT[act], _abi.ACT[last_act], beta, ptr(grad_y.contiguous()), ptr(gp), ptr(gx),
         ptr(ws), wsb)
    return gp, gx


def integrate_fwd(values, weight, ray_index, num_rays, ray_sample_offset=None, stream=None):
    C = values.shape[1]
    out = torch.empty(C, num_rays, dtype=_F32, device=values.device)
    call("neat_integrate_fwd", stream_ptr(stream), ptr(values.contiguous()), ptr(weight.contiguous()),
         ptr(ray_index.contiguous()), values.shape[0], C, num_rays, ptr(ray_sample_offset), ptr(out))
    return out


def integrate_bwd(grad_out, weight, ray_index, stream=None):
    C, R = grad_out.shape
    gv = torch.empty(weight.shape[0], C, dtype=_F32, device=grad_out.device)
    call("neat_integrate_bwd", stream_ptr(stream), ptr(grad_out.contiguous()), ptr(weight.contiguous()),
         ptr(ray_index.contiguous()), weight.shape[0], C, R, ptr(gv))
    return gv


def mse_fp64(pred, target, grad_scale=1.0, want_grad=True, stream=None):
    n = pred.numel()
    loss = torch.empty(1, dtype=torch.float64, device=pred.device)
    grad = torch.empty_like(pred) if want_grad else None
    wsb = _abi.lib().neat_mse_workspace_bytes(n)
    ws = torch.empty(wsb, dtype=torch.uint8, device=pred.device)
    call("neat_mse_fp64", stream_ptr(stream), ptr(pred.contiguous()), ptr(target.contiguous()), n, grad_scale,
         ptr(loss), ptr(grad), ptr(ws), wsb)
    return loss, grad


def grid_to_bf16(grid, stream=None, out=None):
    """Round-to-nearest-even BF16 bit patterns (int16) of an FP32 grid, same shape and strides (neat_grid_to_bf16);
    written into `out` when given (the per-optimizer-step refresh of an existing copy)."""
    if not grid.is_contiguous(memory_format=torch.channels_last_3d) and not grid.is_contiguous():
        raise ValueError("grid_to_bf16: grid must be dense")
    if out is None:
        out = torch.empty_like(grid, dtype=torch.int16)
    elif out.dtype != torch.int16 or out.shape != grid.shape or out.stride() != grid.stride():
        raise ValueError("grid_to_bf16: out must be int16 with the grid's shape and strides")
    call("neat_grid_to_bf16", stream_ptr(stream), ptr(grid), grid.numel(), ptr(out))
    return out


class Projector:
    """Fused projector over one batch's canonical segment list (neat_projector_plan / _fwd / _bwd).

    Holds references to every tensor whose address is in the argument struct, and the workspace the plan fills. The
    backward follows the forward (the BF16 relu backward reads what the forward left in the workspace), as in the
    reference's autograd; calling it first raises."""

    def __init__(self, pos_min, pos_max, prefix, origin, direction, segs, grid, params, mlp, jitter=False, seed=0,
                 stream=None, precision="fp32", grid_bf16=None):
        self.stream = stream
        N, C, G = grid.shape[0], grid.shape[1], grid.shape[2]
        R = origin.shape[0]
        K = segs["seg_node"].shape[0]
        S = int(segs["ray_sample_offset"][R].item()) if "num_samples" not in segs else segs["num_samples"]
        self._keep = [pos_min, pos_max, prefix, origin, direction, grid, params] + [segs[k] for k in (
            "ray_seg_offset", "seg_node", "seg_ray", "seg_count", "seg_tmin", "seg_tmax")]
        a = _abi.ProjectorArgs()
        a.node_pos_min, a.node_pos_max, a.node_active_prefix_sum = pos_min.data_ptr(), pos_max.data_ptr(), \
            prefix.data_ptr()
        a.num_nodes = pos_min.shape[0]
        a.ray_origin, a.ray_direction, a.num_rays = origin.data_ptr(), direction.data_ptr(), R
        a.ray_seg_offset = segs["ray_seg_offset"].data_ptr()
        a.num_segments = K
        for f in ("seg_node", "seg_ray", "seg_count", "seg_tmin", "seg_tmax"):
            setattr(a, f, segs[f].data_ptr())
        if segs.get("seg_cdf") is not None:
            cdf = segs["seg_cdf"]
            if cdf.shape != (K, 8) or cdf.dtype != _F32 or not cdf.is_contiguous():
                raise ValueError("seg_cdf must be a contiguous float32 [num_segments, 8] tensor")
            self._keep.append(cdf)
            a.seg_cdf = cdf.data_ptr()
        a.num_samples, a.jitter, a.seed = S, int(bool(jitter)), seed
        a.grid, a.num_slots, a.channels, a.grid_size = grid.data_ptr(), N, C, G
        for i, st in enumerate(grid.stride()):
            a.grid_strides[i] = st
        a.in_features, a.hidden, a.hidden_layers, a.out_features = mlp["in"], mlp["hidden"], mlp["layers"], mlp["out"]
        a.params, a.act, a.last_act, a.softplus_beta = params.data_ptr(), _abi.ACT[mlp["act"]], \
            _abi.ACT[mlp["last_act"]], mlp["beta"]
        a.precision = _abi.PREC[precision]
        if precision == "bf16":
            if grid_bf16 is None:
                grid_bf16 = grid_to_bf16(grid, stream)
            if grid_bf16.dtype != torch.int16 or grid_bf16.stride() != grid.stride() or \
                    grid_bf16.shape != grid.shape:
                raise ValueError("grid_bf16 must be an int16 tensor with the grid's shape and strides")
            self._keep.append(grid_bf16)
            a.grid_bf16 = grid_bf16.data_ptr()
        self.grid_bf16 = grid_bf16
        self.args = a
        self.grid, self.params, self.R, self.S = grid, params, R, S
        self._forwarded = False
        nbytes = _abi.lib().neat_projector_workspace_bytes(ctypes.byref(a))
        if nbytes == 0:
            raise _abi.NeatError(_abi.NEAT_E_UNSUPPORTED, "neat_projector_workspace_bytes", _abi.last_error())
        self.ws = torch.empty(nbytes, dtype=torch.uint8, device=grid.device)
        call("neat_projector_plan", stream_ptr(stream), ctypes.byref(a), ptr(self.ws), nbytes)

    def forward(self, compensated=False):
        proj = torch.empty(self.R, dtype=_F32, device=self.grid.device)
        call("neat_projector_fwd", stream_ptr(self.stream), ctypes.byref(self.args), int(bool(compensated)), ptr(proj),
             ptr(self.ws), self.ws.numel())
        self._forwarded = True
        return proj

    def _require_forward(self, what):
        if not self._forwarded:
            raise RuntimeError(f"Projector.{what}: run forward() first (the backward reads what it leaves)")

    def backward_stage(self, stage, grad_proj, grad_grid, grad_params):
        """One stage of the backward (1 decoder, 2 parameter gradient, 3 grid gradient); 1..3 in order = backward."""
        self._require_forward("backward_stage")
        call("neat_projector_bwd_stage", stream_ptr(self.stream), ctypes.byref(self.args), int(stage),
             ptr(grad_proj) if grad_proj is not None else None, ptr(grad_grid), ptr(grad_params), ptr(self.ws),
             self.ws.numel())

    def sample_values(self):
        """(values [S] in the plan's node-major order, seg_value_offset [K]): the decoded field at every sample."""
        values = torch.empty(max(self.S, 1), dtype=_F32, device=self.grid.device)
        off = torch.empty(max(self.args.num_segments, 1), dtype=torch.int64, device=self.grid.device)
        call("neat_projector_sample_values", stream_ptr(self.stream), ctypes.byref(self.args), ptr(values), ptr(off),
             ptr(self.ws), self.ws.numel())
        self._forwarded = True  # the same forward kernels, leaving the same per-sample cache
        return values, off

    def ray_grad(self):
        """(dL/dorigin, dL/ddirection) [R, 3] each of the last backward's loss, the samples' t and weights fixed
        (neat_projector_ray_grad; call after backward or backward_stage(1))."""
        go = torch.empty(self.R, 3, dtype=_F32, device=self.grid.device)
        gd = torch.empty(self.R, 3, dtype=_F32, device=self.grid.device)
        call("neat_projector_ray_grad", stream_ptr(self.stream), ctypes.byref(self.args), ptr(go), ptr(gd),
             ptr(self.ws), self.ws.numel())
        return go, gd

    def backward(self, grad_proj):
        self._require_forward("backward")
        grad_proj = grad_proj.contiguous()
        gg = torch.empty_like(self.grid)
        gp = torch.empty_like(self.params)
        call("neat_projector_bwd", stream_ptr(self.stream), ctypes.byref(self.args), ptr(grad_proj), ptr(gg), ptr(gp),
             ptr(self.ws), self.ws.numel())
        return gg, gp


def adam_step(param, grad, exp_avg, exp_avg_sq, lr, step, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0,
              stream=None):
    """One in-place Adam step (neat_adam_step) on dense FP32 GPU tensors of one layout."""
    for t in (grad, exp_avg, exp_avg_sq):
        if t.shape != param.shape or t.stride() != param.stride() or t.dtype != _F32:
            raise ValueError("adam_step: every buffer must match the parameter's shape, strides and dtype")
    call("neat_adam_step", stream_ptr(stream), ptr(param), ptr(grad), ptr(exp_avg), ptr(exp_avg_sq), param.numel(),
         float(lr), float(beta1), float(beta2), float(eps), float(weight_decay), int(step))


# ---- learned proposal --------------------------------------------------------------------------------------------
def proposal_num_params():
    return int(_abi.lib().neat_proposal_num_params())


def proposal_fwd(origin, direction, segs, params, num_samples=8, want_logits=False, stream=None):
    """(seg_cdf [K, 8], seg_count_out [K] int32, logits or None) for a canonical segment list (neat_proposal_fwd)."""
    K = segs["seg_ray"].shape[0]
    dev = origin.device
    cdf = torch.empty((K, 8), dtype=_F32, device=dev)
    cnt = torch.empty(K, dtype=torch.int32, device=dev)
    lg = torch.empty((K, 8), dtype=_F32, device=dev) if want_logits else None
    call("neat_proposal_fwd", stream_ptr(stream), ptr(origin), ptr(direction), ptr(segs["seg_ray"]),
         ptr(segs["seg_tmin"]), ptr(segs["seg_tmax"]), ptr(segs["seg_count"]), K, ptr(params), int(num_samples),
         ptr(cdf), ptr(cnt), ptr(lg) if lg is not None else None)
    return cdf, cnt, lg


def proposal_targets(values, seg_value_offset, seg_count, stream=None):
    K = seg_count.shape[0]
    q = torch.empty((K, 8), dtype=_F32, device=seg_count.device)
    call("neat_proposal_targets", stream_ptr(stream), ptr(values), ptr(seg_value_offset), ptr(seg_count), K, ptr(q))
    return q


def proposal_bwd(origin, direction, segs, params, target, scale=None, stream=None):
    """Gradient of scale * sum_k CE(target_k, softmax(logits_k)) (default scale: 1 / K, the mean)."""
    K = segs["seg_ray"].shape[0]
    scale = 1.0 / max(K, 1) if scale is None else scale
    gp = torch.empty_like(params)
    nbytes = int(_abi.lib().neat_proposal_workspace_bytes(K))
    ws = torch.empty(max(nbytes, 16), dtype=torch.uint8, device=params.device)
    call("neat_proposal_bwd", stream_ptr(stream), ptr(origin), ptr(direction), ptr(segs["seg_ray"]),
         ptr(segs["seg_tmin"]), ptr(segs["seg_tmax"]), K, ptr(params), ptr(target), float(scale), ptr(gp), ptr(ws),
         ws.numel())
    return gp


def proposal_roughness_update(values, seg_value_offset, segs, prefix, num_slots, roughness, decay, stream=None):
    """Decay the per-slot roughness table and raise it to the field's variation along the given uniform segments."""
    K = segs["seg_count"].shape[0]
    call("neat_proposal_roughness_update", stream_ptr(stream), ptr(values), ptr(seg_value_offset),
         ptr(segs["seg_count"]), ptr(segs["seg_node"]), K, ptr(prefix), prefix.shape[0], int(num_slots),
         float(decay), ptr(roughness))


def proposal_gate(segs, seg_count_out, seg_cdf, prefix, num_slots, roughness, tau, tau_uniform=None, mid_samples=8,
                  stream=None):
    """neat_proposal_gate: tau_uniform None is tau (two tiers: proposal or uniform)."""
    K = segs["seg_count"].shape[0]
    call("neat_proposal_gate", stream_ptr(stream), ptr(segs["seg_node"]), ptr(segs["seg_count"]), K, ptr(prefix),
         prefix.shape[0], int(num_slots), ptr(roughness), float(tau), float(tau if tau_uniform is None else tau_uniform),
         int(mid_samples), ptr(seg_cdf), ptr(seg_count_out))


# ---- volume regulariser (GeometryExEx::VolumeRegularizer, geometry_ex_ex.cpp:86-160) ---------------------------------
def workspace(nbytes, device):
    return torch.empty(max(int(nbytes), 16), dtype=torch.uint8, device=device)


def neighbor_samples(pos_min, pos_max, flags, active_ids, grid_size, epsilon=0.001, stream=None):
    """HyperTreeBaseImpl::NodeNeighborSamples (HyperTree.cu:830-940) in canonical order: SampleList columns of the
    neighbour rows (two per pair: coordinate, node id, ray index, local index, weight), shrunk to the pairs found."""
    A = int(active_ids.shape[0])
    dev = pos_min.device
    max_pairs = A * 6 * grid_size * grid_size
    rows = 2 * max_pairs
    coord = torch.empty(rows, 3, dtype=_F32, device=dev)
    node = torch.empty(rows, dtype=torch.int64, device=dev)
    ray = torch.empty(rows, dtype=torch.int64, device=dev)
    local = torch.empty(rows, dtype=torch.int64, device=dev)
    weight = torch.empty(rows, dtype=_F32, device=dev)
    npairs = torch.zeros(1, dtype=torch.int64, device=dev)
    ws = workspace(_abi.lib().neat_neighbor_pairs_workspace_bytes(A, grid_size), dev)
    call("neat_neighbor_pairs", stream_ptr(stream), ptr(pos_min), ptr(pos_max), ptr(flags), pos_min.shape[0],
         ptr(active_ids.contiguous()), A, grid_size, float(epsilon), max_pairs, ptr(coord), ptr(node), ptr(ray),
         ptr(local), ptr(weight), ptr(npairs), ptr(ws), ws.numel())
    n = 2 * int(npairs.item())
    return dict(global_coordinate=coord[:n], node_id=node[:n], ray_index=ray[:n], local_index_in_ray=local[:n],
                weight=weight[:n], num_pairs=n // 2)


def edge_loss_fwd(features, seed, weight, stream=None):
    """mean over pairs of mean_c |f[2k] - f[2k+1]| * r_k, times weight * C; FP64 scalar tensor."""
    P, C = features.shape[0] // 2, features.shape[1]
    out = torch.empty(1, dtype=torch.float64, device=features.device)
    ws = workspace(_abi.lib().neat_reg_workspace_bytes(P), features.device)
    call("neat_edge_loss_fwd", stream_ptr(stream), ptr(features.contiguous()), P, C, int(seed), float(weight),
         ptr(out), ptr(ws), ws.numel())
    return out


def edge_loss_bwd(features, seed, weight, grad_loss, stream=None):
    P, C = features.shape[0] // 2, features.shape[1]
    g = torch.empty_like(features)
    call("neat_edge_loss_bwd", stream_ptr(stream), ptr(features.contiguous()), P, C, int(seed), float(weight),
         ptr(grad_loss.to(_F32).contiguous()), ptr(g))
    return g


def tv_loss_fwd(grid, slots, weight, stream=None):
    """Total variation of grid[slots] (TVLoss::forward) times weight; FP64 scalar tensor."""
    N, C, G = grid.shape[0], grid.shape[1], grid.shape[2]
    st = _strides(grid)
    A = int(slots.shape[0])
    out = torch.empty(1, dtype=torch.float64, device=grid.device)
    ws = workspace(_abi.lib().neat_tv_workspace_bytes(A, C, G), grid.device)
    call("neat_tv_loss_fwd", stream_ptr(stream), ptr(grid), N, C, G, st.ctypes.data, ptr(slots.contiguous()), A,
         float(weight), ptr(out), ptr(ws), ws.numel())
    return out


def tv_loss_bwd(grid, slots, weight, grad_loss, grad_grid, stream=None):
    """Adds d tv / d grid * grad_loss to grad_grid (same strides as grid)."""
    N, C, G = grid.shape[0], grid.shape[1], grid.shape[2]
    assert grad_grid.stride() == grid.stride()
    st = _strides(grid)
    call("neat_tv_loss_bwd", stream_ptr(stream), ptr(grid), N, C, G, st.ctypes.data, ptr(slots.contiguous()),
         int(slots.shape[0]), float(weight), ptr(grad_loss.to(_F32).contiguous()), ptr(grad_grid))


def edge_loss_direct(pos_min, pos_max, flags, prefix, active_ids, grid, seed, weight, epsilon=0.001, want_loss=True,
                     grad_loss=None, grad_grid=None, stream=None):
    """The edge loss in one pass over the candidate face points (neat_edge_loss_direct): returns the FP64 loss tensor
    (or None) and adds d loss / d grid * grad_loss to grad_grid when given (same strides as grid)."""
    N, C, G = grid.shape[0], grid.shape[1], grid.shape[2]
    A = int(active_ids.shape[0])
    st = _strides(grid)
    loss = torch.empty(1, dtype=torch.float64, device=grid.device) if want_loss else None
    if grad_grid is not None:
        assert grad_grid.stride() == grid.stride()
        grad_loss = grad_loss.to(_F32).contiguous()
    ws = workspace(_abi.lib().neat_edge_loss_direct_workspace_bytes(A, G), grid.device)
    call("neat_edge_loss_direct", stream_ptr(stream), ptr(pos_min), ptr(pos_max), ptr(flags), pos_min.shape[0],
         ptr(active_ids.contiguous()), A, ptr(prefix), ptr(grid), N, C, G, st.ctypes.data, float(epsilon), int(seed),
         float(weight), ptr(loss), ptr(grad_loss) if grad_grid is not None else None, ptr(grad_grid), ptr(ws),
         ws.numel())
    return loss


def tv_loss_fwd_bwd(grid, slots, weight, grad_loss, grad_grid, overwrite=False, stream=None):
    """Total variation and its gradient in one pass: returns the FP64 loss; writes (overwrite) or adds
    d tv / d grid * grad_loss into grad_grid's given slots."""
    N, C, G = grid.shape[0], grid.shape[1], grid.shape[2]
    assert grad_grid.stride() == grid.stride()
    st = _strides(grid)
    A = int(slots.shape[0])
    out = torch.empty(1, dtype=torch.float64, device=grid.device)
    ws = workspace(_abi.lib().neat_tv_workspace_bytes(A, C, G), grid.device)
    call("neat_tv_loss_fwd_bwd", stream_ptr(stream), ptr(grid), N, C, G, st.ctypes.data, ptr(slots.contiguous()), A,
         float(weight), ptr(out), ptr(grad_loss.to(_F32).contiguous()), ptr(grad_grid), int(bool(overwrite)), ptr(ws),
         ws.numel())
    return out


def volume_regularizer(pos_min, pos_max, flags, prefix, active_ids, grid, loss_tv=1e-4, loss_edge=1e-3, seed=0,
                       epsilon=0.001, stream=None):
    """GeometryExEx::VolumeRegularizer's value and grid gradient: tv (loss_tv > 0) plus the edge loss (loss_edge > 0)
    along the faces of neighbouring active nodes, the edge loss in one fused pass. Returns (loss FP64 scalar tensor,
    grad_grid)."""
    one = torch.ones(1, dtype=_F32, device=grid.device)
    loss = torch.zeros(1, dtype=torch.float64, device=grid.device)
    if loss_tv > 0:
        # the active nodes hold slots 0 .. A-1 (prefix is the exclusive prefix sum of the active flags), which the
        # total variation's gradient overwrites; the remaining slots are cleared
        A = int(active_ids.shape[0])
        gg = torch.empty_like(grid)
        if grid.shape[0] > A:
            gg[A:].zero_()
        loss += tv_loss_fwd_bwd(grid, prefix[active_ids], loss_tv, one, gg, overwrite=True, stream=stream)
    else:
        gg = torch.zeros_like(grid)
    if loss_edge > 0:
        loss += edge_loss_direct(pos_min, pos_max, flags, prefix, active_ids, grid, seed, loss_edge, epsilon,
                                 grad_loss=one, grad_grid=gg, stream=stream)
    return loss, gg


def volume_regularizer_unfused(pos_min, pos_max, flags, prefix, active_ids, grid, loss_tv=1e-4, loss_edge=1e-3,
                               seed=0, epsilon=0.001, stream=None):
    """The same value and gradient through the reference's chain of steps (neighbour rows, local coordinates, grid
    sample, edge loss, and back through the grid sample)."""
    G = grid.shape[2]
    gg = torch.zeros_like(grid)
    one = torch.ones(1, dtype=_F32, device=grid.device)
    loss = torch.zeros(1, dtype=torch.float64, device=grid.device)
    if loss_tv > 0:
        slots = prefix[active_ids]
        loss += tv_loss_fwd(grid, slots, loss_tv, stream)
        tv_loss_bwd(grid, slots, loss_tv, one, gg, stream)
    if loss_edge > 0:
        smp = neighbor_samples(pos_min, pos_max, flags, active_ids, G, epsilon, stream)
        if smp["num_pairs"] > 0:
            uv = local_samples(smp["global_coordinate"], smp["node_id"], pos_min, pos_max, stream)
            slot = node_to_slot(smp["node_id"], prefix, stream)
            f = gs3d_fwd(grid, slot, uv, stream)
            loss += edge_loss_fwd(f, seed, loss_edge, stream)
            gf = edge_loss_bwd(f, seed, loss_edge, one, stream)
            g2, _ = gs3d_bwd(grid, slot, uv, gf, want_uv=False, stream=stream)
            gg += g2
    return loss, gg


# ---- per-node loss of the evaluation pass ------------------------------------------------------------------------------
def node_loss_accumulate(segs, per_ray_loss, num_nodes, jitter=False, seed=0, stream=None):
    """AccumulateSampleLossPerNode (geometry.cpp:38-58) over a canonical segment list (neat_node_loss_accumulate):
    (loss_sum, weight_sum), float [num_nodes]."""
    dev = per_ray_loss.device
    K = int(segs["seg_node"].shape[0])
    R = int(segs["ray_seg_offset"].shape[0]) - 1
    ls = torch.empty(num_nodes, dtype=_F32, device=dev)
    wsum = torch.empty(num_nodes, dtype=_F32, device=dev)
    ws = workspace(_abi.lib().neat_node_loss_workspace_bytes(K, num_nodes), dev)
    call("neat_node_loss_accumulate", stream_ptr(stream), ptr(segs["ray_seg_offset"]), R, ptr(segs["seg_node"]),
         ptr(segs["seg_ray"]), ptr(segs["seg_count"]), ptr(segs["seg_tmin"]), ptr(segs["seg_tmax"]), K,
         int(bool(jitter)), int(seed), ptr(per_ray_loss.to(_F32).contiguous()), num_nodes, ptr(ls), ptr(wsum),
         ptr(ws), ws.numel())
    return ls, wsum


# ---- camera poses of the ray generation ---------------------------------------------------------------------------------
def rotate_points(rotation, points, index, stream=None):
    """q[index] * p in FP64 (neat_rotate_points_fwd)."""
    p = points.to(torch.float64).contiguous()
    idx = index.to(torch.int32).contiguous()
    out = torch.empty_like(p)
    call("neat_rotate_points_fwd", stream_ptr(stream), ptr(p), ptr(idx), ptr(rotation.contiguous()), p.shape[0],
         rotation.shape[0], ptr(out))
    return out


def rotate_points_bwd(rotation, points, index, grad_out, stream=None):
    """(grad_tangent [poses, 3], grad_points [N, 3]) of rotate_points (neat_rotate_points_bwd)."""
    p = points.to(torch.float64).contiguous()
    idx = index.to(torch.int32).contiguous()
    gt = torch.zeros(rotation.shape[0], 3, dtype=torch.float64, device=p.device)
    gp = torch.empty_like(p)
    call("neat_rotate_points_bwd", stream_ptr(stream), ptr(p), ptr(idx), ptr(rotation.contiguous()), p.shape[0],
         rotation.shape[0], ptr(grad_out.to(torch.float64).contiguous()), ptr(gt), ptr(gp))
    return gt, gp


def so3_apply_tangent(tangent, rotation, stream=None):
    """rotation <- exp(tangent) * rotation, tangent <- 0, in place (neat_so3_apply_tangent)."""
    assert tangent.is_contiguous() and rotation.is_contiguous()
    call("neat_so3_apply_tangent", stream_ptr(stream), ptr(tangent), ptr(rotation), rotation.shape[0])
