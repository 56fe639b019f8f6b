"""ctypes binding of libneat_amd.so (the C ABI declared in include/neat_amd.h).

torch is imported first so that the HIP runtime torch ships is the one the library binds to. Loading fails loudly
when the library has not been built: there is no CPU fallback.
"""
import ctypes
import os

import torch  # noqa: F401  (loads torch's HIP runtime before libneat_amd)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libneat_amd.so")

NEAT_OK, NEAT_E_ARG, NEAT_E_OVERFLOW, NEAT_E_HIP, NEAT_E_UNSUPPORTED = 0, 1, 2, 3, 4
ACT = {"id": 0, "relu": 1, "silu": 2, "softplus": 3, "abs": 4}

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64
_F = ctypes.c_float
_SZ = ctypes.c_size_t

SIGNATURES = {
    "neat_last_error": (ctypes.c_char_p, []),
    "neat_abi_version": (ctypes.c_int, []),
    "neat_build_id": (ctypes.c_char_p, []),
    "neat_tree_num_nodes": (_I64, [_I32]),
    "neat_tree_build_host": (ctypes.c_int, [_I32, _P, _P, _P, _P, _P, _P]),
    "neat_tree_update_active_host": (_I64, [_P, _I32, _P, _P]),
    "neat_tree_flags_host": (ctypes.c_int, [_P, _I32, _I32, _P]),
    "neat_tree_validate_host": (ctypes.c_int, [_P, _P, _P, _I32, _I32]),
    "neat_raysample_workspace_bytes": (_SZ, [_I64]),
    "neat_raysample_count": (ctypes.c_int, [_P, _P, _P, _P, _I32, _I32, _P, _P, _I64, _I32, _P, _P, _P, _P, _SZ]),
    "neat_raysample_segments": (ctypes.c_int, [_P, _P, _P, _P, _I32, _I32, _P, _P, _I64, _I32, _P, _P, _I64,
                                               _P, _P, _P, _P, _P, _P]),
    "neat_raysample_count_noalloc": (ctypes.c_int, [_P, _P, _P, _P, _I32, _I32, _P, _P, _I64, _I32, _P, _P, _P]),
    "neat_raysample_scratch_bytes": (_SZ, [_I64, _I32]),
    "neat_raysample_walk": (ctypes.c_int, [_P, _P, _P, _P, _I32, _I32, _P, _P, _I64, _I32, _P, _P, _P, _P, _SZ, _P,
                                           _SZ]),
    "neat_raysample_compact": (ctypes.c_int, [_P, _P, _P, _P, _I32, _I32, _P, _P, _I64, _I32, _P, _P, _P, _SZ, _I64,
                                              _P, _P, _P, _P, _P, _P]),
    "neat_raysample_emit": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _I64, _I32, _U64,
                                           _P, _P, _P, _P, _P, _P]),
    "neat_local_samples": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P, _I32, _P]),
    "neat_local_samples_bwd": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P, _I32, _P]),
    "neat_node_to_slot": (ctypes.c_int, [_P, _P, _I64, _P, _I32, _P]),
    "neat_gs3d_fwd": (ctypes.c_int, [_P, _P, _I64, _I32, _I32, _I32, _I32, _P, _P, _P, _I64, _P]),
    "neat_gs3d_bwd": (ctypes.c_int, [_P, _P, _I64, _I32, _I32, _I32, _I32, _P, _P, _P, _I64, _P, _P, _P]),
    "neat_decoder_num_params": (_I64, [_I32, _I32, _I32, _I32]),
    "neat_decoder_fwd": (ctypes.c_int, [_P, _P, _I64, _I32, _I32, _I32, _I32, _P, _I32, _I32, _F, _P]),
    "neat_decoder_workspace_bytes": (_SZ, [_I64, _I32, _I32, _I32, _I32]),
    "neat_decoder_bwd": (ctypes.c_int, [_P, _P, _I64, _I32, _I32, _I32, _I32, _P, _I32, _I32, _F, _P, _P, _P,
                                        _P, _SZ]),
    "neat_integrate_fwd": (ctypes.c_int, [_P, _P, _P, _P, _I64, _I32, _I64, _P, _P]),
    "neat_integrate_bwd": (ctypes.c_int, [_P, _P, _P, _P, _I64, _I32, _I64, _P]),
    "neat_mse_workspace_bytes": (_SZ, [_I64]),
    "neat_mse_fp64": (ctypes.c_int, [_P, _P, _P, _I64, _F, _P, _P, _P, _SZ]),
}



class ProjectorArgs(ctypes.Structure):
    """Mirror of neat_projector_args (include/neat_amd.h)."""
    _fields_ = [
        ("node_pos_min", _P), ("node_pos_max", _P), ("node_active_prefix_sum", _P), ("num_nodes", _I32),
        ("ray_origin", _P), ("ray_direction", _P), ("num_rays", _I64), ("ray_seg_offset", _P),
        ("num_segments", _I64), ("seg_node", _P), ("seg_ray", _P), ("seg_count", _P), ("seg_tmin", _P),
        ("seg_tmax", _P), ("num_samples", _I64), ("jitter", _I32), ("seed", _U64),
        ("grid", _P), ("num_slots", _I64), ("channels", _I32), ("grid_size", _I32), ("grid_strides", _I64 * 5),
        ("in_features", _I32), ("hidden", _I32), ("hidden_layers", _I32), ("out_features", _I32),
        ("params", _P), ("act", _I32), ("last_act", _I32), ("softplus_beta", _F),
        ("precision", _I32), ("grid_bf16", _P), ("seg_cdf", _P),
    ]


PREC = {"fp32": 0, "bf16": 1, "fp32_valu": 2}


_PA = ctypes.POINTER(ProjectorArgs)
SIGNATURES.update({
    "neat_projector_workspace_bytes": (_SZ, [_PA]),
    "neat_projector_plan": (ctypes.c_int, [_P, _PA, _P, _SZ]),
    "neat_projector_fwd": (ctypes.c_int, [_P, _PA, _I32, _P, _P, _SZ]),
    "neat_projector_bwd": (ctypes.c_int, [_P, _PA, _P, _P, _P, _P, _SZ]),
    "neat_projector_ray_grad": (ctypes.c_int, [_P, _PA, _P, _P, _P, _SZ]),
    "neat_grid_to_bf16": (ctypes.c_int, [_P, _P, _I64, _P]),
    "neat_projector_bwd_stage": (ctypes.c_int, [_P, _PA, _I32, _P, _P, _P, _P, _SZ]),
    "neat_projector_sample_values": (ctypes.c_int, [_P, _PA, _P, _P, _P, _SZ]),
    "neat_proposal_num_params": (_I64, []),
    "neat_proposal_fwd": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _I64, _P, _I32, _P, _P, _P]),
    "neat_proposal_targets": (ctypes.c_int, [_P, _P, _P, _P, _I64, _P]),
    "neat_proposal_workspace_bytes": (_SZ, [_I64]),
    "neat_proposal_roughness_update": (ctypes.c_int, [_P, _P, _P, _P, _P, _I64, _P, _I32, _I64, _F, _P]),
    "neat_proposal_gate": (ctypes.c_int, [_P, _P, _P, _I64, _P, _I32, _I64, _P, _F, _F, _I32, _P, _P]),
    "neat_proposal_bwd": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _I64, _P, _P, _F, _P, _P, _SZ]),
    "neat_adam_step": (ctypes.c_int, [_P, _P, _P, _P, _P, _I64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, _I64]),
    "neat_neighbor_pairs_workspace_bytes": (_SZ, [_I64, _I32]),
    "neat_neighbor_pairs": (ctypes.c_int, [_P, _P, _P, _P, _I32, _P, _I64, _I32, _F, _I64, _P, _P, _P, _P, _P, _P, _P,
                                           _SZ]),
    "neat_reg_workspace_bytes": (_SZ, [_I64]),
    "neat_tv_workspace_bytes": (_SZ, [_I64, _I32, _I32]),
    "neat_rotate_points_fwd": (ctypes.c_int, [_P, _P, _P, _P, _I64, _I64, _P]),
    "neat_rotate_points_bwd": (ctypes.c_int, [_P, _P, _P, _P, _I64, _I64, _P, _P, _P]),
    "neat_so3_apply_tangent": (ctypes.c_int, [_P, _P, _P, _I64]),
    "neat_uniform_volume_samples": (ctypes.c_int, [_P, _P, _P, _P, _I32, _I64, _P, _I32, _P, _P, _P, _P]),
    "neat_tv_loss_fwd_bwd": (ctypes.c_int, [_P, _P, _I64, _I32, _I32, _P, _P, _I64, ctypes.c_double, _P, _P, _P,
                                            _I32, _P, _SZ]),
    "neat_edge_loss_fwd": (ctypes.c_int, [_P, _P, _I64, _I32, _U64, ctypes.c_double, _P, _P, _SZ]),
    "neat_edge_loss_bwd": (ctypes.c_int, [_P, _P, _I64, _I32, _U64, ctypes.c_double, _P, _P]),
    "neat_tv_loss_fwd": (ctypes.c_int, [_P, _P, _I64, _I32, _I32, _P, _P, _I64, ctypes.c_double, _P, _P, _SZ]),
    "neat_tv_loss_bwd": (ctypes.c_int, [_P, _P, _I64, _I32, _I32, _P, _P, _I64, ctypes.c_double, _P, _P]),
    "neat_edge_loss_direct_workspace_bytes": (_SZ, [_I64, _I32]),
    "neat_edge_loss_direct": (ctypes.c_int, [_P, _P, _P, _P, _I32, _P, _I64, _P, _P, _I64, _I32, _I32, _P, _F, _U64,
                                             ctypes.c_double, _P, _P, _P, _P, _SZ]),
    "neat_node_loss_workspace_bytes": (_SZ, [_I64, _I32]),
    "neat_node_loss_accumulate": (ctypes.c_int, [_P, _P, _I64, _P, _P, _P, _P, _P, _I64, _I32, _U64, _P, _I32, _P, _P,
                                                 _P, _SZ]),
})

_LIB = None


class NeatError(RuntimeError):
    def __init__(self, code, fn, msg):
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m neat_amd.build` (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        from .buildinfo import source_id
        built, want = L.neat_build_id().decode(), source_id("abi")
        if built != want:
            raise ImportError(f"{LIB_PATH} was built from other sources (build id {built}, sources {want}); rebuild "
                              "with `python -m neat_amd.build`")
        _LIB = L
    return _LIB


def build_id():
    """Source hash compiled into the loaded library (equal to the sources' by construction of lib())."""
    return lib().neat_build_id().decode()


def last_error():
    return lib().neat_last_error().decode()


def check(code, fn):
    if code != NEAT_OK:
        raise NeatError(code, fn, last_error())


def ptr(t):
    """Device (or host) address of a tensor, None for None."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def call(name, *args):
    """Calls an entry point and raises NeatError on a non-zero status."""
    f = getattr(lib(), name)
    r = f(*args)
    if f.restype is ctypes.c_int:
        check(r, name)
    return r
