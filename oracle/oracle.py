"""numpy front end of the CPU oracle (liboracle.so, built from neat_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
always as the checker. The product package never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ACT = {"id": 0, "relu": 1, "silu": 2, "softplus": 3, "abs": 4}

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64
_F = ctypes.c_float


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        sig = {
            "oracle_tree_num_nodes": (_I64, [_I32]),
            "oracle_tree_build": (None, [_I32, _P, _P, _P, _P, _P, _P]),
            "oracle_update_active": (_I64, [_P, _I64, _P, _P]),
            "oracle_intersect": (_I32, [_P, _P, _P, _P, _P]),
            "oracle_jitter": (_F, [_U64, _I64, _I32, _I32]),
            "oracle_raysample_count": (_I64, [_P, _P, _P, _I64, _P, _P, _I64, _I32, _P, _P, _I32]),
            "oracle_raysample_segments": (None, [_P, _P, _P, _I64, _P, _P, _I64, _I32, _P, _P,
                                                 _P, _P, _P, _P, _P, _P, _I32]),
            "oracle_raysample_emit": (None, [_P, _P, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _U64,
                                             _P, _P, _P, _P, _P, _P]),
            "oracle_local_samples": (None, [_P, _P, _I64, _P, _P, _P]),
            "oracle_grid_sample_fwd_f32": (None, [_P, _I64, _I32, _I32, _I32, _I32, _P, _P, _P, _I64, _P]),
            "oracle_grid_sample_fwd_f64": (None, [_P, _I64, _I32, _I32, _I32, _I32, _P, _P, _P, _I64, _P]),
            "oracle_grid_sample_bwd_f32": (None, [_P, _I64, _I32, _I32, _I32, _I32, _P, _P, _P, _I64, _P, _P, _P]),
            "oracle_grid_sample_bwd_f64": (None, [_P, _I64, _I32, _I32, _I32, _I32, _P, _P, _P, _I64, _P, _P, _P]),
            "oracle_decoder_num_params": (_I64, [_I32, _I32, _I32, _I32]),
            "oracle_decoder_fwd_f64": (None, [_P, _I64, _I32, _I32, _I32, _I32, _P, _I32, _I32, _F, _P]),
            "oracle_decoder_fwd_f32": (None, [_P, _I64, _I32, _I32, _I32, _I32, _P, _I32, _I32, _F, _P]),
            "oracle_decoder_bwd_f64": (None, [_P, _I64, _I32, _I32, _I32, _I32, _P, _I32, _I32, _F, _P, _P, _P]),
            "oracle_decoder_bwd_f32": (None, [_P, _I64, _I32, _I32, _I32, _I32, _P, _I32, _I32, _F, _P, _P, _P]),
            "oracle_integrate_fwd": (None, [_P, _P, _P, _I64, _I32, _I64, _P]),
            "oracle_projector_fwd": (None, [_I32, _P, _P, _P, _P, _P, _I64, _P, _I64, _P, _P, _P, _P, _P, _P, _I32,
                                            _U64, _P, _I64, _I32, _I32, _I32, _I32, _P, _I32, _I32, _I32, _I32,
                                            _P, _I32, _I32, _F, _P, _I32]),
            "oracle_projector_bwd": (None, [_I32, _P, _P, _P, _P, _P, _I64, _P, _I64, _P, _P, _P, _P, _P, _P, _I32,
                                            _U64, _P, _I64, _I32, _I32, _I32, _I32, _P, _I64, _I32, _I32, _I32,
                                            _I32, _P, _I32, _I32, _F, _P, _P, _P, _I32]),
            "oracle_projector_ray_grad": (None, [_I32, _P, _P, _P, _P, _P, _I64, _P, _I64, _P, _P, _P, _P, _P, _P,
                                                 _I32, _U64, _P, _I64, _I32, _I32, _I32, _I32, _P, _I32, _I32, _I32,
                                                 _I32, _P, _I32, _I32, _F, _P, _P, _P, _I32]),
            "oracle_max_threads": (_I32, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def max_threads():
    return int(lib().oracle_max_threads())


# ---- tree -----------------------------------------------------------------------------------------------------
def build_tree(max_depth):
    n = int(lib().oracle_tree_num_nodes(max_depth))
    t = dict(parent=np.empty(n, np.int32), children=np.empty((n, 8), np.int32),
             pos_min=np.empty((n, 3), np.float32), pos_max=np.empty((n, 3), np.float32),
             scale=np.empty(n, np.float32), depth=np.empty(n, np.int32))
    lib().oracle_tree_build(max_depth, _p(t["parent"]), _p(t["children"]), _p(t["pos_min"]), _p(t["pos_max"]),
                            _p(t["scale"]), _p(t["depth"]))
    t["max_depth"] = max_depth
    return t


def update_active(node_active):
    na = _c(node_active, np.int32)
    ids = np.empty(na.shape[0], np.int64)
    prefix = np.empty(na.shape[0], np.int64)
    c = lib().oracle_update_active(_p(na), na.shape[0], _p(ids), _p(prefix))
    return ids[:c].copy(), prefix


def intersect(bmin, bmax, o, d):
    out = np.empty(2, np.float32)
    hit = lib().oracle_intersect(_p(_c(bmin, np.float32)), _p(_c(bmax, np.float32)), _p(_c(o, np.float32)),
                                 _p(_c(d, np.float32)), _p(out))
    return bool(hit), float(out[0]), float(out[1])


def jitter(seed, ray, rank, j):
    return float(lib().oracle_jitter(seed, ray, rank, j))


# ---- ray sampling ----------------------------------------------------------------------------------------------
def raysample(tree, active_ids, origin, direction, max_spn, jitter=False, seed=0, nthreads=0, emit=True):
    """Canonical segment list and SampleList columns (HyperTree.cu:206-438)."""
    L = lib()
    nthreads = nthreads or max_threads()
    o = _c(origin, np.float32)
    d = _c(direction, np.float32)
    R = o.shape[0]
    ids = _c(active_ids, np.int64)
    pmin = _c(tree["pos_min"], np.float32)
    pmax = _c(tree["pos_max"], np.float32)
    nseg = np.empty(R, np.int64)
    nsamp = np.empty(R, np.int64)
    L.oracle_raysample_count(_p(pmin), _p(pmax), _p(ids), ids.shape[0], _p(o), _p(d), R, max_spn, _p(nseg),
                             _p(nsamp), nthreads)
    seg_off = np.zeros(R + 1, np.int64)
    seg_off[1:] = np.cumsum(nseg)
    samp_off = np.zeros(R + 1, np.int64)
    samp_off[1:] = np.cumsum(nsamp)
    K = int(seg_off[-1])
    S = int(samp_off[-1])
    r = dict(ray_seg_offset=seg_off, ray_sample_offset=samp_off, seg_node=np.empty(K, np.int32),
             seg_ray=np.empty(K, np.int32), seg_count=np.empty(K, np.int32), seg_tmin=np.empty(K, np.float32),
             seg_tmax=np.empty(K, np.float32), seg_sample_offset=np.empty(K, np.int64))
    L.oracle_raysample_segments(_p(pmin), _p(pmax), _p(ids), ids.shape[0], _p(o), _p(d), R, max_spn, _p(seg_off),
                                _p(samp_off), _p(r["seg_node"]), _p(r["seg_ray"]), _p(r["seg_count"]),
                                _p(r["seg_tmin"]), _p(r["seg_tmax"]), _p(r["seg_sample_offset"]), nthreads)
    if emit:
        r.update(t=np.empty(S, np.float32), weight=np.empty(S, np.float32), coord=np.empty((S, 3), np.float32),
                 ray_index=np.empty(S, np.int64), local_index=np.empty(S, np.int64), node_id=np.empty(S, np.int64))
        L.oracle_raysample_emit(_p(o), _p(d), K, _p(r["seg_node"]), _p(r["seg_ray"]), _p(r["seg_count"]),
                                _p(r["seg_tmin"]), _p(r["seg_tmax"]), _p(r["seg_sample_offset"]), _p(seg_off),
                                _p(samp_off), int(bool(jitter)), seed, _p(r["t"]), _p(r["weight"]), _p(r["coord"]),
                                _p(r["ray_index"]), _p(r["local_index"]), _p(r["node_id"]))
    return r


def local_samples(tree, coord, node_id):
    c = _c(coord, np.float32)
    n = _c(node_id, np.int64)
    out = np.empty_like(c)
    lib().oracle_local_samples(_p(c), _p(n), n.shape[0], _p(_c(tree["pos_min"], np.float32)),
                               _p(_c(tree["pos_max"], np.float32)), _p(out))
    return out


# ---- grid sample -------------------------------------------------------------------------------------------------
def _grid_args(grid):
    g = np.asarray(grid)
    strides = np.array([s // g.itemsize for s in g.strides], np.int64)
    return g, strides


def grid_sample_fwd(grid, index, uv):
    g, strides = _grid_args(grid)
    f64 = g.dtype == np.float64
    dt = np.float64 if f64 else np.float32
    idx = _c(index, np.int64)
    u = _c(uv, dt)
    N, C, D, H, W = g.shape
    out = np.empty((idx.shape[0], C), dt)
    fn = lib().oracle_grid_sample_fwd_f64 if f64 else lib().oracle_grid_sample_fwd_f32
    fn(_p(g), N, C, D, H, W, _p(strides), _p(idx), _p(u), idx.shape[0], _p(out))
    return out


def grid_sample_bwd(grid, index, uv, grad_out):
    g, strides = _grid_args(grid)
    f64 = g.dtype == np.float64
    dt = np.float64 if f64 else np.float32
    idx = _c(index, np.int64)
    u = _c(uv, dt)
    go = _c(grad_out, dt)
    N, C, D, H, W = g.shape
    gg = np.zeros_like(g)  # same strides as the grid
    guv = np.zeros((idx.shape[0], 3), dt)
    fn = lib().oracle_grid_sample_bwd_f64 if f64 else lib().oracle_grid_sample_bwd_f32
    fn(_p(g), N, C, D, H, W, _p(strides), _p(idx), _p(u), idx.shape[0], _p(go), _p(gg), _p(guv))
    return gg, guv


# ---- decoder -----------------------------------------------------------------------------------------------------
def decoder_num_params(d_in, hidden, layers, d_out):
    return int(lib().oracle_decoder_num_params(d_in, hidden, layers, d_out))


def decoder_fwd(x, params, d_in, hidden, layers, d_out, act="relu", last_act="softplus", beta=2.0, f64=True):
    dt = np.float64 if f64 else np.float32
    xx = _c(x, dt)
    p = _c(params, dt)
    y = np.empty((xx.shape[0], d_out), dt)
    fn = lib().oracle_decoder_fwd_f64 if f64 else lib().oracle_decoder_fwd_f32
    fn(_p(xx), xx.shape[0], d_in, hidden, layers, d_out, _p(p), ACT[act], ACT[last_act], beta, _p(y))
    return y


def decoder_bwd(x, params, grad_y, d_in, hidden, layers, d_out, act="relu", last_act="softplus", beta=2.0,
                f64=True):
    dt = np.float64 if f64 else np.float32
    xx = _c(x, dt)
    p = _c(params, dt)
    gy = _c(grad_y, dt)
    gp = np.zeros_like(p)
    gx = np.zeros_like(xx)
    fn = lib().oracle_decoder_bwd_f64 if f64 else lib().oracle_decoder_bwd_f32
    fn(_p(xx), xx.shape[0], d_in, hidden, layers, d_out, _p(p), ACT[act], ACT[last_act], beta, _p(gy), _p(gp),
       _p(gx))
    return gp, gx


def integrate_fwd(values, weight, ray_index, num_rays):
    v = _c(values, np.float32)
    w = _c(weight, np.float32)
    ri = _c(ray_index, np.int64)
    C = v.shape[1]
    out = np.empty((C, num_rays), np.float32)
    lib().oracle_integrate_fwd(_p(v), _p(w), _p(ri), ri.shape[0], C, num_rays, _p(out))
    return out


# ---- fused projector ---------------------------------------------------------------------------------------------
def _proj_common(prec, tree, prefix, origin, direction, segs, grid, mlp, jitter, seed):
    o = _c(origin, np.float32)
    d = _c(direction, np.float32)
    g, strides = _grid_args(grid)
    assert g.dtype == np.float32
    N, C, D, H, W = g.shape
    keep = dict(o=o, d=d, g=g, strides=strides, pmin=_c(tree["pos_min"], np.float32),
                pmax=_c(tree["pos_max"], np.float32), prefix=_c(prefix, np.int64),
                seg_off=_c(segs["ray_seg_offset"], np.int64), sn=_c(segs["seg_node"], np.int32),
                sr=_c(segs["seg_ray"], np.int32), sc=_c(segs["seg_count"], np.int32),
                st0=_c(segs["seg_tmin"], np.float32), st1=_c(segs["seg_tmax"], np.float32),
                params=_c(mlp["params"], np.float32))
    # optional proposal CDF [K, 8] (rows with cdf[7] == 0 sample uniformly)
    keep["cdf"] = _c(segs["seg_cdf"], np.float32) if segs.get("seg_cdf") is not None else None
    args = [prec, _p(keep["pmin"]), _p(keep["pmax"]), _p(keep["prefix"]), _p(o), _p(d), o.shape[0],
            _p(keep["seg_off"]), keep["sn"].shape[0], _p(keep["sn"]), _p(keep["sr"]), _p(keep["sc"]),
            _p(keep["st0"]), _p(keep["st1"]), _p(keep["cdf"]) if keep["cdf"] is not None else None,
            int(bool(jitter)), seed, _p(g), N, C, D, H, W, _p(strides)]
    return keep, args


def projector_fwd(tree, prefix, origin, direction, segs, grid, mlp, jitter=False, seed=0, prec=0, nthreads=0):
    """Per-ray projection (float64 array). prec 0: FP32 restatement, 1: FP64 oracle (structure frozen, t re-derived
    by the FP64 slab test), 2: FP64 on the segments' own t values."""
    keep, args = _proj_common(prec, tree, prefix, origin, direction, segs, grid, mlp, jitter, seed)
    R = keep["o"].shape[0]
    proj = np.empty(R, np.float64)
    lib().oracle_projector_fwd(*args, mlp["in"], mlp["hidden"], mlp["layers"], mlp["out"], _p(keep["params"]),
                               ACT[mlp["act"]], ACT[mlp["last_act"]], mlp["beta"], _p(proj),
                               nthreads or max_threads())
    return proj


def projector_ray_grad(tree, prefix, origin, direction, segs, grid, mlp, grad_proj, jitter=False, seed=0, prec=0,
                       nthreads=0):
    """Gradients of sum(grad_proj * proj) w.r.t. the ray origins and directions [R, 3] each, the samples' t and
    weights held fixed (prec 0: FP32 restatement, 1: FP64 on the segments' t values)."""
    keep, args = _proj_common(prec, tree, prefix, origin, direction, segs, grid, mlp, jitter, seed)
    R = keep["o"].shape[0]
    gp = _c(grad_proj, np.float64)
    go = np.zeros((R, 3), np.float64)
    gd = np.zeros((R, 3), np.float64)
    lib().oracle_projector_ray_grad(*args, mlp["in"], mlp["hidden"], mlp["layers"], mlp["out"], _p(keep["params"]),
                                    ACT[mlp["act"]], ACT[mlp["last_act"]], mlp["beta"], _p(gp), _p(go), _p(gd),
                                    nthreads or max_threads())
    return go, gd


def projector_bwd(tree, prefix, origin, direction, segs, grid, mlp, grad_proj, jitter=False, seed=0, prec=0,
                  nthreads=0):
    """Gradients of sum(grad_proj * proj) w.r.t. the grid (same layout as `grid`) and the packed parameters."""
    keep, args = _proj_common(prec, tree, prefix, origin, direction, segs, grid, mlp, jitter, seed)
    g = keep["g"]
    gp = _c(grad_proj, np.float64)
    # the gradient shares the grid's strides: allocate a buffer covering the strided extent
    span = int(sum((s - 1) * st for s, st in zip(g.shape, keep["strides"]))) + 1
    gg_flat = np.zeros(span, np.float64)
    gpar = np.zeros(keep["params"].shape[0], np.float64)
    lib().oracle_projector_bwd(*args, span, mlp["in"], mlp["hidden"], mlp["layers"], mlp["out"],
                               _p(keep["params"]), ACT[mlp["act"]], ACT[mlp["last_act"]], mlp["beta"], _p(gp),
                               _p(gg_flat), _p(gpar), nthreads or max_threads())
    gg = np.lib.stride_tricks.as_strided(gg_flat, shape=g.shape, strides=tuple(int(s) * 8 for s in keep["strides"]))
    return gg.copy(), gpar
