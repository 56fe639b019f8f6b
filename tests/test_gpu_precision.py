"""GPU checks of the projector's arithmetic forms: FP32 on the matrix cores (the default), FP32 on the vector ALUs,
and BF16 (BF16 grid copy and hidden-layer weights/activations on the matrix cores, FP32 accumulation, FP32 output
layer), each compared with the FP32 and the FP64 oracle on the same canonical segment list.

Bars: FP32 forms within 1e-5 relative (norm-wise) of the FP32 oracle; BF16 projections within 1e-3 relative of the
FP64 oracle (all measured errors are printed). BF16 gradients: within 1e-2 of the FP64 oracle for a smooth hidden
activation (silu); for relu the bound against the FP64 oracle on the FP32 inputs is 1e-1, because relu' is a step:
BF16 rounding moves pre-activations near 0 across it, and the resulting gradient error grows like the square root of
the fraction of flipped units (about sqrt(2^-9 x density at 0)), not like the rounding error itself (the FP32 path
shows the same effect at 4e-6). Rounding the inputs alone does as much: the FP64 oracle itself, evaluated on the
BF16-rounded grid and weights, is about 3.7e-2 away on C1 and C2, and the BF16 form's distance to that evaluation
(its arithmetic: BF16 features and activations) is of the same size and held to it."""
import numpy as np
import pytest
import torch

from neat_amd import ops, pipeline, workloads as W

pytestmark = pytest.mark.gpu
DEV = "cuda"


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def bf16_bits_np(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def scene(cfg, nrays, seed, hidden=None, act=None):
    sc = W.make_scene(cfg, seed=seed)
    if hidden is not None or act is not None:
        h = hidden or cfg.hidden
        sc["params"] = W.decoder_params(cfg.channels, h, cfg.layers, 1, seed + 2)
        sc["mlp"] = dict(sc["mlp"], hidden=h, act=act or sc["mlp"]["act"])
    ds = pipeline.to_device(sc, cfg)
    o, d = W.batch_rays(cfg, nrays, seed=seed + 7)
    return sc, ds, o, d, torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV)


def fused(ds, to, td, segs, jitter, seed, grad, precision, compensated=False):
    pj = ops.Projector(ds.pos_min, ds.pos_max, ds.prefix, to, td, segs, ds.grid, ds.params, ds.mlp, jitter, seed,
                       precision=precision)
    proj = pj.forward(compensated=compensated)
    gg, gp = pj.backward(grad)
    torch.cuda.synchronize()
    return proj.cpu().numpy(), gg.cpu().numpy(), gp.cpu().numpy()


def oracle_both(oracle, sc, o, d, segs, jitter, seed, gproj, nthreads=0):
    hseg = {k: v.cpu().numpy() for k, v in segs.items() if torch.is_tensor(v)}
    mlp = dict(sc["mlp"], params=sc["params"])
    out = {}
    for prec in (0, 1):
        p = oracle.projector_fwd(sc["tree"], sc["prefix"], o, d, hseg, sc["grid"], mlp, jitter=jitter, seed=seed,
                                 prec=prec, nthreads=nthreads)
        gg, gp = oracle.projector_bwd(sc["tree"], sc["prefix"], o, d, hseg, sc["grid"], mlp,
                                      gproj.astype(np.float64), jitter=jitter, seed=seed, prec=prec,
                                      nthreads=nthreads)
        out[prec] = (p, gg, gp)
    return out


def test_grid_to_bf16_rounds_to_nearest_even():
    print("case: grid_to_bf16 bit patterns", flush=True)
    rng = np.random.default_rng(0)
    x = rng.normal(size=(3, 8, 5, 5, 5)).astype(np.float32)
    # ties: exact halfway values with even and odd upper halves, and values that round across an exponent step
    flat = x.reshape(-1)
    flat[:4] = np.array([0x3F808000, 0x3F818000, 0x3F7FFFFF, 0xBF808001], np.uint32).view(np.float32)
    g = torch.from_numpy(x).to(DEV).contiguous(memory_format=torch.channels_last_3d)
    b = ops.grid_to_bf16(g)
    assert b.stride() == g.stride() and b.dtype == torch.int16
    got = b.cpu().contiguous().numpy().view(np.uint16)
    assert np.array_equal(got, bf16_bits_np(x))
    assert list(got.reshape(-1)[:4]) == [0x3F80, 0x3F82, 0x3F80, 0xBF81]


@pytest.mark.parametrize("precision", ["fp32", "fp32_valu"])
@pytest.mark.parametrize("variant", [dict(), dict(hidden=64, act="silu"), dict(hidden=32, act="silu")])
def test_fp32_forms_match_oracle(oracle, precision, variant):
    print("case: fp32 forms C1 geometry", precision, variant, flush=True)
    cfg = W.C1
    sc, ds, o, d, to, td = scene(cfg, 5000, 11, **variant)
    segs = pipeline.sample(ds, to, td, jitter=True, seed=6, emit=False)
    gproj = np.random.default_rng(3).normal(size=o.shape[0]).astype(np.float32)
    proj, gg, gp = fused(ds, to, td, segs, True, 6, torch.from_numpy(gproj).to(DEV), precision)
    ref = oracle_both(oracle, sc, o, d, segs, True, 6, gproj)
    e32 = [rel(proj, ref[0][0]), rel(gg, ref[0][1]), rel(gp, ref[0][2])]
    e64 = [rel(proj, ref[1][0]), rel(gg, ref[1][1]), rel(gp, ref[1][2])]
    print(precision, "errors vs fp32 oracle (proj, grid, params)", e32, "vs fp64 oracle", e64)
    assert max(e32) <= 1e-5


@pytest.mark.parametrize("cfg_name,nrays", [("C1", 6000), ("C2", 640)])
@pytest.mark.parametrize("act", ["relu", "silu"])
def test_bf16_error_vs_fp64_oracle(oracle, cfg_name, nrays, act):
    print("case: bf16 vs fp64 oracle", cfg_name, nrays, act, flush=True)
    cfg = getattr(W, cfg_name)
    sc, ds, o, d, to, td = scene(cfg, nrays, 21, act=act)
    segs = pipeline.sample(ds, to, td, jitter=True, seed=8, emit=False)
    gproj = np.random.default_rng(4).normal(size=o.shape[0]).astype(np.float32)
    proj, gg, gp = fused(ds, to, td, segs, True, 8, torch.from_numpy(gproj).to(DEV), "bf16")
    ref = oracle_both(oracle, sc, o, d, segs, True, 8, gproj)
    e64 = [rel(proj, ref[1][0]), rel(gg, ref[1][1]), rel(gp, ref[1][2])]
    print("bf16 errors vs fp64 oracle (proj, grid, params)", e64)
    assert e64[0] <= 1e-3
    gbar = 1e-2 if act == "silu" else 1e-1
    assert e64[1] <= gbar and e64[2] <= gbar


def bf16_round_np(x):
    return (bf16_bits_np(x).astype(np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("cfg_name,nrays", [("C1", 6000), ("C2", 640)])
def test_bf16_relu_gradient_vs_fp64_on_its_inputs(oracle, cfg_name, nrays):
    """The BF16 form's relu gradients split into the two errors they carry. Its inputs are rounded to BF16 (grid copy,
    first- and hidden-layer weights; biases are carried as BF16 pairs, about 16 bits, the output layer in FP32): the
    FP64 oracle evaluated on those rounded inputs differs from the FP64 oracle on the FP32 inputs by the rounding of
    the inputs alone, a property of the function, not of the arithmetic (relu' flips for units near 0; about 3.7e-2
    on the grid gradient here, printed). The GPU's arithmetic error is its distance to the FP64 oracle on the rounded
    inputs: the BF16 features and hidden activations flip units the same way (measured 3e-2), so it is held to no
    more than the inputs' rounding does (x 1.25), with the decoder gradient within 1e-2; a wrong operand or fragment
    index would be off by O(1)."""
    print("case: bf16 relu gradient vs fp64 on bf16-rounded inputs", cfg_name, nrays, flush=True)
    cfg = getattr(W, cfg_name)
    sc, ds, o, d, to, td = scene(cfg, nrays, 21, act="relu")
    segs = pipeline.sample(ds, to, td, jitter=True, seed=8, emit=False)
    gproj = np.random.default_rng(4).normal(size=o.shape[0]).astype(np.float32)
    proj, gg, gp = fused(ds, to, td, segs, True, 8, torch.from_numpy(gproj).to(DEV), "bf16")
    h, c = cfg.hidden, cfg.channels
    prm = sc["params"].copy()
    prm[:h * c] = bf16_round_np(prm[:h * c])                                    # W0
    prm[h * c + h:h * c + h + h * h] = bf16_round_np(prm[h * c + h:h * c + h + h * h])  # W1
    sr = dict(sc, grid=bf16_round_np(sc["grid"]), params=prm)
    hseg = {k: v.cpu().numpy() for k, v in segs.items() if torch.is_tensor(v)}
    ref_in = oracle.projector_bwd(sr["tree"], sr["prefix"], o, d, hseg, sr["grid"], dict(sr["mlp"], params=prm),
                                  gproj.astype(np.float64), jitter=True, seed=8, prec=1)
    ref = oracle.projector_bwd(sc["tree"], sc["prefix"], o, d, hseg, sc["grid"], dict(sc["mlp"], params=sc["params"]),
                               gproj.astype(np.float64), jitter=True, seed=8, prec=1)
    floor = (rel(ref_in[0], ref[0]), rel(ref_in[1], ref[1]))
    e_in = (rel(gg, ref_in[0]), rel(gp, ref_in[1]))
    e = (rel(gg, ref[0]), rel(gp, ref[1]))
    print("fp64 on rounded inputs vs fp64 (grid, params) %.3e %.3e; gpu bf16 vs fp64 on rounded inputs %.3e %.3e; "
          "gpu bf16 vs fp64 %.3e %.3e" % (floor + e_in + e), flush=True)
    assert e_in[0] <= 1.25 * floor[0] and e_in[1] <= 1e-2


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_compensated_ray_sums_and_determinism(oracle, precision):
    print("case: compensated sums + determinism", precision, flush=True)
    cfg = W.C1
    sc, ds, o, d, to, td = scene(cfg, 3000, 31)
    segs = pipeline.sample(ds, to, td, jitter=False, seed=0, emit=False)
    gproj = np.ones(o.shape[0], np.float32)
    p1, g1, q1 = fused(ds, to, td, segs, False, 0, torch.from_numpy(gproj).to(DEV), precision, compensated=True)
    p2, g2, q2 = fused(ds, to, td, segs, False, 0, torch.from_numpy(gproj).to(DEV), precision, compensated=True)
    # forward, decoder gradient (fixed-order partial sums) and grid gradient (integer fixed-point scatter) are all
    # bit-reproducible
    assert np.array_equal(p1, p2) and np.array_equal(q1, q2) and np.array_equal(g1, g2)
    ref = oracle_both(oracle, sc, o, d, segs, False, 0, gproj)
    e = rel(p1, ref[1][0])
    print(precision, "compensated projection error vs fp64 oracle", e)
    assert e <= (1e-5 if precision == "fp32" else 1e-3)


@pytest.mark.parametrize("cfg_name,nrays", [("C1", 6000), ("C2", 640)])
def test_bf16_relu_all_units_active(oracle, cfg_name, nrays):
    """relu decoder whose hidden units are all active (biases far above the spread of the pre-activations): relu'
    is then 1 everywhere in both FP64 and BF16, so no unit flips and the BF16 gradient error is the rounding error
    alone, held to the smooth-activation bar (1e-2). Every weight, bias and feature gradient and every sample passes
    through the BF16 products, so a wrong fragment index or operand orientation fails it by O(1)."""
    print("case: bf16 relu all active", cfg_name, nrays, flush=True)
    cfg = getattr(W, cfg_name)
    sc, ds, o, d, to, td = scene(cfg, nrays, 23, act="relu")
    h, c = cfg.hidden, cfg.channels
    prm = sc["params"].copy()
    o_b0 = h * c
    o_w1 = o_b0 + h
    o_b1 = o_w1 + h * h
    o_w2 = o_b1 + h
    prm[o_b0:o_b0 + h] += 2.0
    prm[o_b1:o_b1 + h] += 12.0
    prm[o_w2:o_w2 + h] *= 0.05
    sc["params"] = prm
    ds = pipeline.to_device(sc, cfg)
    segs = pipeline.sample(ds, to, td, jitter=True, seed=9, emit=False)
    gproj = np.random.default_rng(5).normal(size=o.shape[0]).astype(np.float32)
    proj, gg, gp = fused(ds, to, td, segs, True, 9, torch.from_numpy(gproj).to(DEV), "bf16")
    ref = oracle_both(oracle, sc, o, d, segs, True, 9, gproj)
    e64 = [rel(proj, ref[1][0]), rel(gg, ref[1][1]), rel(gp, ref[1][2])]
    # per parameter block, so that a block that is wrong cannot hide under the norm of a larger one
    blocks = {"W0": (0, o_b0), "b0": (o_b0, o_w1), "W1": (o_w1, o_b1), "b1": (o_b1, o_w2), "W2": (o_w2, o_w2 + h),
              "b2": (o_w2 + h, o_w2 + h + 1)}
    eb = {k: rel(gp[a:b], ref[1][2][a:b]) for k, (a, b) in blocks.items()}
    print("bf16 relu all-active errors vs fp64 oracle (proj, grid, params)", e64, "per block", eb)
    assert e64[0] <= 1e-3
    assert e64[1] <= 1e-2 and max(eb.values()) <= 1e-2
