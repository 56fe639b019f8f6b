// Wave-level decoder MLP on CDNA4 matrix cores, for tiles of 32 samples per wave.
//
// The MLP is evaluated transposed, H^T = W X^T: the M dimension is the hidden width (one or two 32-row tiles), N is
// the tile's 32 samples, K the layer input. With 32x32 MFMA shapes, lane l of the wave owns sample n = l % 32 and
// lane group g = l / 32; the accumulator register i of lane l holds row crow(i, g) = (i/4)*8 + g*4 + i%4 of column
// n. The B operand of the next layer needs, per lane, the K slots (l/32)*KS + j of column n: taking the K order of
// the next layer to be "register i of group g" makes the accumulator registers themselves the B operand, so layers
// chain in registers. The weight fragments are laid out with the same K permutation once per workgroup, in LDS.
//
// Precisions: FP32 on v_mfma_f32_32x32x2f32 (KS = 1 slot per lane per step); BF16 on v_mfma_f32_32x32x8bf16_1k for
// the 8-channel first layer and v_mfma_f32_32x32x16_bf16 (KS = 8) for hidden layers, FP32 accumulation.
// Lane layouts are those of the instructions (A: row l%32, K slots (l/32)*KS..; B: K slots (l/32)*KS.., column l%32;
// C: crow(i, l/32), column l%32).
#pragma once

#include <type_traits>

#include "common.hpp"

namespace neat
{

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ int crow(int i, int g) { return ((i >> 2) << 3) + (g << 2) + (i & 3); }

__device__ __forceinline__ short bf16_bits(float f)
{
    // round to nearest even (no NaN inputs on this path)
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

__device__ __forceinline__ float bf16_to_f32(unsigned short b) { return __uint_as_float(((uint32_t)b) << 16); }

struct PrecF32
{
    static constexpr int KS = 1;   // K slots per lane per hidden-layer step
    static constexpr bool kBF16 = false;
    typedef float Frag;            // hidden-layer A/B fragment
    typedef float Frag1;           // first-layer fragment (one channel per step, 4 steps)
    static constexpr int kSteps1 = 4;
    __device__ static __forceinline__ f32x16 mma(Frag a, Frag b, f32x16 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ f32x16 mma1(Frag1 a, Frag1 b, f32x16 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    }
    // B fragment of step st from accumulator registers (after an elementwise map f)
    template <typename F>
    __device__ static __forceinline__ Frag from_acc(const f32x16& z, int st, F&& f)
    {
        return f(z[st], st);
    }
};

struct PrecBF16
{
    static constexpr int KS = 8;
    static constexpr bool kBF16 = true;
    typedef bf16x8 Frag;
    typedef s16x4 Frag1;           // 4 channels per lane, one K = 8 step
    static constexpr int kSteps1 = 1;
    __device__ static __forceinline__ f32x16 mma(Frag a, Frag b, f32x16 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ f32x16 mma1(Frag1 a, Frag1 b, f32x16 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
    }
    template <typename F>
    __device__ static __forceinline__ Frag from_acc(const f32x16& z, int st, F&& f)
    {
        Frag b;
#pragma unroll
        for (int j = 0; j < 8; ++j) b[j] = (__bf16)f(z[st * 8 + j], st * 8 + j);
        return b;
    }
};

// Steps per 32-row source tile of a hidden layer.
template <typename P>
constexpr int steps_per_tile()
{
    return 16 / P::KS;
}

// ---- weight fragments in LDS -------------------------------------------------------------------------------------
// Layout (per workgroup), all indexed [...][lane]:
//   W1f [MT][kSteps1][64]         first layer, dst row 32mt + l%32, channel (l/32)*4 + step (FP32) or + j (BF16)
//   W2f [MT][MT][S][64]           hidden layer, dst row 32mt + l%32, src row 32ms + crow(st*KS + j, l/32)
//   W2Tf[MT][MT][S][64]           transposed: dst = layer-1 row 32mt + l%32, src = layer-2 row 32ms + crow(...)
//   W1Tf[MT][S][64]               dX: dst channel l%32 (< C, else 0), src = layer-1 row 32ms + crow(...)
//   vec  [4][64]                  b1, b2, w3 (fp32), b3 at vec[3][0]
template <typename P, int H>
struct WeightLDS
{
    static constexpr int MT = H / 32;
    static constexpr int S  = steps_per_tile<P>();
    typename P::Frag1 w1[MT][P::kSteps1][64];
    typename P::Frag w2[MT][MT][S][64];
    typename P::Frag w2t[MT][MT][S][64];
    typename P::Frag w1t[MT][S][64];
    float vec[4][64];
};

// Packed parameter layout (C -> H -> H -> 1): W0 [H][C], b0 [H], W1 [H][H], b1 [H], W2 [1][H], b2 [1].
template <typename P, int H, int C>
__device__ void load_weights(const float* __restrict__ params, WeightLDS<P, H>& w)
{
    constexpr int MT = H / 32, S = steps_per_tile<P>();
    const float* W0 = params;
    const float* b0 = W0 + H * C;
    const float* W1 = b0 + H;
    const float* b1 = W1 + H * H;
    const float* W2 = b1 + H;
    const float* b2 = W2 + H;
    for (int t = threadIdx.x; t < MT * 64; t += blockDim.x)
    {
        const int mt = t / 64, l = t % 64, g = l >> 5, r = 32 * mt + (l & 31);
        if constexpr (P::kBF16)
        {
            s16x4 f;
#pragma unroll
            for (int j = 0; j < 4; ++j) f[j] = bf16_bits(W0[r * C + 4 * g + j]);
            w.w1[mt][0][l] = f;
        }
        else
        {
#pragma unroll
            for (int s = 0; s < 4; ++s) w.w1[mt][s][l] = W0[r * C + 4 * g + s];
        }
    }
    for (int t = threadIdx.x; t < MT * MT * S * 64; t += blockDim.x)
    {
        const int l = t % 64, st = (t / 64) % S, ms = (t / (64 * S)) % MT, mt = t / (64 * S * MT);
        const int g = l >> 5, dst = 32 * mt + (l & 31);
        typename P::Frag a, b;
#pragma unroll
        for (int j = 0; j < P::KS; ++j)
        {
            const int src = 32 * ms + crow(st * P::KS + j, g);
            float va      = W1[dst * H + src];  // forward: W1[dst][src]
            float vb      = W1[src * H + dst];  // transposed: W1^T[dst][src] = W1[src][dst]
            if constexpr (P::kBF16)
            {
                a[j] = (__bf16)va;
                b[j] = (__bf16)vb;
            }
            else
            {
                a = va;
                b = vb;
            }
        }
        w.w2[mt][ms][st][l]  = a;
        w.w2t[mt][ms][st][l] = b;
    }
    for (int t = threadIdx.x; t < MT * S * 64; t += blockDim.x)
    {
        const int l = t % 64, st = (t / 64) % S, ms = t / (64 * S);
        const int g = l >> 5, c = l & 31;
        typename P::Frag f;
#pragma unroll
        for (int j = 0; j < P::KS; ++j)
        {
            const int src = 32 * ms + crow(st * P::KS + j, g);
            float v       = c < C ? W0[src * C + c] : 0.f;
            if constexpr (P::kBF16)
                f[j] = (__bf16)v;
            else
                f = v;
        }
        w.w1t[ms][st][l] = f;
    }
    for (int t = threadIdx.x; t < 64; t += blockDim.x)
    {
        w.vec[0][t] = t < H ? b0[t] : 0.f;
        w.vec[1][t] = t < H ? b1[t] : 0.f;
        w.vec[2][t] = t < H ? W2[t] : 0.f;
        w.vec[3][t] = t == 0 ? b2[0] : 0.f;
    }
}

// Per-tile forward state of a lane: pre-activations of both hidden layers (C layout) and the output pre-activation.
template <int MT>
struct Fwd
{
    f32x16 z1[MT];
    f32x16 z2[MT];
    float z;
};

// First layer input: the lane's 4 channels (group g holds channels 4g..4g+3 of its sample).
template <typename P>
__device__ __forceinline__ typename P::Frag1 x_step(const float (&x4)[4], int s)
{
    if constexpr (P::kBF16)
    {
        s16x4 f;
#pragma unroll
        for (int j = 0; j < 4; ++j) f[j] = bf16_bits(x4[j]);
        return f;
    }
    else
    {
        return x4[s];
    }
}

template <typename P, int H, int ACT>
__device__ __forceinline__ void tile_forward_mfma(const WeightLDS<P, H>& w, const float (&x4)[4], float beta,
                                                  Fwd<H / 32>& f)
{
    constexpr int MT = H / 32, S = steps_per_tile<P>();
    const int l = threadIdx.x & 63, g = l >> 5;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
    {
        f32x16 acc = {};
#pragma unroll
        for (int s = 0; s < P::kSteps1; ++s) acc = P::mma1(w.w1[mt][s][l], x_step<P>(x4, s), acc);
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[i] += w.vec[0][32 * mt + crow(i, g)];
        f.z1[mt] = acc;
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
    {
        f32x16 acc = {};
#pragma unroll
        for (int ms = 0; ms < MT; ++ms)
        {
#pragma unroll
            for (int st = 0; st < S; ++st)
            {
                auto b = P::from_acc(f.z1[ms], st, [&](float v, int) { return act_f(ACT, v, beta); });
                acc    = P::mma(w.w2[mt][ms][st][l], b, acc);
            }
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[i] += w.vec[1][32 * mt + crow(i, g)];
        f.z2[mt] = acc;
    }
    float part = 0.f;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
    {
#pragma unroll
        for (int i = 0; i < 16; ++i) part = fmaf(w.vec[2][32 * mt + crow(i, g)], act_f(ACT, f.z2[mt][i], beta), part);
    }
    f.z = part + __shfl_xor(part, 32) + w.vec[3][0];
}

// ---- weight-gradient products over a tile's samples -----------------------------------------------------------
__device__ __forceinline__ void wave_lds_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

constexpr int kRS = 36;  // row stride (floats) of the backward's [row][sample] staging buffers (16-byte rows)

// A / B fragment of a weight-gradient product over the tile's 32 samples, from a [row][sample] staging buffer:
// FP32 step st covers samples 2st + l/32; BF16 step st covers samples 16st + 8(l/32) + j.
template <typename P>
__device__ __forceinline__ typename P::Frag stage_frag(const float* buf, int row, int st, int hg, bool zero)
{
    typename P::Frag f;
    if constexpr (P::kBF16)
    {
        // masked lanes read row 0 (always inside the buffer) and discard it
        const float4* src = reinterpret_cast<const float4*>(buf + (zero ? 0 : row) * kRS + 16 * st + 8 * hg);
        const float4 u = src[0], v = src[1];
        const float e[8] = {u.x, u.y, u.z, u.w, v.x, v.y, v.z, v.w};
#pragma unroll
        for (int j = 0; j < 8; ++j) f[j] = (__bf16)(zero ? 0.f : e[j]);
    }
    else
    {
        f = zero ? 0.f : buf[row * kRS + 2 * st + hg];
    }
    return f;
}

// Rows of the X staging tile: the 8 features, then the columns that turn the X product into the bias and output
// gradients (E = dZ1 [x; 1]^T + dZ2 [e10]^T + h2 [dz e9]^T, with the B operand masked per product).
constexpr int kXOnes1 = 8, kXDz = 9, kXOnes2 = 10, kXRows = 11;

// The backward's [row][sample] staging buffers: BF16 (rows of 40) for the BF16 form, where the values are rounded
// once when staged and read as ready operands, FP32 (rows of kRS) otherwise.
template <typename P>
struct Stage
{
    static constexpr int kStride = P::kBF16 ? 40 : kRS;  // elements per row; 16-byte rows either way
    typedef typename std::conditional<P::kBF16, unsigned short, float>::type E;
    // elements of one wave's buffers for hidden width h: dZ / activations (2h rows) and the X tile
    static constexpr int elems(int h) { return (2 * h + kXRows) * kStride; }
    __device__ static __forceinline__ void put(E* b, int row, int col, float v)
    {
        if constexpr (P::kBF16)
            b[row * kStride + col] = (unsigned short)bf16_bits(v);
        else
            b[row * kStride + col] = v;
    }
    // A / B fragment over the tile's samples (see stage_frag); masked lanes read row 0 and discard it
    __device__ static __forceinline__ typename P::Frag frag(const E* b, int row, int st, int hg, bool zero)
    {
        if constexpr (P::kBF16)
        {
            const uint4 u = *reinterpret_cast<const uint4*>(b + (zero ? 0 : row) * kStride + 16 * st + 8 * hg);
            typename P::Frag f;
            __builtin_memcpy(&f, &u, sizeof(f));
            if (zero) f = typename P::Frag{};
            return f;
        }
        else
        {
            return zero ? 0.f : b[row * kStride + 2 * st + hg];
        }
    }
};

// One wave's staging buffers: sA (H rows), sB (H rows), sX (kXRows rows). The constant rows of sX are set here.
template <typename P, int H>
struct BwdStage
{
    typedef typename Stage<P>::E E;
    E* sA;
    E* sB;
    E* sX;
    __device__ explicit BwdStage(E* base) : sA(base), sB(base + H * Stage<P>::kStride), sX(sB + H * Stage<P>::kStride)
    {
        const int l = threadIdx.x & 63;
        if (l < 32)
        {
            Stage<P>::put(sX, kXOnes1, l, 1.f);
            Stage<P>::put(sX, kXOnes2, l, 1.f);
        }
    }
};

// A wave's running parameter-gradient products: dW2 (hidden layer), E (first layer in columns 0..7, its bias in
// column kXOnes1, the output-layer weights in column kXDz, the hidden bias in column kXOnes2), b3 (output bias).
template <int MT>
struct BwdAcc
{
    f32x16 dW2[MT][MT];
    f32x16 E[MT];
    float b3;
    __device__ BwdAcc() : b3(0.f)
    {
#pragma unroll
        for (int a = 0; a < MT; ++a)
        {
            E[a] = f32x16{};
#pragma unroll
            for (int b = 0; b < MT; ++b) dW2[a][b] = f32x16{};
        }
    }
};

// Backward of one tile of 32 samples on a wave: recompute the forward (tile_forward_mfma), take dz = gy *
// last_act'(z) (0 for invalid lanes), back-propagate on the matrix cores, and add the tile's parameter-gradient
// products to acc (products over the tile's samples, staged through LDS as [row][sample]). Returns dL/dx in
// registers 0..3 of each lane (channels 4g..4g+3 of its sample, g = lane / 32).
template <typename P, int H, int ACT>
__device__ __forceinline__ f32x16 tile_backward_mfma(const WeightLDS<P, H>& w, const float (&x4)[4], float gy,
                                                     bool valid, int last_act, float beta, const BwdStage<P, H>& sb,
                                                     BwdAcc<H / 32>& acc)
{
    constexpr int MT = H / 32, S = steps_per_tile<P>();
    constexpr int KSTEPS = 32 / (2 * P::KS);  // sample steps of a weight-gradient product (FP32 16, BF16 2)
    typedef Stage<P> ST;
    const int l = threadIdx.x & 63, n = l & 31, hg = l >> 5;
    Fwd<MT> f;
    tile_forward_mfma<P, H, ACT>(w, x4, beta, f);
    const float dz = valid ? gy * act_d(last_act, f.z, beta) : 0.f;
    if (hg == 0) acc.b3 += dz;
    // X tile and the first staged pair: dZ2 (A) and act(z1) (B)
#pragma unroll
    for (int j = 0; j < 4; ++j) ST::put(sb.sX, 4 * hg + j, n, x4[j]);
    if (hg == 0) ST::put(sb.sX, kXDz, n, dz);
    f32x16 dZ2[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
    {
#pragma unroll
        for (int i2 = 0; i2 < 16; ++i2)
        {
            const int r = 32 * mt + crow(i2, hg);
            dZ2[mt][i2] = w.vec[2][r] * dz * act_d(ACT, f.z2[mt][i2], beta);
            ST::put(sb.sA, r, n, dZ2[mt][i2]);
            ST::put(sb.sB, r, n, act_f(ACT, f.z1[mt][i2], beta));
        }
    }
    wave_lds_fence();
    // dW(hidden) += dZ2 act(z1)^T; E += dZ2 e10^T (the hidden bias gradient, column 10)
#pragma unroll
    for (int st = 0; st < KSTEPS; ++st)
    {
        const auto bx = ST::frag(sb.sX, n, st, hg, n != kXOnes2);
        typename P::Frag bf[MT];
#pragma unroll
        for (int m1 = 0; m1 < MT; ++m1) bf[m1] = ST::frag(sb.sB, 32 * m1 + n, st, hg, false);
#pragma unroll
        for (int m2 = 0; m2 < MT; ++m2)
        {
            const auto a = ST::frag(sb.sA, 32 * m2 + n, st, hg, false);
            acc.E[m2]    = P::mma(a, bx, acc.E[m2]);
#pragma unroll
            for (int m1 = 0; m1 < MT; ++m1) acc.dW2[m2][m1] = P::mma(a, bf[m1], acc.dW2[m2][m1]);
        }
    }
    wave_lds_fence();
    // E += act(z2) (dz e9)^T: the output-layer weight gradient, column 9
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
    {
#pragma unroll
        for (int i2 = 0; i2 < 16; ++i2) ST::put(sb.sB, 32 * mt + crow(i2, hg), n, act_f(ACT, f.z2[mt][i2], beta));
    }
    wave_lds_fence();
#pragma unroll
    for (int st = 0; st < KSTEPS; ++st)
    {
        const auto bx = ST::frag(sb.sX, n, st, hg, n != kXDz);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) acc.E[mt] = P::mma(ST::frag(sb.sB, 32 * mt + n, st, hg, false), bx, acc.E[mt]);
    }
    // dZ1 = (W1^T dZ2) * act'(z1), in registers
    f32x16 dZ1[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
    {
        f32x16 a = {};
#pragma unroll
        for (int ms = 0; ms < MT; ++ms)
        {
#pragma unroll
            for (int st = 0; st < S; ++st)
                a = P::mma(w.w2t[mt][ms][st][l], P::from_acc(dZ2[ms], st, [](float v, int) { return v; }), a);
        }
#pragma unroll
        for (int i2 = 0; i2 < 16; ++i2) a[i2] *= act_d(ACT, f.z1[mt][i2], beta);
        dZ1[mt] = a;
    }
    // dL/dx: rows 0..7 of W0^T dZ1, i.e. registers 0..3 of each lane group (channels 4g..4g+3)
    f32x16 dx = {};
#pragma unroll
    for (int ms = 0; ms < MT; ++ms)
    {
#pragma unroll
        for (int st = 0; st < S; ++st)
            dx = P::mma(w.w1t[ms][st][l], P::from_acc(dZ1[ms], st, [](float v, int) { return v; }), dx);
    }
    // E += dZ1 [x; 1]^T: the first-layer weight gradient (columns 0..7) and bias gradient (column 8)
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
    {
#pragma unroll
        for (int i2 = 0; i2 < 16; ++i2) ST::put(sb.sA, 32 * mt + crow(i2, hg), n, dZ1[mt][i2]);
    }
    wave_lds_fence();
#pragma unroll
    for (int st = 0; st < KSTEPS; ++st)
    {
        const auto bx = ST::frag(sb.sX, n, st, hg, n > kXOnes1);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) acc.E[mt] = P::mma(ST::frag(sb.sA, 32 * mt + n, st, hg, false), bx, acc.E[mt]);
    }
    wave_lds_fence();
    return dx;
}

// A wave's parameter-gradient products written to its slice `part` of the partials, in the packed layout
// (C -> H -> H -> 1: W0, b0, W1, b1, W2, b2). Every entry of the slice is written.
template <int H, int C>
__device__ __forceinline__ void write_param_partials(BwdAcc<H / 32>& acc, float* part)
{
    constexpr int MT = H / 32;
    const int l = threadIdx.x & 63, n = l & 31, hg = l >> 5;
    const int oW0 = 0, ob0 = H * C, oW1 = ob0 + H, ob1 = oW1 + H * H, oW2 = ob1 + H, ob2 = oW2 + H;
#pragma unroll
    for (int m2 = 0; m2 < MT; ++m2)
    {
#pragma unroll
        for (int m1 = 0; m1 < MT; ++m1)
        {
#pragma unroll
            for (int i2 = 0; i2 < 16; ++i2) part[oW1 + (32 * m2 + crow(i2, hg)) * H + 32 * m1 + n] = acc.dW2[m2][m1][i2];
        }
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
    {
#pragma unroll
        for (int i2 = 0; i2 < 16; ++i2)
        {
            const int r = 32 * mt + crow(i2, hg);
            if (n < C) part[oW0 + r * C + n] = acc.E[mt][i2];
            if (n == kXOnes1) part[ob0 + r] = acc.E[mt][i2];
            if (n == kXDz) part[oW2 + r] = acc.E[mt][i2];
            if (n == kXOnes2) part[ob1 + r] = acc.E[mt][i2];
        }
    }
    float b3 = acc.b3;
#pragma unroll
    for (int off = 16; off >= 1; off >>= 1) b3 += __shfl_xor(b3, off);
    if (l == 0) part[ob2] = b3;
}

// ---- BF16 relu decoders: weights for products in both orientations --------------------------------------------
// The A and B fragments of v_mfma_f32_32x32x16_bf16 share one register layout (lane l%32 = row of A / column of B,
// lane l/32 and slot j = k), so mma(a, b) and mma(b, a) give a product and its transpose from the same registers.
// Each weight fragment below therefore serves a product in either orientation. Biases ride in spare k slots as a
// (hi, lo) pair of BF16 values against constant 1 slots, so they cost no elementwise adds.
//   w1e[MT][64]          first layer: dst row 32mt + l%32; group 0 slots 0..3 = channels 0..3, 4 = bias hi,
//                        5 = bias lo; group 1 slots 0..3 = channels 4..7
//   w2[MT][MT][2][64]    hidden layer (as WeightLDS::w2)
//   w2b[MT][64]          hidden bias: group 0 slots 0, 1 = bias hi, lo
//   w2t[MT][MT][2][64]   transposed hidden layer with each source row scaled by its output-layer weight:
//                        dst = first-layer unit 32mt + l%32, src = hidden unit 32ms + crow(8st + j, l/32)
//   w1t[MT][2][64]       transposed first layer (as WeightLDS::w1t)
//   w3[64], b3           output layer (FP32)
// WeightLDSReluFwd holds what the forward reads (w1e, w2, w2b, w3, b3); WeightLDSRelu adds w2t and w1t.
template <int H>
struct WeightLDSReluFwd  // what the forward needs
{
    static constexpr int MT = H / 32;
    bf16x8 w1e[MT][64];
    bf16x8 w2[MT][MT][2][64];
    bf16x8 w2b[MT][64];
    float w3[64];
    float b3;
};

template <int H>
struct WeightLDSRelu : WeightLDSReluFwd<H>
{
    static constexpr int MT = H / 32;
    bf16x8 w2t[MT][MT][2][64];
    bf16x8 w1t[MT][2][64];
    // output layer as an A fragment: row 0 holds w3 in the K order of an h fragment of tile mt, step st; other rows 0
    bf16x8 w3f[MT][2][64];
};

// (hi, lo) BF16 split of an FP32 value into slots k, k + 1 of a fragment: hi + lo carries about 16 significant bits
__device__ __forceinline__ void bf16_split(float v, bf16x8& f, int k)
{
    const __bf16 hi = (__bf16)v;
    f[k]            = hi;
    f[k + 1]        = (__bf16)(v - (float)hi);
}

// Packed parameter layout as load_weights (C = 8).
template <int H, typename WS>
__device__ void load_weights_relu(const float* __restrict__ params, WS& w)
{
    constexpr bool kBwd = std::is_same<WS, WeightLDSRelu<H>>::value;
    constexpr int MT = H / 32, C = 8;
    const float* W0 = params;
    const float* b0 = W0 + H * C;
    const float* W1 = b0 + H;
    const float* b1 = W1 + H * H;
    const float* W2 = b1 + H;
    const float* b2 = W2 + H;
    for (int t = threadIdx.x; t < MT * 64; t += blockDim.x)
    {
        const int mt = t / 64, l = t % 64, g = l >> 5, r = 32 * mt + (l & 31);
        bf16x8 f = {}, fb = {};
#pragma unroll
        for (int j = 0; j < 4; ++j) f[j] = (__bf16)W0[r * C + 4 * g + j];
        if (g == 0)
        {
            bf16_split(b0[r], f, 4);
            bf16_split(b1[r], fb, 0);
        }
        w.w1e[mt][l] = f;
        w.w2b[mt][l] = fb;
    }
    for (int t = threadIdx.x; t < MT * MT * 2 * 64; t += blockDim.x)
    {
        const int l = t % 64, st = (t / 64) % 2, ms = (t / 128) % MT, mt = t / (128 * MT);
        const int g = l >> 5, dst = 32 * mt + (l & 31);
        bf16x8 a, b;
#pragma unroll
        for (int j = 0; j < 8; ++j)
        {
            const int src = 32 * ms + crow(st * 8 + j, g);
            a[j]          = (__bf16)W1[dst * H + src];
            b[j]          = (__bf16)(W1[src * H + dst] * W2[src]);
        }
        w.w2[mt][ms][st][l] = a;
        if constexpr (kBwd) w.w2t[mt][ms][st][l] = b;
    }
    if constexpr (kBwd)
    for (int t = threadIdx.x; t < MT * 2 * 64; t += blockDim.x)
    {
        const int l = t % 64, st = (t / 64) % 2, ms = t / 128;
        const int g = l >> 5, c = l & 31;
        bf16x8 f;
#pragma unroll
        for (int j = 0; j < 8; ++j) f[j] = (__bf16)(c < C ? W0[(32 * ms + crow(st * 8 + j, g)) * C + c] : 0.f);
        w.w1t[ms][st][l] = f;
    }
    if constexpr (kBwd)
    for (int t = threadIdx.x; t < MT * 2 * 64; t += blockDim.x)
    {
        const int l = t % 64, st = (t / 64) % 2, mt = t / 128;
        const int g = l >> 5, m = l & 31;
        bf16x8 f;
#pragma unroll
        for (int j = 0; j < 8; ++j) f[j] = (__bf16)(m == 0 ? W2[32 * mt + crow(st * 8 + j, g)] : 0.f);
        w.w3f[mt][st][l] = f;
    }
    for (int t = threadIdx.x; t < 64; t += blockDim.x) w.w3[t] = t < H ? W2[t] : 0.f;
    if (threadIdx.x == 0) w.b3 = b2[0];
}


}  // namespace neat
