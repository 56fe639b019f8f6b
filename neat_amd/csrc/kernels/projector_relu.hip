// Backward of the fused projector for BF16 decoders with relu hidden layers, in a file of its own: it is compiled
// with the matrix-core results in VGPRs (-mllvm -amdgpu-mfma-vgpr-form), since nearly every product here is
// consumed by vector ALU code, which cannot read accumulation registers directly.
#include "projector_dev.hpp"

namespace neat
{

// Backward for BF16 relu decoders (the configuration the projector benchmarks). Per wave and tile of 32 samples,
// with "h" fragments holding hidden units in the registers and samples on the lanes, and "s" fragments the transpose
// (what the weight-gradient products, which contract over samples, take):
//   forward  Z1h = W0 [x;1]; H1 = relu(Z1) (BF16); Z2h = W1 H1 + b1; z = w2 . relu(Z2h) + b2
//   backward dz = dL/dz; D2 = dz * [Z2 > 0] (the output weights are folded into w2t and the final row scale)
//            dW1 += D2s H1s^T; E += D2s [e10] + relu(Z2s) [dz e9] + D1s [x; 1]  (bias / output / first-layer grads)
//            D1h = (w2t D2h) * [H1 > 0]; dL/dfeature = W0^T D1h
// Every product runs once, in h form; the s fragments of H1, relu(Z2) and D1 come from LDS transposes
// (transpose_tile), and the s tile T of [x; 1; dz; 1] from one product against a 0/1 routing fragment. The relu
// masks come from the BF16 activations themselves, so only BF16 fragments stay live across the tile.
typedef short s16x8 __attribute__((ext_vector_type(8)));

// Elementwise work on BF16 fragments runs as packed 16-bit integer operations on their bit patterns (two lanes of
// a register per instruction; integer ops, not packed FP32): a BF16 value is negative exactly when its int16
// pattern is, and after relu a nonzero pattern means an active unit.

// BF16 fragment of z slots 8st..8st+7, rounded to nearest even, one conversion per pair (converting a 2-vector;
// converting single values makes the compiler convert each alone and repack the halves)
__device__ __forceinline__ bf16x8 bf16_frag(const f32x16& z, int st)
{
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    u32x4 r;
#pragma unroll
    for (int q = 0; q < 4; ++q)
    {
        const f32x2 v = {z[8 * st + 2 * q], z[8 * st + 2 * q + 1]};
        r[q]          = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
    }
    return __builtin_bit_cast(bf16x8, r);
}

// relu on BF16 patterns: max(int16, 0) maps every negative value, -0 included, to +0
__device__ __forceinline__ bf16x8 relu_bf16(const bf16x8& f)
{
    return __builtin_bit_cast(bf16x8, __builtin_elementwise_max(__builtin_bit_cast(s16x8, f), (s16x8)0));
}

// f where the relu activation h is nonzero, else 0: f * min(h, 1) per 16-bit lane, two lanes per instruction. The
// constant 1 is hidden from the compiler (an SGPR set by an s_mov it cannot see through); knowing that min(h, 1) is
// 0 or 1 it would rewrite the product as per-lane compares and selects.
__device__ __forceinline__ uint32_t opaque_one2()
{
    uint32_t v;
    asm volatile("s_mov_b32 %0, 0x10001" : "=s"(v));
    return v;
}

typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ bf16x8 keep_active(const bf16x8& f, const bf16x8& h, uint32_t one2)
{
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 ones = {one2, one2, one2, one2};
    const u16x8 m    = __builtin_elementwise_min(__builtin_bit_cast(u16x8, h), __builtin_bit_cast(u16x8, ones));
    return __builtin_bit_cast(bf16x8, __builtin_bit_cast(u16x8, f) * m);
}

// relu for the FP32 output-layer dot product (one instruction; -0 may pass through, which adds nothing)
__device__ __forceinline__ float relu_f32(float z) { return __builtin_amdgcn_fmed3f(z, 0.f, 3.402823466e38f); }

// The softplus output and its derivative (the reference's threshold 20 on beta z, DecodeFeatures,
// geometry.cpp:123-143) on the hardware's exp2 / log2 / reciprocal: the BF16 form is held to its error bar against
// the FP64 oracle, not to the FP32 oracle's rounding, and the accurate log1p costs about 70 instructions per sample.
// log(1 + e) loses e's low bits once e < 2^-24 (beta z below about -17), where the value is below 4e-8 / beta.
__device__ __forceinline__ float softplus_fast(float z, float beta, float inv_beta)
{
    const float t = z * beta;
    return t > 20.f ? z : __logf(1.f + __expf(t)) * inv_beta;
}

__device__ __forceinline__ float softplus_d_fast(float z, float beta)
{
    const float t = z * beta;
    if (t > 20.f) return 1.f;
    const float e = __expf(t);
    return e * __builtin_amdgcn_rcpf(e + 1.f);
}

// ---- fragment transposes through LDS ------------------------------------------------------------------------------
// An "h" fragment (samples on the lanes: lane (n, g) holds rows 16st + 8(j/4) + 4g + j%4 of sample n for slot j) and
// the "s" fragment of the same 32 x 32 tile (lane (m, g) holds samples crow(8st + j, g) of row m) are transposes of
// each other. The h fragments of a tile go to a per-wave LDS tile [sample][row] (rows padded to 36 elements, 72 B,
// so consecutive samples start on different banks) with two 8-byte stores each; the s fragments come back with
// ds_read_b64_tr_b16, which hands lane 16G + i of a 16-lane group element e of the 8 bytes that lane 16G + 4e + i/4
// points at, shifted by i % 4 elements: with lane 16G + 4e + q pointing at (sample 16st + 8r + 4(G/2) + e,
// row 16(G%2) + 4q), lane (m, g) receives rows m of samples 16st + 8r + 4g + e, the s-fragment slots 4r + e.
constexpr int kTrRow = 36;  // bf16 elements per sample row of a transpose tile

typedef short s16x4_t __attribute__((ext_vector_type(4)));
#define NEAT_LDS __attribute__((address_space(3)))

__device__ __forceinline__ void tr_store(unsigned short* tile, const bf16x8& f, int st)
{
    const int l = threadIdx.x & 63, n = l & 31, g = l >> 5;
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 v = __builtin_bit_cast(u32x4, f);
    unsigned short* row = tile + n * kTrRow + 16 * st + 4 * g;
    *reinterpret_cast<uint2*>(row)     = make_uint2(v[0], v[1]);
    *reinterpret_cast<uint2*>(row + 8) = make_uint2(v[2], v[3]);
}

__device__ __forceinline__ bf16x8 tr_load(const unsigned short* tile, int st)
{
    const int l = threadIdx.x & 63, G = l >> 4, e = (l >> 2) & 3, q = l & 3;
    typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
    u32x2 r[2];
#pragma unroll
    for (int h = 0; h < 2; ++h)
    {
        const unsigned short* src = tile + (16 * st + 8 * h + 4 * (G >> 1) + e) * kTrRow + 16 * (G & 1) + 4 * q;
        const s16x4_t v = __builtin_amdgcn_ds_read_tr16_b64_v4i16((NEAT_LDS s16x4_t*)(src));
        r[h]            = __builtin_bit_cast(u32x2, v);
    }
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 o = {r[0][0], r[0][1], r[1][0], r[1][1]};
    return __builtin_bit_cast(bf16x8, o);
}

// s fragments (st = 0, 1) of the tile whose h fragments are hf, through this wave's transpose tile
__device__ __forceinline__ void transpose_tile(unsigned short* tile, const bf16x8 (&hf)[2], bf16x8 (&sf)[2])
{
    wave_lds_fence();  // the previous reads of the tile are issued before these stores
    tr_store(tile, hf[0], 0);
    tr_store(tile, hf[1], 1);
    wave_lds_fence();
    sf[0] = tr_load(tile, 0);
    sf[1] = tr_load(tile, 1);
}

// Node-resident like the forward: workgroup b walks forward chunks [wg_first[b], wg_first[b + 1]) (two waves per
// SIMD, one workgroup per CU). Each sample's features and weight come from what the forward left in the plan's
// caches (feat_cache, w_cache: the very bits it computed), so the backward stages no node grid and recomputes no
// sample geometry; the packed cell coordinates the scatter reads are the forward's as well.
template <int H>
__global__ void __launch_bounds__(kNodeThreads, 2) k_mfma_bwd_relu(ProjGeom g, Plan p, const float* __restrict__ params,
                                                                   const float* grad_proj, float* dfeat,
                                                                   unsigned int* dmax, float* partials,
                                                                   int32_t P_total)
{
    constexpr int MT = H / 32, C = 8;
    typedef PrecBF16 P;
    __shared__ WeightLDSRelu<H> w;
    __shared__ ChunkCtx cx;
    __shared__ __align__(16) float s_dz[kNodeWaves][32];
    __shared__ __align__(16) unsigned short s_tr[kNodeWaves][32 * kTrRow];
    load_weights_relu<H>(params, w);
    __syncthreads();  // the final partial write reads w.w3 even in a workgroup that gets no chunk
    const int l = threadIdx.x & 63, wave = threadIdx.x >> 6, n = l & 31, hg = l >> 5;
    // constant fragments: ONES (group 0 slots 0, 1) against the hidden bias; IDB routes the slots of the feature
    // fragment to columns of the s-tile T (channels 0..7, 1 -> 8, 1 -> 10, dz -> 9)
    bf16x8 ones = {}, idb = {};
    if (hg == 0)
    {
        ones[0] = (__bf16)1.f;
        ones[1] = (__bf16)1.f;
    }
    {
        const int col[2][8] = {{0, 1, 2, 3, 8, 10, 9, -1}, {4, 5, 6, 7, -1, -1, -1, -1}};
#pragma unroll
        for (int j = 0; j < 8; ++j) idb[j] = (__bf16)(col[hg][j] == n ? 1.f : 0.f);
    }
    f32x16 dW2[MT][MT], E[MT];
#pragma unroll
    for (int a = 0; a < MT; ++a)
    {
        E[a] = f32x16{};
#pragma unroll
        for (int b = 0; b < MT; ++b) dW2[a][b] = f32x16{};
    }
    float b3 = 0.f;
    float amax = 0.f;  // largest |dL/dfeature| this lane wrote (the batch's goes to dmax[0], each slot's to 1 + slot)
    float* sdz          = s_dz[wave];
    unsigned short* trt = s_tr[wave];
    const uint32_t one2 = opaque_one2();
    const uint2* fc     = reinterpret_cast<const uint2*>(p.feat_cache);  // [S][2]: a lane group's 4 channels each
    const int32_t cend  = p.wg_first[blockIdx.x + 1];
    for (int32_t c = p.wg_first[blockIdx.x]; c < cend; ++c)
    {
        const int32_t slot = p.chunk1_slot[c];
        const int32_t q0 = p.chunk1_q0[c], q1 = p.chunk1_q1[c], nseg = q1 - q0;
        if (nseg == 0) continue;
        const int64_t base = p.nm_soff[q0];
        __syncthreads();  // the previous chunk's readers of cx are done
        load_chunk_ctx(g, p, q0, q1, base, grad_proj, cx);
        __syncthreads();
        const int32_t nsamp = cx.beg[nseg];
        float cmax = 0.f;  // this chunk's
        for (int32_t w0 = 0; w0 < nsamp; w0 += kFwdChunk)
        {
            const int32_t wn = min(kFwdChunk, nsamp - w0);
            expand_window(cx, nseg, w0, wn);
            __syncthreads();
            auto tile = [&](int32_t i, const float (&x4)[4], float gy) {
                const bool valid = i < wn;
                // feature fragment: channels 4hg..4hg+3, group 0 adds the two bias slots (1, 1)
                bf16x8 X = {};
#pragma unroll
                for (int j = 0; j < 4; ++j) X[j] = (__bf16)x4[j];
                if (hg == 0)
                {
                    X[4] = (__bf16)1.f;
                    X[5] = (__bf16)1.f;
                }
                // first layer (h)
                bf16x8 H1h[MT][2];
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                {
                    const f32x16 zh = P::mma(w.w1e[mt][l], X, f32x16{});
#pragma unroll
                    for (int st = 0; st < 2; ++st) H1h[mt][st] = relu_bf16(bf16_frag(zh, st));
                }
                // hidden layer (h): relu(Z2), the output pre-activation and dz, D2 = dz [Z2 > 0]. Here the output
                // layer runs on the matrix cores too (row 0 of w3f against relu(Z2) in BF16): z only enters dz, where
                // its BF16 rounding is of the size of the BF16 rounding of dz itself; the forward keeps it in FP32.
                bf16x8 H2h[MT][2], D2h[MT][2];
                float dz;
                {
                    f32x16 zacc = {};
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
                    {
                        f32x16 acc = P::mma(w.w2b[mt][l], ones, f32x16{});
#pragma unroll
                        for (int ms = 0; ms < MT; ++ms)
                        {
#pragma unroll
                            for (int st = 0; st < 2; ++st) acc = P::mma(w.w2[mt][ms][st][l], H1h[ms][st], acc);
                        }
#pragma unroll
                        for (int st = 0; st < 2; ++st)
                        {
                            H2h[mt][st] = relu_bf16(bf16_frag(acc, st));
                            zacc        = P::mma(w.w3f[mt][st][l], H2h[mt][st], zacc);
                        }
                    }
                    // row 0 of the product sits in register 0 of the lane group 0; group 1 takes it from there
                    // (the shuffle runs on every lane: a lane reading an inactive lane's register gets no data)
                    const float z0 = zacc[0];
                    const float zx = __shfl_xor(z0, 32);
                    const float z  = (hg == 0 ? z0 : zx) + w.b3;
                    dz = valid ? gy * softplus_d_fast(z, g.beta) : 0.f;
                    bf16x8 dz8;
#pragma unroll
                    for (int j = 0; j < 8; ++j) dz8[j] = (__bf16)dz;
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
                    {
#pragma unroll
                        for (int st = 0; st < 2; ++st) D2h[mt][st] = keep_active(dz8, H2h[mt][st], one2);
                    }
                }
                if (hg == 0) b3 += dz;
                // dz per register of an s tile (sample crow(i, hg)), through LDS
                wave_lds_fence();
                if (hg == 0) sdz[n] = dz;
                wave_lds_fence();
                bf16x8 dzs[2];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                {
                    const float4 v = *reinterpret_cast<const float4*>(&sdz[8 * q + 4 * hg]);
                    dzs[q >> 1][4 * (q & 1) + 0] = (__bf16)v.x;
                    dzs[q >> 1][4 * (q & 1) + 1] = (__bf16)v.y;
                    dzs[q >> 1][4 * (q & 1) + 2] = (__bf16)v.z;
                    dzs[q >> 1][4 * (q & 1) + 3] = (__bf16)v.w;
                }
                // the s tile T of [x; 1; dz; 1] (columns 0..7, 8, 9, 10) and its masked B fragments
                bf16x8 B08[2], B9[2], B10[2];
                {
                    if (hg == 0) X[6] = (__bf16)dz;
                    const f32x16 T = P::mma(X, idb, f32x16{});
#pragma unroll
                    for (int st = 0; st < 2; ++st)
                    {
                        const bf16x8 t = bf16_frag(T, st);
                        B08[st] = n <= 8 ? t : bf16x8{};
                        B9[st]  = n == 9 ? t : bf16x8{};
                        B10[st] = n == 10 ? t : bf16x8{};
                    }
                }
                // s fragments of H1 (hidden-weight gradient) by transposes
                bf16x8 H1s[MT][2];
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) transpose_tile(trt, H1h[mt], H1s[mt]);
                // relu(Z2) and D2 in s form; output-weight and hidden-bias gradients; hidden-weight gradient
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                {
                    bf16x8 H2s[2];
                    transpose_tile(trt, H2h[mt], H2s);
#pragma unroll
                    for (int st = 0; st < 2; ++st)
                    {
                        const bf16x8 d2 = keep_active(dzs[st], H2s[st], one2);
                        E[mt] = P::mma(H2s[st], B9[st], E[mt]);
                        E[mt] = P::mma(d2, B10[st], E[mt]);
#pragma unroll
                        for (int m1 = 0; m1 < MT; ++m1) dW2[mt][m1] = P::mma(d2, H1s[m1][st], dW2[mt][m1]);
                    }
                }
                // first layer: D1 = (w2t D2) [H1 > 0] in h form, dL/dfeature = W0^T D1, and in s form the first-layer
                // weight / bias gradients
                f32x16 dx = {};
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                {
                    f32x16 acc = {};
#pragma unroll
                    for (int ms = 0; ms < MT; ++ms)
                    {
#pragma unroll
                        for (int st = 0; st < 2; ++st) acc = P::mma(w.w2t[mt][ms][st][l], D2h[ms][st], acc);
                    }
                    bf16x8 D1h[2], D1s[2];
#pragma unroll
                    for (int st = 0; st < 2; ++st)
                    {
                        D1h[st] = keep_active(bf16_frag(acc, st), H1h[mt][st], one2);
                        dx      = P::mma(w.w1t[mt][st][l], D1h[st], dx);
                    }
                    transpose_tile(trt, D1h, D1s);
#pragma unroll
                    for (int st = 0; st < 2; ++st) E[mt] = P::mma(D1s[st], B08[st], E[mt]);
                }
                if (valid)
                {
                    // dL/dfeature [S][8] in BF16: this lane's channels 4hg..4hg+3
                    *reinterpret_cast<uint2*>(reinterpret_cast<unsigned short*>(dfeat) + (base + w0 + i) * 8 +
                                              4 * hg) = bf16x4_bits(dx[0], dx[1], dx[2], dx[3]);
                    cmax = fmaxf(cmax, fmaxf(fmaxf(fabsf(dx[0]), fabsf(dx[1])), fmaxf(fabsf(dx[2]), fabsf(dx[3]))));
                }
            };
            for (int32_t p0 = wave * 64; p0 < wn; p0 += kNodeThreads)
            {
                // both tiles' code inline (256 VGPRs, no spills), so that the scheduler can overlap them
#pragma unroll
                for (int h = 0; h < 2; ++h)
                {
                    if (h == 1 && p0 + 32 >= wn) break;
                    // this lane's 4 channels of sample i and dL/dy = dL/dproj x weight (zeros past the window)
                    const int32_t i = p0 + 32 * h + n;
                    float xs[4]     = {0.f, 0.f, 0.f, 0.f};
                    float gy        = 0.f;
                    if (i < wn)
                    {
                        const uint2 raw = fc[(base + w0 + i) * 2 + hg];
                        xs[0]           = __uint_as_float(raw.x << 16);
                        xs[1]           = __uint_as_float(raw.x & 0xFFFF0000u);
                        xs[2]           = __uint_as_float(raw.y << 16);
                        xs[3]           = __uint_as_float(raw.y & 0xFFFF0000u);
                        gy              = cx.gproj[cx.sq[i]] * p.w_cache[base + w0 + i];
                    }
                    tile(i, xs, gy);
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) cmax = fmaxf(cmax, __shfl_xor(cmax, off));
        if (l == 0) atomicMax(dmax + 1 + slot, __float_as_uint(cmax));
        amax = fmaxf(amax, cmax);
    }
    // ---- this wave's slice of the partial parameter gradient (packed layout); rows scaled by the output weights
    // where the output weights were folded out of D2 ----
    float* part   = partials + ((int64_t)blockIdx.x * kNodeWaves + wave) * P_total;
    const int oW0 = 0, ob0 = H * C, oW1 = ob0 + H, ob1 = oW1 + H * H, oW2 = ob1 + H, ob2 = oW2 + H;
#pragma unroll
    for (int m2 = 0; m2 < MT; ++m2)
    {
#pragma unroll
        for (int i2 = 0; i2 < 16; ++i2)
        {
            const int r    = 32 * m2 + crow(i2, hg);
            const float s3 = w.w3[r];
#pragma unroll
            for (int m1 = 0; m1 < MT; ++m1) part[oW1 + r * H + 32 * m1 + n] = dW2[m2][m1][i2] * s3;
            if (n < C) part[oW0 + r * C + n] = E[m2][i2];
            if (n == 8) part[ob0 + r] = E[m2][i2];
            if (n == 9) part[oW2 + r] = E[m2][i2];
            if (n == 10) part[ob1 + r] = E[m2][i2] * s3;
        }
    }
#pragma unroll
    for (int off = 16; off >= 1; off >>= 1) b3 += __shfl_xor(b3, off);
    if (l == 0) part[ob2] = b3;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) amax = fmaxf(amax, __shfl_xor(amax, off));
    if (l == 0) atomicMax(dmax, __float_as_uint(amax));
}

// Forward for BF16 relu decoders: k_mfma_fwd's structure (persistent workgroups over the chunks, tiles of 32
// samples per wave, per-segment sums sequential in j) with the first- and hidden-layer biases in the products'
// spare k slots and the elementwise work on packed BF16 fragments.
// Node-resident like the backward.
template <int H>
__global__ void __launch_bounds__(kNodeThreads, 2) k_mfma_fwd_relu(ProjGeom g, Plan p, const float* __restrict__ params,
                                                                   const void* grid_lowp, float* seg_partial,
                                                                   float* sample_out)
{
    constexpr int MT = H / 32;
    typedef PrecBF16 P;
    __shared__ WeightLDSReluFwd<H> w;
    __shared__ ChunkCtx cx;
    __shared__ float s_contrib[kFwdChunk];
    __shared__ float s_seg[kSegCap];
    __shared__ uint4 s_grid[kStagedVerts];
    load_weights_relu<H>(params, w);
    const int l = threadIdx.x & 63, wave = threadIdx.x >> 6, n = l & 31, hg = l >> 5;
    const float inv_beta = 1.f / g.beta;
    bf16x8 ones = {};
    if (hg == 0)
    {
        ones[0] = (__bf16)1.f;
        ones[1] = (__bf16)1.f;
    }
    const int32_t nvert = g.G * g.G * g.G;
    int32_t staged      = -1;
    const int32_t cend  = p.wg_first[blockIdx.x + 1];
    for (int32_t c = p.wg_first[blockIdx.x]; c < cend; ++c)
    {
        const int32_t slot = p.chunk1_slot[c];
        const int32_t q0 = p.chunk1_q0[c], q1 = p.chunk1_q1[c], nseg = q1 - q0;
        if (nseg == 0) continue;
        const int64_t base = p.nm_soff[q0];
        __syncthreads();  // the previous chunk's readers of cx / s_seg / s_grid are done (and the weights are in place)
        if (slot != staged)
        {
            stage_node_grid(static_cast<const uint4*>(grid_lowp) + (int64_t)slot * nvert, s_grid, nvert);
            staged = slot;
        }
        load_chunk_ctx(g, p, q0, q1, base, nullptr, cx);
        for (int32_t t = threadIdx.x; t < nseg; t += blockDim.x) s_seg[t] = 0.f;
        NodeBox nb;
        node_bounds(g, g.seg_node[p.perm[q0]], nb);
        __syncthreads();
        const int32_t nsamp = cx.beg[nseg];
        for (int32_t w0 = 0; w0 < nsamp; w0 += kFwdChunk)
        {
            const int32_t wn = min(kFwdChunk, nsamp - w0);
            expand_window(cx, nseg, w0, wn);
            __syncthreads();
            auto tile = [&](int32_t i, const float (&x4)[4], float wgt) {
                bf16x8 X = {};
#pragma unroll
                for (int j = 0; j < 4; ++j) X[j] = (__bf16)x4[j];
                if (hg == 0)
                {
                    X[4] = (__bf16)1.f;
                    X[5] = (__bf16)1.f;
                }
                bf16x8 H1h[MT][2];
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                {
                    const f32x16 zh = P::mma(w.w1e[mt][l], X, f32x16{});
#pragma unroll
                    for (int st = 0; st < 2; ++st) H1h[mt][st] = relu_bf16(bf16_frag(zh, st));
                }
                float part = 0.f;
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                {
                    f32x16 acc = P::mma(w.w2b[mt][l], ones, f32x16{});
#pragma unroll
                    for (int ms = 0; ms < MT; ++ms)
                    {
#pragma unroll
                        for (int st = 0; st < 2; ++st) acc = P::mma(w.w2[mt][ms][st][l], H1h[ms][st], acc);
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                    {
                        const float4 w3 = *reinterpret_cast<const float4*>(&w.w3[32 * mt + 8 * q + 4 * hg]);
                        part = fmaf(w3.x, relu_f32(acc[4 * q + 0]), part);
                        part = fmaf(w3.y, relu_f32(acc[4 * q + 1]), part);
                        part = fmaf(w3.z, relu_f32(acc[4 * q + 2]), part);
                        part = fmaf(w3.w, relu_f32(acc[4 * q + 3]), part);
                    }
                }
                const float z = part + __shfl_xor(part, 32) + w.b3;
                if (i < wn && hg == 0)
                {
                    const float y = softplus_fast(z, g.beta, inv_beta);
                    s_contrib[i]  = y * wgt;
                    if (sample_out) sample_out[base + w0 + i] = y;
                }
            };
            for (int32_t p0 = wave * 64; p0 < wn; p0 += kNodeThreads)
            {
                // the fetch also leaves each sample's packed cell coordinate and weight for the backward
                PairIn cur;
                fetch_pair<true, false>(g, cx, p0, w0, wn, base, nb, s_grid, reinterpret_cast<float4*>(p.cell_cache),
                                        cur, p.w_cache);
                if (p.feat_cache)
                {
                    // and this lane's 4 channels of both samples, in BF16 as the first layer takes them
                    unsigned short* fc = reinterpret_cast<unsigned short*>(p.feat_cache) + (base + w0 + p0 + n) * 8 +
                                         4 * hg;
                    if (p0 + n < wn)
                        *reinterpret_cast<uint2*>(fc) = bf16x4_bits(cur.xa[0], cur.xa[1], cur.xa[2], cur.xa[3]);
                    if (p0 + 32 + n < wn)
                        *reinterpret_cast<uint2*>(fc + 32 * 8) = bf16x4_bits(cur.xb[0], cur.xb[1], cur.xb[2], cur.xb[3]);
                }
                tile(p0 + n, cur.xa, cur.va);
                if (p0 + 32 < wn) tile(p0 + 32 + n, cur.xb, cur.vb);
            }
            __syncthreads();
            for (int32_t t = threadIdx.x; t < nseg; t += blockDim.x)
            {
                const int32_t a = max(cx.beg[t], w0), b = min(cx.beg[t + 1], w0 + wn);
                if (a >= b) continue;
                float acc = 0.f;
                for (int32_t i = a; i < b; ++i) acc += s_contrib[i - w0];
                s_seg[t] += acc;
            }
            __syncthreads();
        }
        for (int32_t t = threadIdx.x; t < nseg; t += blockDim.x) seg_partial[cx.k[t]] = s_seg[t];
    }
}

int launch_mfma_fwd_relu(hipStream_t s, int H, const ProjGeom& g, const Plan& p, const float* params,
                         const void* grid_bf16, float* seg_partial, float* sample_out)
{
    if (H == 64)
        k_mfma_fwd_relu<64><<<p.nwg, kNodeThreads, 0, s>>>(g, p, params, grid_bf16, seg_partial, sample_out);
    else
        k_mfma_fwd_relu<32><<<p.nwg, kNodeThreads, 0, s>>>(g, p, params, grid_bf16, seg_partial, sample_out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mfma_bwd_relu(hipStream_t s, int H, const ProjGeom& g, const Plan& p, const float* params,
                         const float* grad_proj, float* dfeat, unsigned int* dmax, float* partials, int32_t P_total)
{
    if (!p.feat_cache || !p.w_cache) return -1;
    if (H == 64)
        k_mfma_bwd_relu<64><<<p.nwg, kNodeThreads, 0, s>>>(g, p, params, grad_proj, dfeat, dmax, partials, P_total);
    else
        k_mfma_bwd_relu<32><<<p.nwg, kNodeThreads, 0, s>>>(g, p, params, grad_proj, dfeat, dmax, partials, P_total);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace neat
