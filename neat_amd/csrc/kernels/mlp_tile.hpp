// Tile-level FP32 decoder MLP used by the unfused decoder kernels and the fused projector.
//
// A tile is 64 samples processed by one wave, thread s = sample s. Layer activations live in LDS as feature-major
// columns (buf[k * kS + s], padded stride kS = 65 so per-thread accesses are bank-conflict free); a layer loads its
// thread's input column into registers once, then loops over output blocks of 4 with the weights read through
// uniform addresses (scalar loads). The weight gradient, a reduction over the tile's samples, is accumulated by
// letting thread o own output row o (hidden layers) or input column k (output layer).
//
// Packed parameter layout (FCBlockImpl, ImplicitNet.h:18-71): for each Linear layer l, W_l [dout, din] row-major,
// then b_l [dout]. Weight (o, k) of a layer is read as w[k * ks + o * os]: (ks, os) = (1, din) for the packed layout,
// (dout, 1) for a transposed copy Wt_l [din, dout] at the same offsets (contiguous in o, so wider scalar loads).
#pragma once

#include "common.hpp"

namespace neat
{

constexpr int kTile   = 64;  // samples per tile (one wave)
constexpr int kS      = 65;  // LDS column stride in floats
constexpr int kMaxIn  = 16;
constexpr int kMaxOut = 4;

struct MLP
{
    int in, hidden, layers, out;
    __host__ __device__ int nlin() const { return layers + 2; }
    __host__ __device__ int din(int l) const { return l == 0 ? in : hidden; }
    __host__ __device__ int dout(int l) const { return l == nlin() - 1 ? out : hidden; }
    __host__ __device__ int wofs(int l) const
    {
        int o = 0;
        for (int k = 0; k < l; ++k) o += dout(k) * din(k) + dout(k);
        return o;
    }
    __host__ __device__ int bofs(int l) const { return wofs(l) + dout(l) * din(l); }
    __host__ __device__ int nparams() const { return wofs(nlin()); }
};

struct WLayout
{
    bool transposed;
    __device__ __forceinline__ int ks(int din, int dout) const { return transposed ? dout : 1; }
    __device__ __forceinline__ int os(int din, int dout) const { return transposed ? 1 : din; }
};

// dense layer for the thread's sample: out[o] = b[o] + sum_k Wt[k][o] * f(in[k]), f = act or identity.
template <int DIN>
__device__ __forceinline__ void dense_col(const float* __restrict__ wt, const float* __restrict__ b, int din, int dout,
                                          WLayout wl, const float* in_col, bool apply_act, int act, float beta,
                                          float* out_col)
{
    const int s  = threadIdx.x & (kTile - 1);
    const int ks = wl.ks(din, dout), os = wl.os(din, dout);
    float h[DIN];
#pragma unroll
    for (int k = 0; k < DIN; ++k)
    {
        float v = k < din ? in_col[k * kS + s] : 0.f;
        h[k]    = apply_act ? act_f(act, v, beta) : v;
    }
    for (int o0 = 0; o0 < dout; o0 += 4)
    {
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < DIN; ++k)
        {
            if (k < din)
            {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (o0 + j < dout) acc[j] = fmaf(wt[k * ks + (o0 + j) * os], h[k], acc[j]);
            }
        }
        // the bias is added after the sum, like addmm with beta = 1
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (o0 + j < dout) out_col[(o0 + j) * kS + s] = acc[j] + b[o0 + j];
    }
}

// Back-propagation through a dense layer for the thread's sample:
// dprev[k] = (sum_o W[o][k] * d[o]) * f'(zprev[k]) (f' = 1 when zprev is the raw input), written over zprev_col.
template <int DOUT>
__device__ __forceinline__ void back_col(const float* __restrict__ wt, int din, int dout, WLayout wl,
                                         const float* d_col, float* zprev_col, bool apply_act, int act, float beta)
{
    const int s  = threadIdx.x & (kTile - 1);
    const int ks = wl.ks(din, dout), os = wl.os(din, dout);
    float d[DOUT];
#pragma unroll
    for (int o = 0; o < DOUT; ++o) d[o] = o < dout ? d_col[o * kS + s] : 0.f;
    for (int k0 = 0; k0 < din; k0 += 4)
    {
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int o = 0; o < DOUT; ++o)
        {
            if (o < dout)
            {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (k0 + j < din) acc[j] = fmaf(wt[(k0 + j) * ks + o * os], d[o], acc[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            if (k0 + j < din)
            {
                float* p = zprev_col + (k0 + j) * kS + s;
                *p       = apply_act ? acc[j] * act_d(act, *p, beta) : acc[j];
            }
        }
    }
}

// gW[o][k] += sum_s d[o][s] * f(a[k][s]) and gb[o] += sum_s d[o][s] over the first ns samples; thread o owns row o.
__device__ __forceinline__ void grad_rows(const float* d_buf, const float* a_buf, int ns, int dout, int din,
                                          bool apply_act, int act, float beta, float* gW, float* gb)
{
    constexpr int kChunk = 16;
    const int o          = threadIdx.x & (kTile - 1);
    if (o >= dout) return;
    float accb = 0.f;
    for (int s = 0; s < ns; ++s) accb += d_buf[o * kS + s];
    gb[o] += accb;
    for (int k0 = 0; k0 < din; k0 += kChunk)
    {
        float acc[kChunk];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) acc[j] = 0.f;
        for (int s = 0; s < ns; ++s)
        {
            float dv = d_buf[o * kS + s];
#pragma unroll
            for (int j = 0; j < kChunk; ++j)
            {
                if (k0 + j < din)
                {
                    float a = a_buf[(k0 + j) * kS + s];
                    acc[j]  = fmaf(dv, apply_act ? act_f(act, a, beta) : a, acc[j]);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kChunk; ++j)
            if (k0 + j < din) gW[o * din + k0 + j] += acc[j];
    }
}

// Output layer (few rows): thread k owns column k.
__device__ __forceinline__ void grad_cols(const float* d_buf, const float* a_buf, int ns, int dout, int din,
                                          int act, float beta, float* gW, float* gb)
{
    const int k = threadIdx.x & (kTile - 1);
    if (k < din)
    {
        float acc[kMaxOut] = {0.f, 0.f, 0.f, 0.f};
        for (int s = 0; s < ns; ++s)
        {
            float a = act_f(act, a_buf[k * kS + s], beta);
#pragma unroll
            for (int o = 0; o < kMaxOut; ++o)
                if (o < dout) acc[o] = fmaf(d_buf[o * kS + s], a, acc[o]);
        }
#pragma unroll
        for (int o = 0; o < kMaxOut; ++o)
            if (o < dout) gW[o * din + k] += acc[o];
    }
    if (k < dout)
    {
        float accb = 0.f;
        for (int s = 0; s < ns; ++s) accb += d_buf[k * kS + s];
        gb[k] += accb;
    }
}

// LDS footprint of one tile: X [kMaxIn], A [H], B [H], D [kMaxOut] columns.
struct TileBufs
{
    float* X;
    float* A;
    float* B;
    float* D;
};
constexpr int kTileFloats = (kMaxIn + 64 + 64 + kMaxOut) * kS;

__device__ __forceinline__ TileBufs tile_bufs(float* base)
{
    return TileBufs{base, base + kMaxIn * kS, base + (kMaxIn + 64) * kS, base + (kMaxIn + 128) * kS};
}

// Forward of the tile (X must hold the inputs). Leaves z0 in A, z1 (layers == 1) in B, and returns the output
// pre-activations in D. Requires H <= 64, layers <= 1.
template <int H>
__device__ __forceinline__ void tile_forward(const MLP& m, const float* __restrict__ wt, WLayout wl, int act,
                                             float beta, const TileBufs& t)
{
    dense_col<kMaxIn>(wt + m.wofs(0), wt + m.bofs(0), m.in, H, wl, t.X, false, act, beta, t.A);
    const float* last_in = t.A;
    if (m.layers == 1)
    {
        dense_col<H>(wt + m.wofs(1), wt + m.bofs(1), H, H, wl, t.A, true, act, beta, t.B);
        last_in = t.B;
    }
    const int L = m.nlin() - 1;
    dense_col<H>(wt + m.wofs(L), wt + m.bofs(L), H, m.out, wl, last_in, true, act, beta, t.D);
}

// Backward of the tile after tile_forward. D must hold dL/dz_out (pre-activation gradient of the output layer).
// Accumulates the parameter gradient of the first ns samples into gacc (packed layout) and leaves dL/dx in X.
template <int H>
__device__ __forceinline__ void tile_backward(const MLP& m, const float* __restrict__ wt, WLayout wl, int act,
                                              float beta, const TileBufs& t, int ns, float* gacc)
{
    const int L          = m.nlin() - 1;
    float* last_in       = m.layers == 1 ? t.B : t.A;
    // output layer gradient (one wave: LDS writes of this tile are visible after the barrier)
    __syncthreads();
    grad_cols(t.D, last_in, ns, m.out, H, act, beta, gacc + m.wofs(L), gacc + m.bofs(L));
    __syncthreads();
    back_col<kMaxOut>(wt + m.wofs(L), H, m.out, wl, t.D, last_in, true, act, beta);  // last_in := delta
    __syncthreads();
    if (m.layers == 1)
    {
        grad_rows(t.B, t.A, ns, H, H, true, act, beta, gacc + m.wofs(1), gacc + m.bofs(1));
        __syncthreads();
        back_col<H>(wt + m.wofs(1), H, H, wl, t.B, t.A, true, act, beta);  // A := delta of layer 0
        __syncthreads();
    }
    grad_rows(t.A, t.X, ns, H, m.in, false, act, beta, gacc + m.wofs(0), gacc + m.bofs(0));
    __syncthreads();
    back_col<H>(wt + m.wofs(0), m.in, H, wl, t.A, t.X, false, act, beta);  // X := dL/dx
    __syncthreads();
}

// Transposed copy of the packed parameters: Wt_l [din, dout] at W_l's offset, biases copied.
static __global__ void __launch_bounds__(256) k_transpose_params(const float* params, MLP m, float* wt)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m.nparams()) return;
    int l = 0;
    while (l + 1 < m.nlin() && p >= m.wofs(l + 1)) ++l;
    int rel = p - m.wofs(l);
    int din = m.din(l), dout = m.dout(l);
    if (rel < din * dout)
    {
        int o = rel / din, k = rel % din;       // params: W[o][k]
        wt[m.wofs(l) + k * dout + o] = params[p];  // wt: Wt[k][o]
    }
    else
    {
        wt[p] = params[p];  // bias
    }
}

// Fixed-order sum of num_wg partial slices [num_wg][P]: a block covers 32 parameters with 8 groups of 32 threads;
// group q sums slices q, q + 8, q + 16, ... in order (loads of one thread are independent, so they overlap), then
// the 8 group sums are added in group order. Deterministic for a given num_wg.
static __global__ void __launch_bounds__(256) k_dec_reduce_partials(const float* partials, int num_wg, int P, float* out)
{
    __shared__ float s[8][32];
    const int lane = threadIdx.x & 31, q = threadIdx.x >> 5;
    const int p    = blockIdx.x * 32 + lane;
    float acc      = 0.f;
    if (p < P)
    {
#pragma unroll 8
        for (int w = q; w < num_wg; w += 8) acc += partials[(int64_t)w * P + p];
    }
    s[q][lane] = acc;
    __syncthreads();
    if (q == 0 && p < P)
    {
        float t = s[0][lane];
#pragma unroll
        for (int k = 1; k < 8; ++k) t += s[k][lane];
        out[p] = t;
    }
}

// grid of k_dec_reduce_partials for P parameters (256 threads per block)
static inline unsigned reduce_partials_blocks(int P) { return (unsigned)((P + 31) / 32); }

}  // namespace neat
