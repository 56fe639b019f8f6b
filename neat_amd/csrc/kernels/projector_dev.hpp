// Device-side pieces shared by the projector's kernels (projector.hip, projector_relu.hip): the argument structs,
// per-sample geometry, corner gathers, and the chunk context and pair fetch of the matrix-core kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "mfma_mlp.hpp"
#include "mlp_tile.hpp"

namespace neat
{

constexpr int kFwdChunk   = 4096;  // samples per forward / backward-pass-1 chunk (and per forward window)
constexpr int kSegCost    = 16;    // chunk cost of a segment, in samples
constexpr int kSegCap     = kFwdChunk / kSegCost;  // segments per forward chunk, at most
constexpr int kFwdThreads = 256;

// Node-resident workgroups (the BF16 relu kernels): a workgroup of kNodeThreads (one fits a CU) walks a contiguous,
// cost-balanced run of forward chunks. The forward keeps the current slot's BF16 node grid in LDS, so every corner
// gather reads LDS and each node grid is fetched from HBM about once per pass; the backward reads the per-sample
// features the forward left (Plan::feat_cache). There are kNodeWGPerCU runs per CU, so the hardware's dispatch of
// the next waiting workgroup to the first free CU evens out what the cost model misses.
constexpr int kNodeThreads  = 512;
constexpr int kNodeWaves    = kNodeThreads / 64;
constexpr int kNodeWGPerCU  = 4;
constexpr int kNodeWGMax    = 1024;  // workgroups a plan can describe
constexpr int kStagedVerts = 4913;  // vertices of the largest staged node grid (17^3), 16 B each (8 BF16 channels)

struct ProjGeom
{
    const float* pmin;
    const float* pmax;
    const int64_t* prefix;
    int32_t num_nodes;
    const float* ro;
    const float* rd;
    int64_t R;
    const int64_t* ray_seg_offset;
    int64_t K;
    const int32_t* seg_node;
    const int32_t* seg_ray;
    const int32_t* seg_count;
    const float* seg_tmin;
    const float* seg_tmax;
    int32_t jitter;
    uint64_t seed;
    const float* grid;  // channels-last [N, G, G, G, C]
    int64_t N;
    int32_t G;
    float beta;
    const float* seg_cdf;  // optional [K][8] proposal CDF (row used when cdf[7] > 0)
    int64_t S;             // samples of the segment list
};

struct Plan
{
    int32_t* perm;         // [K] node-major position -> canonical segment index
    int64_t* nm_soff;      // [K + 1] node-major sample offsets
    int32_t* slot_begin;   // [N + 1] node-major segment range of each slot
    int64_t* chunk_off1;   // [N + 1] first forward chunk of each slot
    int64_t* chunk_off2;   // [N + 1] first scatter chunk of each slot
    int32_t* chunk1_slot;  // [max1]
    int32_t* chunk1_q0;
    int32_t* chunk1_q1;
    int32_t* chunk2_slot;  // [max2]
    int32_t* chunk2_q0;
    int32_t* chunk2_q1;
    int64_t max1, max2;
    int32_t chunk2;  // samples per scatter chunk
    int32_t* wg_first;  // [nwg + 1] first forward chunk of each node-resident workgroup
    int32_t nwg;        // node-resident workgroups the plan balanced the chunks over
    // per-segment records in node-major order, written once by the plan so that a chunk's context is one coalesced
    // read (the canonical arrays and the rays are read in scattered order there, once per plan)
    float4* rec_ot;  // [K] (origin, tmin)
    float4* rec_dt;  // [K] (direction, tmax)
    int4* rec_i;     // [K] (canonical index, ray, rank in the ray, sample count)
    uint4* rec_h;    // [K] (ray jitter hash low, high, kind, step bits); kind 0 uniform, 1 proposal cells, 2 CDF
    float4* rec_c;   // [2K] kind 1: cell boundaries 1..7 (and 0); kind 2: the CDF row; null without proposal rows
    int32_t* redo_slots;  // [N] slots with no or several scatter chunks (the scatter's reduction), *nredo of them
    int32_t* nredo;
    // what the BF16 relu forward leaves for its backward, per sample in node-major order: the features as 8 BF16
    // channels (the decoder's first-layer input, bit for bit), the sample weight, and the packed cell coordinate
    // (pack_cell_coord; the scatter's input as well). Null where the workspace has no room (4 channels).
    uint4* feat_cache;
    float* w_cache;
    uint2* cell_cache;
};

// ---- per-sample geometry (bit-identical to the unfused sampler + ComputeLocalSamples chain) -----------------------
struct Sample
{
    float w;
    float uv[3];
    int64_t ray;
};

// Proposal-guided placement: n equal-probability cells of the piecewise-constant density over 8 equal bins whose CDF
// is cdf (cell ends pinned to tmin / tmax); the sample sits at t1 (1 - a) + t2 a in its cell, weighted by the cell's
// length. Same operations as the oracle's proposal_cell / sample_tw.
__device__ __forceinline__ float inv_cdf8(const float (&cdf)[8], float u)
{
    int b    = 0;
    float lo = 0.f, hi = cdf[0];
#pragma unroll
    for (int i = 0; i < 7; ++i)
    {
        if (b == i && cdf[i] <= u)
        {
            lo = cdf[i];
            hi = cdf[i + 1];
            b  = i + 1;
        }
    }
    const float pb = hi - lo;
    float f        = pb > 0.f ? (u - lo) / pb : 0.f;
    f              = fminf(fmaxf(f, 0.f), 1.f);
    return ((float)b + f) * 0.125f;
}

__device__ __forceinline__ float proposal_cell(const float (&cdf)[8], float tmin, float tmax, int32_t n, int32_t j)
{
    if (j <= 0) return tmin;
    if (j >= n) return tmax;
    return tmin + (tmax - tmin) * inv_cdf8(cdf, (float)j / (float)n);
}

// t and w of sample j of a segment: uniform (HyperTree.cu:339-389) or inside the proposal's cell j.
__device__ __forceinline__ void sample_tw(const ProjGeom& g, bool prop, const float (&cdf)[8], float tmin, float tmax,
                                          float step, int32_t n, int32_t j, int64_t ray, int32_t rank, float& t,
                                          float& w)
{
    const float a = seg_a(g.jitter, g.seed, ray, rank, j);
    if (!prop)
    {
        t = sample_t(tmin, step, j, a);
        w = sample_w(tmin, tmax, step, n, j, t, g.jitter, g.seed, ray, rank);
        return;
    }
    const float t1 = proposal_cell(cdf, tmin, tmax, n, j);
    const float t2 = proposal_cell(cdf, tmin, tmax, n, j + 1);
    t              = t1 * (1.0f - a) + t2 * a;
    w              = t2 - t1;
}

__device__ __forceinline__ bool load_cdf(const ProjGeom& g, int64_t k, float (&cdf)[8])
{
    if (!g.seg_cdf) return false;
    const float4* r = reinterpret_cast<const float4*>(g.seg_cdf + 8 * k);
    const float4 u = r[0], v = r[1];
    cdf[0] = u.x; cdf[1] = u.y; cdf[2] = u.z; cdf[3] = u.w;
    cdf[4] = v.x; cdf[5] = v.y; cdf[6] = v.z; cdf[7] = v.w;
    return cdf[7] > 0.f;
}

__device__ __forceinline__ Sample sample_geo(const ProjGeom& g, int32_t k, int32_t j)
{
    Sample s;
    const int64_t ray  = g.seg_ray[k];
    const int32_t n    = g.seg_count[k];
    const int32_t rank = (int32_t)(k - g.ray_seg_offset[ray]);
    const int32_t node = g.seg_node[k];
    const float tmin = g.seg_tmin[k], tmax = g.seg_tmax[k];
    const float step = (tmax - tmin) / (float)n;
    float cdf[8];
    const bool prop = load_cdf(g, k, cdf);
    float t;
    sample_tw(g, prop, cdf, tmin, tmax, step, n, j, ray, rank, t, s.w);
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        float p   = g.ro[3 * ray + d] + g.rd[3 * ray + d] * t;
        float mn  = g.pmin[3 * node + d], mx = g.pmax[3 * node + d];
        s.uv[d]   = (p - mn) / (mx - mn) * 2 - 1;
    }
    s.ray = ray;
    return s;
}

// Trilinear corners in a cubic channels-last node grid (offsets in floats from the node's first voxel).
struct Corners
{
    float w[8];
    int32_t off[8];
};

__device__ __forceinline__ int clamp_g(int v, int G) { return min(G - 1, max(v, 0)); }

// Continuous grid coordinate of a local coordinate (align_corners, IndirectGridSample3D.cu:448-455).
__device__ __forceinline__ float grid_coord(float u, int G) { return ((u + 1) / 2) * (float)(G - 1); }

__device__ __forceinline__ Corners corners_ix(float ix, float iy, float iz, int G, int C)
{
    Corners c;
    int x0 = (int)floorf(ix), y0 = (int)floorf(iy), z0 = (int)floorf(iz);
    int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
    c.w[0] = ((float)x1 - ix) * ((float)y1 - iy) * ((float)z1 - iz);
    c.w[1] = (ix - (float)x0) * ((float)y1 - iy) * ((float)z1 - iz);
    c.w[2] = ((float)x1 - ix) * (iy - (float)y0) * ((float)z1 - iz);
    c.w[3] = (ix - (float)x0) * (iy - (float)y0) * ((float)z1 - iz);
    c.w[4] = ((float)x1 - ix) * ((float)y1 - iy) * (iz - (float)z0);
    c.w[5] = (ix - (float)x0) * ((float)y1 - iy) * (iz - (float)z0);
    c.w[6] = ((float)x1 - ix) * (iy - (float)y0) * (iz - (float)z0);
    c.w[7] = (ix - (float)x0) * (iy - (float)y0) * (iz - (float)z0);
    const int cx0 = clamp_g(x0, G), cx1 = clamp_g(x1, G);
    const int cy0 = clamp_g(y0, G), cy1 = clamp_g(y1, G);
    const int cz0 = clamp_g(z0, G), cz1 = clamp_g(z1, G);
    const int zs = G * G * C, ys = G * C;
    c.off[0] = cz0 * zs + cy0 * ys + cx0 * C;
    c.off[1] = cz0 * zs + cy0 * ys + cx1 * C;
    c.off[2] = cz0 * zs + cy1 * ys + cx0 * C;
    c.off[3] = cz0 * zs + cy1 * ys + cx1 * C;
    c.off[4] = cz1 * zs + cy0 * ys + cx0 * C;
    c.off[5] = cz1 * zs + cy0 * ys + cx1 * C;
    c.off[6] = cz1 * zs + cy1 * ys + cx0 * C;
    c.off[7] = cz1 * zs + cy1 * ys + cx1 * C;
    return c;
}

__device__ __forceinline__ Corners corners_cl(const float uv[3], int G, int C)
{
    return corners_ix(grid_coord(uv[0], G), grid_coord(uv[1], G), grid_coord(uv[2], G), G, C);
}

// ---- the BF16 form's per-sample hand-off from the backward to the grid-gradient scatter (24 B per sample) --------
// dL/dfeature as 8 BF16 channels ([S][8], round to nearest even) and the grid coordinate packed into 8 B: the cell
// (x0 + 1, y0 + 1, z0 + 1 in 5 bits each, x0 = floor(ix) clamped to [-1, 30]; grids here have G <= 17) and the
// position inside it (ix - x0, ...) as 16-bit fractions in units of 2^-16 (rounded to nearest, at most 65535):
// x = fx | fy << 16, y = fz | cell << 16, so y >> 16 alone tells the cells apart. The fractions' quantum, 2^-16 of a
// cell, is far below the BF16 rounding of dL/dfeature that the same weights multiply.
__device__ __forceinline__ uint2 pack_cell_coord(float ix, float iy, float iz)
{
    auto frac16 = [](float f, float f0) { return (uint32_t)fminf(__builtin_rintf((f - f0) * 65536.f), 65535.f); };
    auto cell5  = [](float f0) { return (uint32_t)min(31, max(0, (int)f0 + 1)); };
    const float x0 = floorf(ix), y0 = floorf(iy), z0 = floorf(iz);
    return make_uint2(frac16(ix, x0) | frac16(iy, y0) << 16,
                      frac16(iz, z0) | (cell5(x0) | cell5(y0) << 5 | cell5(z0) << 10) << 16);
}

__device__ __forceinline__ uint2 bf16x4_bits(float a, float b, float c, float d)
{
    typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
    typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
    const bf16x4_t v = {(__bf16)a, (__bf16)b, (__bf16)c, (__bf16)d};
    const u32x2_t u  = __builtin_bit_cast(u32x2_t, v);
    return make_uint2(u[0], u[1]);
}

__device__ __forceinline__ void bf16x8_floats(const uint4& v, float (&x)[8])
{
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int h = 0; h < 4; ++h)
    {
        x[2 * h]     = __uint_as_float(w[h] << 16);
        x[2 * h + 1] = __uint_as_float(w[h] & 0xFFFF0000u);
    }
}

template <int C>
__device__ __forceinline__ void gather_cl(const float* node_grid, const Corners& c, float (&x)[C])
{
#pragma unroll
    for (int q = 0; q < C; q += 4)
    {
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 8; ++k)
        {
            float4 v = *reinterpret_cast<const float4*>(node_grid + c.off[k] + q);
            acc.x += v.x * c.w[k];
            acc.y += v.y * c.w[k];
            acc.z += v.z * c.w[k];
            acc.w += v.w * c.w[k];
        }
        x[q]     = acc.x;
        x[q + 1] = acc.y;
        x[q + 2] = acc.z;
        x[q + 3] = acc.w;
    }
}

// Register-resident decoder C -> H -> H -> 1 (one hidden layer), packed parameters read with uniform loads.
// Returns the output pre-activation.
template <int C, int H, int ACT>
__device__ __forceinline__ float mlp_forward_reg(const float* __restrict__ P, const float (&x)[C], float beta)
{
    const float* W0 = P;
    const float* b0 = W0 + H * C;
    const float* W1 = b0 + H;
    const float* b1 = W1 + H * H;
    const float* W2 = b1 + H;
    const float* b2 = W2 + H;
    float h0[H];
#pragma unroll
    for (int o = 0; o < H; ++o)
    {
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < C; ++k) acc = fmaf(W0[o * C + k], x[k], acc);
        h0[o] = act_f(ACT, acc + b0[o], beta);
    }
    float acc2 = 0.f;
#pragma unroll
    for (int o = 0; o < H; ++o)
    {
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < H; ++k) acc = fmaf(W1[o * H + k], h0[k], acc);
        float h1 = act_f(ACT, acc + b1[o], beta);
        acc2     = fmaf(W2[o], h1, acc2);
    }
    return acc2 + b2[0];
}

// Node-major position of the segment holding chunk-relative sample s: the last q in [q0, q1) with
// nm_soff[q] - base <= s.
__device__ __forceinline__ int32_t find_segment(const int64_t* nm_soff, int32_t q0, int32_t q1, int64_t target)
{
    int32_t lo = q0, hi = q1 - 1;
    while (lo < hi)
    {
        int32_t mid = (lo + hi + 1) >> 1;
        if (nm_soff[mid] <= target)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// ---- matrix-core path ---------------------------------------------------------------------------------------------------
// Gather of the lane's 4 channels (group g: channels 4g..4g+3) of an 8-channel node grid, FP32 or BF16 storage,
// trilinear weights and accumulation in FP32 (same corner order as the reference).
template <bool kBF16Grid>
__device__ __forceinline__ void gather_half(const void* node_grid, const Corners& c, int g, float (&x4)[4])
{
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < 8; ++k)
    {
        float4 v;
        if constexpr (kBF16Grid)
        {
            uint2 raw = *reinterpret_cast<const uint2*>(static_cast<const unsigned short*>(node_grid) + c.off[k] + 4 * g);
            v = make_float4(__uint_as_float(raw.x << 16), __uint_as_float(raw.x & 0xFFFF0000u),
                            __uint_as_float(raw.y << 16), __uint_as_float(raw.y & 0xFFFF0000u));
        }
        else
        {
            v = *reinterpret_cast<const float4*>(static_cast<const float*>(node_grid) + c.off[k] + 4 * g);
        }
        if constexpr (kBF16Grid)
        {
            // the BF16 form is not bound to the oracle's rounding: fused multiply-adds
            acc.x = fmaf(v.x, c.w[k], acc.x);
            acc.y = fmaf(v.y, c.w[k], acc.y);
            acc.z = fmaf(v.z, c.w[k], acc.z);
            acc.w = fmaf(v.w, c.w[k], acc.w);
        }
        else
        {
            acc.x += v.x * c.w[k];
            acc.y += v.y * c.w[k];
            acc.z += v.z * c.w[k];
            acc.w += v.w * c.w[k];
        }
    }
    x4[0] = acc.x;
    x4[1] = acc.y;
    x4[2] = acc.z;
    x4[3] = acc.w;
}

__device__ __forceinline__ const void* node_grid_ptr(const ProjGeom& g, const void* grid_lowp, bool bf16, int32_t slot)
{
    const int64_t off = (int64_t)slot * g.G * g.G * g.G * 8;
    return bf16 ? static_cast<const void*>(static_cast<const unsigned short*>(grid_lowp) + off)
                : static_cast<const void*>(g.grid + off);
}

// Segment state of one forward chunk, staged in LDS once per chunk from the plan's node-major records: what the
// chunk's samples need from the segment list and the rays, so that a sample's geometry costs LDS reads.
struct ChunkCtx
{
    int32_t beg[kSegCap + 1];  // chunk-relative first sample of each segment; beg[nseg] = the chunk's sample count
    int32_t k[kSegCap];        // canonical segment index
    int32_t ray[kSegCap];
    int32_t rank[kSegCap];
    int32_t n[kSegCap];
    float tmin[kSegCap], tmax[kSegCap], step[kSegCap];
    float gproj[kSegCap];  // dL/dprojection of the segment's ray (backward)
    float o[3][kSegCap], d[3][kSegCap];
    float cell[8][kSegCap];     // kind 1: proposal cell boundaries 1..7; kind 2: the proposal CDF row
    uint32_t hray[2][kSegCap];  // jitter_ray_hash of the segment's ray (low, high word)
    uint8_t kind[kSegCap];      // 0 uniform samples, 1 proposal cells precomputed (n <= 8), 2 proposal CDF
    uint16_t sq[kFwdChunk];     // window-relative sample -> chunk-relative segment
};

// Record of the segment at node-major position q (k = perm[q]): what ChunkCtx holds, computed with the same
// operations as sample_geo and proposal_cell so every derived value has the same bits.
__device__ inline void make_record(const ProjGeom& g, const Plan& p, int64_t q)
{
    const int32_t k   = p.perm[q];
    const int32_t ray = g.seg_ray[k];
    const int32_t n   = g.seg_count[k];
    const float tmin = g.seg_tmin[k], tmax = g.seg_tmax[k];
    const float step  = (tmax - tmin) / (float)n;
    const uint64_t hr = jitter_ray_hash(g.seed, ray);
    uint32_t kind     = 0;
    float cdf[8];
    if (load_cdf(g, k, cdf))
    {
        kind = n <= 8 ? 1u : 2u;
        float cell[8];
#pragma unroll
        for (int b = 0; b < 8; ++b) cell[b] = cdf[b];
        if (kind == 1)
        {
#pragma unroll
            for (int j = 1; j <= 7; ++j) cell[j - 1] = proposal_cell(cdf, tmin, tmax, n, j);
            cell[7] = 0.f;
        }
        p.rec_c[2 * q]     = make_float4(cell[0], cell[1], cell[2], cell[3]);
        p.rec_c[2 * q + 1] = make_float4(cell[4], cell[5], cell[6], cell[7]);
    }
    p.rec_ot[q] = make_float4(g.ro[3 * (int64_t)ray], g.ro[3 * (int64_t)ray + 1], g.ro[3 * (int64_t)ray + 2], tmin);
    p.rec_dt[q] = make_float4(g.rd[3 * (int64_t)ray], g.rd[3 * (int64_t)ray + 1], g.rd[3 * (int64_t)ray + 2], tmax);
    p.rec_i[q]  = make_int4(k, ray, (int32_t)(k - g.ray_seg_offset[ray]), n);
    p.rec_h[q]  = make_uint4((uint32_t)hr, (uint32_t)(hr >> 32), kind, __float_as_uint(step));
}

__device__ inline void load_chunk_ctx(const ProjGeom& g, const Plan& p, int32_t q0, int32_t q1, int64_t base,
                                      const float* grad_proj, ChunkCtx& c)
{
    const int32_t nseg = q1 - q0;
    for (int32_t t = threadIdx.x; t < nseg; t += blockDim.x)
    {
        const int32_t q = q0 + t;
        const float4 ot = p.rec_ot[q], dt = p.rec_dt[q];
        const int4 ri   = p.rec_i[q];
        const uint4 rh  = p.rec_h[q];
        c.beg[t] = (int32_t)(p.nm_soff[q] - base);
        if (t == nseg - 1) c.beg[nseg] = (int32_t)(p.nm_soff[q1] - base);
        c.k[t]       = ri.x;
        c.ray[t]     = ri.y;
        c.rank[t]    = ri.z;
        c.n[t]       = ri.w;
        c.tmin[t]    = ot.w;
        c.tmax[t]    = dt.w;
        c.step[t]    = __uint_as_float(rh.w);
        c.gproj[t]   = grad_proj ? grad_proj[ri.y] : 0.f;
        c.hray[0][t] = rh.x;
        c.hray[1][t] = rh.y;
        c.kind[t]    = (uint8_t)rh.z;
        c.o[0][t] = ot.x;
        c.o[1][t] = ot.y;
        c.o[2][t] = ot.z;
        c.d[0][t] = dt.x;
        c.d[1][t] = dt.y;
        c.d[2][t] = dt.z;
        if (rh.z != 0)
        {
            const float4 u = p.rec_c[2 * q], v = p.rec_c[2 * q + 1];
            c.cell[0][t] = u.x;
            c.cell[1][t] = u.y;
            c.cell[2][t] = u.z;
            c.cell[3][t] = u.w;
            c.cell[4][t] = v.x;
            c.cell[5][t] = v.y;
            c.cell[6][t] = v.z;
            c.cell[7][t] = v.w;
        }
    }
}

// sq for the window [w0, w0 + wn) of the chunk's samples (needs beg[] complete).
__device__ inline void expand_window(ChunkCtx& c, int32_t nseg, int32_t w0, int32_t wn)
{
    for (int32_t t = threadIdx.x; t < nseg; t += blockDim.x)
    {
        const int32_t a = max(c.beg[t], w0), b = min(c.beg[t + 1], w0 + wn);
        for (int32_t i = a; i < b; ++i) c.sq[i - w0] = (uint16_t)t;
    }
}

// Sample j of chunk segment t: the same operations as sample_geo on the same values, so the same bits.
// A chunk's node box. rinv[d] is 1 / (mx[d] - mn[d]) when that extent is a power of two (every octree box is), where
// x * rinv equals x / extent exactly (both round the same exact quotient); 0 otherwise, and then the division runs.
struct NodeBox
{
    float mn[3], mx[3], rinv[3];
};

__device__ __forceinline__ Sample ctx_sample(const ProjGeom& g, const ChunkCtx& c, int32_t t, int32_t j,
                                             const NodeBox& nb)
{
    Sample s;
    const int64_t ray  = c.ray[t];
    const int32_t rank = c.rank[t], n = c.n[t];
    const float tmin = c.tmin[t], tmax = c.tmax[t], step = c.step[t];
    const uint64_t hr = ((uint64_t)c.hray[1][t] << 32) | c.hray[0][t];
    auto a_of         = [&](int32_t jj) { return g.jitter ? jitter_a_h(hr, rank, jj) : 0.5f; };
    float tt;
    const uint32_t kind = c.kind[t];
    // one position formula for both kinds (sample_t is the uniform cell's t1 (1 - a) + t2 a); the kinds differ in
    // the cell ends and the weight, so the lanes of a wave stay together through the hash and the position
    float t1, t2;
    if (kind == 2)
    {
        float cdf[8];
#pragma unroll
        for (int b = 0; b < 8; ++b) cdf[b] = c.cell[b][t];
        t1 = proposal_cell(cdf, tmin, tmax, n, j);
        t2 = proposal_cell(cdf, tmin, tmax, n, j + 1);
    }
    else if (kind == 1)
    {
        // boundaries 1..7 precomputed by the plan (proposal_cell of the same CDF row)
        t1 = j == 0 ? tmin : c.cell[j - 1][t];
        t2 = j + 1 >= n ? tmax : c.cell[j][t];
    }
    else
    {
        t1 = tmin + (float)j * step;
        t2 = tmin + (float)(j + 1) * step;
    }
    const float a = a_of(j);
    tt            = t1 * (1.0f - a) + t2 * a;
    if (kind != 0)
    {
        s.w = t2 - t1;
    }
    else
    {
        // midpoint-cell weight from the neighbours' positions (sample_w), their hashes taken only here
        const float lo = (j == 0) ? tmin : (sample_t(tmin, step, j - 1, a_of(j - 1)) + tt) * 0.5f;
        const float hi = (j == n - 1) ? tmax : (sample_t(tmin, step, j + 1, a_of(j + 1)) + tt) * 0.5f;
        s.w            = hi - lo;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        const float pd  = c.o[d][t] + c.d[d][t] * tt;
        const float rel = nb.rinv[d] != 0.f ? (pd - nb.mn[d]) * nb.rinv[d] : (pd - nb.mn[d]) / (nb.mx[d] - nb.mn[d]);
        s.uv[d]         = rel * 2 - 1;
    }
    s.ray = ray;
    return s;
}

__device__ __forceinline__ void node_bounds(const ProjGeom& g, int32_t node, NodeBox& nb)
{
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        nb.mn[d]         = g.pmin[3 * node + d];
        nb.mx[d]         = g.pmax[3 * node + d];
        const float e    = nb.mx[d] - nb.mn[d];
        const uint32_t b = __float_as_uint(e);
        const bool pow2  = e > 0.f && (b & 0x7FFFFFu) == 0u && (b >> 23) > 1u && (b >> 23) < 253u;
        nb.rinv[d]       = pow2 ? 1.0f / e : 0.f;
    }
}

// Inputs of a pair of 32-sample tiles (samples p0 .. p0 + 63 of the window): every lane computes the geometry of one
// sample (lane l: sample p0 + l) and the lane groups swap their corners, so that each lane holds the corners of
// sample p0 + n (tile A) and p0 + 32 + n (tile B) and gathers its 4 channels of both. v is the sample weight
// (forward) or dL/dy = dL/dproj x weight (backward); samples past the window have zero weights, features and v.
// The backward also stores each sample's grid coordinate for the scatter: (ix, iy, iz, 0) in the FP32 form, the
// packed cell and fractions (pack_cell_coord) in the BF16 form. A BF16 forward given sgc and w_out stores the packed
// coordinate and the sample weight (what its backward reads instead of recomputing them).
struct PairIn
{
    float xa[4], xb[4];
    float va, vb;
};

template <bool kBF16, bool kBwd>
__device__ __forceinline__ void fetch_pair(const ProjGeom& g, const ChunkCtx& cx, int32_t p0, int32_t w0,
                                           int32_t wn, int64_t base, const NodeBox& nb, const void* ngrid,
                                           float4* sgc, PairIn& in, float* w_out = nullptr)
{
    const int l = threadIdx.x & 63, hg = l >> 5;
    const int32_t ii = p0 + l;
    Corners cr;
    float v = 0.f;
    if (ii < wn)
    {
        const int32_t t = cx.sq[ii];
        Sample smp      = ctx_sample(g, cx, t, w0 + ii - cx.beg[t], nb);
        const float ix = grid_coord(smp.uv[0], g.G), iy = grid_coord(smp.uv[1], g.G), iz = grid_coord(smp.uv[2], g.G);
        cr = corners_ix(ix, iy, iz, g.G, 8);
        v  = kBwd ? cx.gproj[t] * smp.w : smp.w;
        if constexpr (kBF16)
        {
            if (kBwd || sgc) reinterpret_cast<uint2*>(sgc)[base + w0 + ii] = pack_cell_coord(ix, iy, iz);
            if (!kBwd && w_out) w_out[base + w0 + ii] = smp.w;
        }
        else if constexpr (kBwd)
        {
            sgc[base + w0 + ii] = make_float4(ix, iy, iz, 0.f);
        }
    }
    else
    {
#pragma unroll
        for (int k = 0; k < 8; ++k)
        {
            cr.off[k] = 0;
            cr.w[k]   = 0.f;
        }
    }
    // the lane groups' halves swapped across (v_permlane32_swap): lanes of group 0 keep their sample as tile A and
    // get group 1's as tile B, lanes of group 1 the other way round
    Corners ca, cb;
#pragma unroll
    for (int k = 0; k < 8; ++k)
    {
        const auto o = __builtin_amdgcn_permlane32_swap((uint32_t)cr.off[k], (uint32_t)cr.off[k], false, false);
        ca.off[k]    = (int)o[0];
        cb.off[k]    = (int)o[1];
        const auto w = __builtin_amdgcn_permlane32_swap(__float_as_uint(cr.w[k]), __float_as_uint(cr.w[k]), false, false);
        ca.w[k]      = __uint_as_float(w[0]);
        cb.w[k]      = __uint_as_float(w[1]);
    }
    const auto vv = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    gather_half<kBF16>(ngrid, ca, hg, in.xa);
    gather_half<kBF16>(ngrid, cb, hg, in.xb);
    in.va = __uint_as_float(vv[0]);
    in.vb = __uint_as_float(vv[1]);
}

// Copies a node's BF16 grid (nvert vertices x 8 channels, 16 B per vertex) into LDS, four loads in flight per thread.
__device__ __forceinline__ void stage_node_grid(const uint4* __restrict__ src, uint4* dst, int32_t nvert)
{
    // every load runs (clamped to the last vertex) so that the four stay in registers and are in flight together
    for (int32_t t0 = threadIdx.x; t0 < nvert; t0 += 4 * (int32_t)blockDim.x)
    {
        uint4 v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = src[min(t0 + k * (int32_t)blockDim.x, nvert - 1)];
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            const int32_t t = t0 + k * (int32_t)blockDim.x;
            if (t < nvert) dst[t] = v[k];
        }
    }
}

// ---- relu BF16 backward (projector_relu.hip) ----------------------------------------------------------------------
// Launches k_mfma_fwd_relu<H> (H = 32 or 64) on the plan's p.nwg node-resident workgroups of kNodeThreads.
int launch_mfma_fwd_relu(hipStream_t s, int H, const ProjGeom& g, const Plan& p, const float* params,
                         const void* grid_bf16, float* seg_partial, float* sample_out);

// Launches k_mfma_bwd_relu<H> (H = 32 or 64) on the plan's p.nwg node-resident workgroups of kNodeThreads; every
// wave writes one parameter-gradient slice (p.nwg * kNodeWaves slices). It reads the per-sample caches that
// k_mfma_fwd_relu wrote into the same plan: the forward must have run on it with the same arguments.
int launch_mfma_bwd_relu(hipStream_t s, int H, const ProjGeom& g, const Plan& p, const float* params,
                         const float* grad_proj, float* dfeat, unsigned int* dmax, float* partials, int32_t P_total);

}  // namespace neat
