// Fused FP32 projector: forward line integrals and backward to grid and decoder gradients over the canonical
// segment list, with the samples recomputed from the segments (never materialised as a SampleList).
//
// It fuses, per sample, the chain the reference runs as separate kernels and ATen ops -- sample position
// (HyperTree.cu:339-358, :426-432), local coordinates (:177-192), slot lookup (HyperTree.h:183-187), trilinear
// gather (IndirectGridSample3D.cu:432-572), decoder + softplus (ImplicitNet.h:18-71, geometry.cpp:102-163) and the
// weighted per-ray sum (geometry.cpp:393-424) -- and the hand-written backward of all of it.
//
// Work is node-major: a plan sorts the segments by grid slot (stable, so the order inside a slot is canonical) and
// cuts every slot's samples into chunks; a workgroup processes one chunk, so all its gathers hit one node's grid.
//   forward  : one thread per sample, register-resident MLP; per-segment partial sums (sequential in j), then
//              per-ray sums over the ray's segments in canonical order. Deterministic.
//   backward : pass 1 recomputes each sample, runs the decoder backward tile by tile (one wave, 64 samples,
//              LDS-staged weight-gradient reduction, per-workgroup partials) and writes dL/dfeature per sample;
//              pass 2 scatters dL/dfeature x corner weight into the node's gradient grid held in LDS and stores the
//              chunk's grid with plain stores; slots cut into several chunks are summed in chunk order. No global
//              atomics anywhere.
#include <hipcub/hipcub.hpp>

#include <type_traits>

#include "projector_dev.hpp"

namespace neat
{

// ---- plan ---------------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_plan_keys(ProjGeom g, int32_t* keys, int32_t* vals)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g.K) return;
    int32_t node = g.seg_node[k];
    int64_t slot = (node >= 0 && node < g.num_nodes) ? g.prefix[node] : -1;
    keys[k]      = (slot >= 0 && slot < g.N) ? (int32_t)slot : (int32_t)g.N;  // invalid slots sort last
    vals[k]      = (int32_t)k;
}

__global__ void __launch_bounds__(256) k_plan_counts(ProjGeom g, const int32_t* perm, int64_t* nm_soff)
{
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < g.K) nm_soff[q] = g.seg_count[perm[q]];
}

__global__ void __launch_bounds__(256) k_plan_slots(ProjGeom g, const int32_t* skeys, Plan p)
{
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s > g.N) return;
    // lower bound of s in the sorted keys
    int64_t lo = 0, hi = g.K;
    while (lo < hi)
    {
        int64_t mid = (lo + hi) >> 1;
        if (skeys[mid] < s)
            lo = mid + 1;
        else
            hi = mid;
    }
    p.slot_begin[s] = (int32_t)lo;
}

// Forward / backward-pass-1 chunks are cut by the cost samples + kSegCost x segments, in steps of kFwdChunk: a
// chunk then holds fewer than kFwdChunk / kSegCost + 1 segments and fewer than kFwdChunk samples plus its last
// segment's.
__global__ void __launch_bounds__(256) k_plan_chunk_counts(ProjGeom g, Plan p)
{
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.N) return;
    const int32_t sb = p.slot_begin[s], se = p.slot_begin[s + 1];
    int64_t n       = p.nm_soff[se] - p.nm_soff[sb];
    p.chunk_off1[s] = (n + (int64_t)kSegCost * (se - sb) + kFwdChunk - 1) / kFwdChunk;
    p.chunk_off2[s] = (n + p.chunk2 - 1) / p.chunk2;
}

// First node-major segment in [lo, hi) whose first sample is >= target.
__device__ __forceinline__ int32_t lower_seg(const int64_t* nm_soff, int32_t lo, int32_t hi, int64_t target)
{
    while (lo < hi)
    {
        int32_t mid = (lo + hi) >> 1;
        if (nm_soff[mid] < target)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// First node-major segment q in [lo, hi) with cost(q) = nm_soff[q] - base + kSegCost (q - sb) >= target.
__device__ __forceinline__ int32_t lower_cost(const int64_t* nm_soff, int32_t sb, int32_t lo, int32_t hi, int64_t base,
                                              int64_t target)
{
    while (lo < hi)
    {
        int32_t mid = (lo + hi) >> 1;
        if (nm_soff[mid] - base + (int64_t)kSegCost * (mid - sb) < target)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// One thread per chunk. Scatter chunk r of a slot covers the slot's segments whose first sample lies in
// [r*T, (r+1)*T) (relative to the slot's first sample); forward chunk r those whose cost lies in that range.
// Threads [0, max1) fill forward chunks, [max1, max1 + max2) scatter chunks.
__global__ void __launch_bounds__(256) k_plan_chunks(ProjGeom g, Plan p)
{
    int64_t i       = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool pass1 = i < p.max1;
    if (!pass1) i -= p.max1;
    const int64_t* off = pass1 ? p.chunk_off1 : p.chunk_off2;
    if (i >= (pass1 ? p.max1 : p.max2) || i >= off[g.N]) return;
    // slot owning chunk i: last s with off[s] <= i
    int64_t lo = 0, hi = g.N - 1;
    while (lo < hi)
    {
        int64_t mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int64_t s    = lo;
    const int64_t r    = i - off[s];
    const int64_t T    = pass1 ? kFwdChunk : p.chunk2;
    const int32_t sb   = p.slot_begin[s], se = p.slot_begin[s + 1];
    const int64_t base = p.nm_soff[sb];
    int32_t q0, q1;
    if (pass1)
    {
        q0 = lower_cost(p.nm_soff, sb, sb, se, base, r * T);
        q1 = (i + 1 < off[s + 1]) ? lower_cost(p.nm_soff, sb, sb, se, base, (r + 1) * T) : se;
    }
    else
    {
        q0 = lower_seg(p.nm_soff, sb, se, base + r * T);
        q1 = (i + 1 < off[s + 1]) ? lower_seg(p.nm_soff, sb, se, base + (r + 1) * T) : se;
    }
    (pass1 ? p.chunk1_slot : p.chunk2_slot)[i] = (int32_t)s;
    (pass1 ? p.chunk1_q0 : p.chunk2_q0)[i]     = q0;
    (pass1 ? p.chunk1_q1 : p.chunk2_q1)[i]     = q1;
}

// Per-segment records in node-major order (make_record), after the sort.
__global__ void __launch_bounds__(256) k_plan_records(ProjGeom g, Plan p)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < g.K) make_record(g, p, q);
}

// Node-resident workgroup ranges: workgroup b takes the forward chunks whose first segment starts at a cost
// (samples + kSegCost x segments, over the node-major list) in [b T / nwg, (b + 1) T / nwg), T the total cost;
// thread b writes the first chunk of workgroup b (b = nwg: the chunk count).
__global__ void __launch_bounds__(256) k_plan_wg(ProjGeom g, Plan p)
{
    const int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > p.nwg) return;
    const int64_t nch    = p.chunk_off1[g.N];
    const int32_t kv     = p.slot_begin[g.N];
    const int64_t total  = p.nm_soff[kv] + (int64_t)kSegCost * kv;
    const int64_t target = total * b / p.nwg;
    int64_t lo = 0, hi = nch;
    while (lo < hi)
    {
        const int64_t mid = (lo + hi) >> 1;
        const int32_t q   = p.chunk1_q0[mid];
        if (p.nm_soff[q] + (int64_t)kSegCost * q < target)
            lo = mid + 1;
        else
            hi = mid;
    }
    p.wg_first[b] = (int32_t)(b == p.nwg ? nch : lo);
}

// ---- forward --------------------------------------------------------------------------------------------------------
template <int C, int H, int ACT>
__global__ void __launch_bounds__(kFwdThreads) k_proj_fwd(ProjGeom g, Plan p, const float* __restrict__ params,
                                                          float* seg_partial)
{
    // The chunk's samples are processed in windows of kFwdChunk (a chunk exceeds one window only when it ends with
    // a segment longer than the remaining space); each segment's sum is accumulated window by window, sequentially
    // in j inside a window. A chunk holds at most kFwdChunk segments (every segment starts inside it).
    __shared__ float s_contrib[kFwdChunk];
    __shared__ float s_seg[kFwdChunk];
    const int64_t c = blockIdx.x;
    if (c >= p.chunk_off1[g.N]) return;
    const int32_t slot = p.chunk1_slot[c];
    const int32_t q0 = p.chunk1_q0[c], q1 = p.chunk1_q1[c];
    const int64_t base  = p.nm_soff[q0];
    const int64_t nsamp = p.nm_soff[q1] - base;
    const float* node_grid = g.grid + (int64_t)slot * g.G * g.G * g.G * C;
    for (int32_t q = q0 + threadIdx.x; q < q1; q += kFwdThreads) s_seg[q - q0] = 0.f;
    for (int64_t w0 = 0; w0 < nsamp; w0 += kFwdChunk)
    {
        const int32_t wn = (int32_t)min<int64_t>(kFwdChunk, nsamp - w0);
        for (int32_t i = threadIdx.x; i < wn; i += kFwdThreads)
        {
            const int64_t s = w0 + i;
            const int32_t q = find_segment(p.nm_soff, q0, q1, base + s);
            const int32_t k = p.perm[q];
            const int32_t j = (int32_t)(base + s - p.nm_soff[q]);
            Sample smp      = sample_geo(g, k, j);
            Corners cr      = corners_cl(smp.uv, g.G, C);
            float x[C];
            gather_cl<C>(node_grid, cr, x);
            float z      = mlp_forward_reg<C, H, ACT>(params, x, g.beta);
            s_contrib[i] = act_f(NEAT_ACT_SOFTPLUS, z, g.beta) * smp.w;
        }
        __syncthreads();
        for (int32_t q = q0 + threadIdx.x; q < q1; q += kFwdThreads)
        {
            const int64_t a = max<int64_t>(p.nm_soff[q] - base, w0);
            const int64_t b = min<int64_t>(p.nm_soff[q + 1] - base, w0 + wn);
            if (a >= b) continue;
            float acc = 0.f;
            for (int64_t s = a; s < b; ++s) acc += s_contrib[s - w0];
            s_seg[q - q0] += acc;
        }
        __syncthreads();
    }
    for (int32_t q = q0 + threadIdx.x; q < q1; q += kFwdThreads) seg_partial[p.perm[q]] = s_seg[q - q0];
}

// proj[r] = sum of the ray's segment partials in canonical order (optionally Neumaier-compensated).
__global__ void __launch_bounds__(256) k_proj_raysum(const int64_t* ray_seg_offset, int64_t R,
                                                     const float* seg_partial, int32_t compensated, float* proj)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    float sum = 0.f, comp = 0.f;
    for (int64_t k = ray_seg_offset[r]; k < ray_seg_offset[r + 1]; ++k)
    {
        float v = seg_partial[k];
        if (compensated)
        {
            float t = sum + v;
            comp += (fabsf(sum) >= fabsf(v)) ? (sum - t) + v : (v - t) + sum;
            sum = t;
        }
        else
        {
            sum += v;
        }
    }
    proj[r] = sum + comp;
}

// ---- backward pass 1: decoder backward per tile, dL/dfeature per sample ------------------------------------------------
template <int C, int H>
__global__ void __launch_bounds__(kTile) k_proj_bw1(ProjGeom g, Plan p, MLP m, const float* __restrict__ wt, int act,
                                                    const float* grad_proj, float* dfeat, float* partials)
{
    __shared__ float lds[kTileFloats];
    TileBufs t     = tile_bufs(lds);
    float* gacc    = partials + (int64_t)blockIdx.x * m.nparams();
    const int lane = threadIdx.x;
    const int64_t nchunks = p.chunk_off1[g.N];
    for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x)
    {
        const int32_t slot = p.chunk1_slot[c];
        const int32_t q0 = p.chunk1_q0[c], q1 = p.chunk1_q1[c];
        const int64_t base  = p.nm_soff[q0];
        const int32_t nsamp = (int32_t)(p.nm_soff[q1] - base);
        const float* node_grid = g.grid + (int64_t)slot * g.G * g.G * g.G * C;
        for (int32_t s0 = 0; s0 < nsamp; s0 += kTile)
        {
            const int32_t s  = s0 + lane;
            const bool valid = s < nsamp;
            const int ns     = min(kTile, nsamp - s0);
            float gy         = 0.f;
            float x[C];
#pragma unroll
            for (int q = 0; q < C; ++q) x[q] = 0.f;
            if (valid)
            {
                const int32_t q = find_segment(p.nm_soff, q0, q1, base + s);
                const int32_t k = p.perm[q];
                const int32_t j = (int32_t)(base + s - p.nm_soff[q]);
                Sample smp      = sample_geo(g, k, j);
                Corners cr      = corners_cl(smp.uv, g.G, C);
                gather_cl<C>(node_grid, cr, x);
                gy = grad_proj[smp.ray] * smp.w;
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < kMaxIn; ++q) t.X[q * kS + lane] = q < C ? x[q] : 0.f;
            __syncthreads();
            tile_forward<H>(m, wt, WLayout{true}, act, g.beta, t);
            t.D[lane] = valid ? gy * act_d(NEAT_ACT_SOFTPLUS, t.D[lane], g.beta) : 0.f;
            tile_backward<H>(m, wt, WLayout{true}, act, g.beta, t, ns, gacc);
            if (valid)
            {
                float* out = dfeat + (base + s) * C;
#pragma unroll
                for (int q = 0; q < C; q += 4)
                    *reinterpret_cast<float4*>(out + q) =
                        make_float4(t.X[q * kS + lane], t.X[(q + 1) * kS + lane], t.X[(q + 2) * kS + lane],
                                    t.X[(q + 3) * kS + lane]);
            }
        }
    }
}

// ---- backward pass 2: scatter into the node's gradient grid in LDS ----------------------------------------------------
// Per-segment state of a sample stream: what the samples of one segment share.
struct SegCtx
{
    int64_t ray;
    int32_t rank, n;
    float tmin, tmax, step;
    float o[3], d[3];
    bool prop;
    float cdf[8];
};

__device__ __forceinline__ SegCtx seg_ctx(const ProjGeom& g, int32_t k)
{
    SegCtx c;
    c.ray  = g.seg_ray[k];
    c.n    = g.seg_count[k];
    c.rank = (int32_t)(k - g.ray_seg_offset[c.ray]);
    c.tmin = g.seg_tmin[k];
    c.tmax = g.seg_tmax[k];
    c.step = (c.tmax - c.tmin) / (float)c.n;
    c.prop = load_cdf(g, k, c.cdf);
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        c.o[d] = g.ro[3 * c.ray + d];
        c.d[d] = g.rd[3 * c.ray + d];
    }
    return c;
}

// Local coordinates of sample j: the same operations as sample_geo, so the same bits.
__device__ __forceinline__ void seg_uv(const ProjGeom& g, const SegCtx& c, const float (&mn)[3], const float (&mx)[3],
                                       int32_t j, float (&uv)[3])
{
    float t;
    if (!c.prop)
    {
        t = sample_t(c.tmin, c.step, j, seg_a(g.jitter, g.seed, c.ray, c.rank, j));
    }
    else
    {
        const float a  = seg_a(g.jitter, g.seed, c.ray, c.rank, j);
        const float t1 = proposal_cell(c.cdf, c.tmin, c.tmax, c.n, j);
        const float t2 = proposal_cell(c.cdf, c.tmin, c.tmax, c.n, j + 1);
        t              = t1 * (1.0f - a) + t2 * a;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        float pd = c.o[d] + c.d[d] * t;
        uv[d]    = (pd - mn[d]) / (mx[d] - mn[d]) * 2 - 1;
    }
}

constexpr int kScatterThreads = 1024;

// One workgroup per scatter chunk (a run of node-major samples of one slot). Lane l handles channel l % C of a
// sample, so one LDS add instruction of a wave covers 64 / C samples x C channels (distinct banks for distinct
// vertices). The chunk's samples are cut into 16 x 64 / C contiguous streams, one per lane group, so the samples
// of one instruction lie far apart along the node-major order (different segments / cells: few same-address
// collisions); a stream walks its samples in order and reloads the segment state only at segment boundaries.
// 2^F with F = 45 - exponent(max |dL/dfeature|): every contribution is below 2^45 in fixed point and a vertex's sum
// of fewer than 2^18 of them stays inside int64; quantisation is 2^-46 of the largest value per converted sum.
__device__ __forceinline__ double fixed_scale(unsigned int maxbits)
{
    const float m = __uint_as_float(maxbits);
    if (!(m > 0.f) || !isfinite(m)) return 1.0;
    int e;
    frexpf(m, &e);
    return ldexp(1.0, 45 - e);
}

template <int C>
__global__ void __launch_bounds__(kScatterThreads) k_proj_bw2(ProjGeom g, Plan p, const float* dfeat,
                                                              const float4* sgc, const unsigned int* dmax,
                                                              float* grad_grid, float* partial_grids)
{
    extern __shared__ float sg[];
    constexpr int kSPW = 64 / C;
    const int64_t c = blockIdx.x;
    if (c >= p.chunk_off2[g.N]) return;
    const int32_t slot  = p.chunk2_slot[c];
    const int32_t q0    = p.chunk2_q0[c], q1 = p.chunk2_q1[c];
    const int32_t gsize = g.G * g.G * g.G * C;
    for (int32_t i = threadIdx.x * 4; i < gsize; i += blockDim.x * 4)
        *reinterpret_cast<float4*>(sg + i) = make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ch       = lane % C;
    const int64_t U    = (int64_t)(blockDim.x >> 6) * kSPW;
    const int64_t u    = (int64_t)wave * kSPW + lane / C;
    const int64_t base = p.nm_soff[q0];
    const int64_t ns   = p.nm_soff[q1] - base;
    const int64_t L    = (ns + U - 1) / U;
    int64_t s          = u * L;
    const int64_t send = min(ns, s + L);
    if constexpr (C == 8)
    {
      if (sgc)
      {
        // The backward kept every sample's grid coordinate. Channel half blockIdx.y of the node grid is accumulated
        // in 64-bit fixed point (value x 2^F, F from the batch's largest |dL/dfeature|): integer adds are exact and
        // order-independent, so the result is bit-reproducible, and LDS integer atomics run several times faster
        // than float ones. Lanes are (sample stream, corner) pairs: the 8 lanes of a stream share its sample's loads,
        // each computes one corner's weight and vertex once and adds all 4 channels of the half. The chunk's samples
        // are cut into 8 x 16 contiguous streams. The LDS sums are planar, [channel][vertex]: one atomic instruction
        // adds one channel for 64 vertices, which then spread over all bank pairs (interleaved channels would put
        // every lane of an instruction on the same eighth of the banks).
        unsigned long long* sq = reinterpret_cast<unsigned long long*>(sg);
        const int32_t nv       = g.G * g.G * g.G;
        const int32_t G        = g.G;
        // Lane = sample stream: the chunk's samples are cut into blockDim.x contiguous streams, and a lane walks its
        // stream in order, with all 8 corners x 4 channels of the current trilinear cell summed in FP32 registers;
        // the cell's sums go to the fixed-point LDS grid (32 integer atomics) only when the stream leaves the cell
        // (about every 4 samples along a ray). The cells, their runs and the order inside a run are fixed by the
        // plan, so the grid gradient stays bit-reproducible.
        const int64_t U2 = blockDim.x;
        const int64_t L2 = (ns + U2 - 1) / U2;
        int64_t s2       = (int64_t)threadIdx.x * L2;
        const int64_t e2 = min(ns, s2 + L2);
        const double scale = fixed_scale(*dmax);
        const float4* dh   = reinterpret_cast<const float4*>(dfeat) + (int64_t)blockIdx.y * g.S + base;
        if (s2 < e2)
        {
            int cell = -1;  // (x0 + 1) + (G + 2) ((y0 + 1) + (G + 2) (z0 + 1)) of the current cell (x0 in [-1, G])
            int cx0 = 0, cx1 = 0, cy0 = 0, cy1 = 0, cz0 = 0, cz1 = 0;
            float acc[8][4];
#pragma unroll
            for (int k = 0; k < 8; ++k) acc[k][0] = acc[k][1] = acc[k][2] = acc[k][3] = 0.f;
            auto flush = [&]() {
                const int vz[2] = {cz0 * G * G, cz1 * G * G}, vy[2] = {cy0 * G, cy1 * G}, vx[2] = {cx0, cx1};
#pragma unroll
                for (int k = 0; k < 8; ++k)
                {
                    const int v = vz[k >> 2] + vy[(k >> 1) & 1] + vx[k & 1];
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                    {
                        atomicAdd(sq + q * nv + v, (unsigned long long)__double2ll_rn((double)acc[k][q] * scale));
                        acc[k][q] = 0.f;
                    }
                }
            };
            // blocks of 8 samples: 128 B of each array per stream and block, a whole cache line of each once the
            // stream is aligned to it (its first block runs to the next 8-sample boundary), so a line is read once
            constexpr int kSB = 8;
            for (int64_t s0 = s2; s0 < e2;)
            {
                const int64_t s1 = min(e2, ((base + s0) / kSB + 1) * kSB - base);
                float4 gb[kSB], db[kSB];
#pragma unroll
                for (int b = 0; b < kSB; ++b)
                {
                    if (s0 + b < s1)
                    {
                        gb[b] = sgc[base + s0 + b];
                        db[b] = dh[s0 + b];
                    }
                }
#pragma unroll
                for (int b = 0; b < kSB; ++b)
                {
                    if (s0 + b >= s1) break;
                    const float4 gc = gb[b];
                    // the trilinear corners of corners_ix, the same operations
                    const int x0 = (int)floorf(gc.x), y0 = (int)floorf(gc.y), z0 = (int)floorf(gc.z);
                    const int c = (x0 + 1) + (G + 2) * ((y0 + 1) + (G + 2) * (z0 + 1));
                    if (c != cell)
                    {
                        if (cell >= 0) flush();
                        cell = c;
                        cx0 = clamp_g(x0, G);
                        cx1 = clamp_g(x0 + 1, G);
                        cy0 = clamp_g(y0, G);
                        cy1 = clamp_g(y0 + 1, G);
                        cz0 = clamp_g(z0, G);
                        cz1 = clamp_g(z0 + 1, G);
                    }
                    const float fx1 = gc.x - (float)x0, fx0 = (float)(x0 + 1) - gc.x;
                    const float fy1 = gc.y - (float)y0, fy0 = (float)(y0 + 1) - gc.y;
                    const float fz1 = gc.z - (float)z0, fz0 = (float)(z0 + 1) - gc.z;
                    const float wxy[4] = {fx0 * fy0, fx1 * fy0, fx0 * fy1, fx1 * fy1};
                    const float4 dv = db[b];
#pragma unroll
                    for (int k = 0; k < 8; ++k)
                    {
                        // (fx * fy) * fz: the product order of corners_ix
                        const float wk = wxy[k & 3] * ((k & 4) ? fz1 : fz0);
                        acc[k][0] += wk * dv.x;
                        acc[k][1] += wk * dv.y;
                        acc[k][2] += wk * dv.z;
                        acc[k][3] += wk * dv.w;
                    }
                }
                s0 = s1;
            }
            if (cell >= 0) flush();
        }
        const int ch0 = 4 * (int)blockIdx.y;
        __syncthreads();
        const bool split = p.chunk_off2[slot + 1] - p.chunk_off2[slot] > 1;
        float* dst       = split ? partial_grids + c * gsize : grad_grid + (int64_t)slot * gsize;
        const double inv = 1.0 / scale;
        for (int32_t v = threadIdx.x; v < g.G * g.G * g.G; v += blockDim.x)
        {
            float4 o;
            o.x = (float)((double)(long long)sq[0 * nv + v] * inv);
            o.y = (float)((double)(long long)sq[1 * nv + v] * inv);
            o.z = (float)((double)(long long)sq[2 * nv + v] * inv);
            o.w = (float)((double)(long long)sq[3 * nv + v] * inv);
            *reinterpret_cast<float4*>(dst + (int64_t)v * C + ch0) = o;
        }
        return;
      }
    }
    if (s < send)
    {
        const int32_t node = g.seg_node[p.perm[q0]];
        float mn[3], mx[3];
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            mn[d] = g.pmin[3 * node + d];
            mx[d] = g.pmax[3 * node + d];
        }
        int32_t q    = find_segment(p.nm_soff, q0, q1, base + s);
        int64_t qbeg = p.nm_soff[q] - base;
        int64_t qend = p.nm_soff[q + 1] - base;
        SegCtx sc    = seg_ctx(g, p.perm[q]);
        for (; s < send; ++s)
        {
            while (s >= qend)
            {
                ++q;
                qbeg = qend;
                qend = p.nm_soff[q + 1] - base;
                if (s < qend) sc = seg_ctx(g, p.perm[q]);
            }
            float uv[3];
            seg_uv(g, sc, mn, mx, (int32_t)(s - qbeg), uv);
            const Corners cr = corners_cl(uv, g.G, C);
            const float dv   = dfeat[(base + s) * C + ch];
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) atomicAdd(sg + cr.off[kk] + ch, cr.w[kk] * dv);
        }
    }
    __syncthreads();
    const bool split = p.chunk_off2[slot + 1] - p.chunk_off2[slot] > 1;
    float* dst       = split ? partial_grids + c * gsize : grad_grid + (int64_t)slot * gsize;
    for (int32_t i = threadIdx.x * 4; i < gsize; i += blockDim.x * 4)
        *reinterpret_cast<float4*>(dst + i) = *reinterpret_cast<const float4*>(sg + i);
}

// The BF16 form's grid-gradient scatter: all 8 channels in one pass, in 32-bit fixed point, from the BF16 backward's
// 24 B per sample (BF16 dL/dfeature [S][8] and the packed cell coordinate, pack_cell_coord). A chunk of n samples
// adds at most n |dL/dfeature| (trilinear weights are at most 1 and sum to 1 per sample) to any vertex, so with the
// scale 2^F, F = 30 - ceil(log2 n) - e (|dL/dfeature| <= 2^e over the slot's samples, dmax[1 + slot]: below 2^e in
// FP32, which the BF16 rounding lifts to 2^e at most), every vertex sum (at most 2^30 plus the flushes' roundings)
// stays inside int32. The quantum is 2^-F, below 2^-29 n max|dL/dfeature|: with scatter chunks of at most 65536
// samples, under 2^-13 of the slot's largest value. (F is clamped to +-100, far outside what float gradients reach,
// so that 2^F stays finite.) Integer adds are exact and order-independent, so the gradient stays bit-reproducible.
// Lanes are sample streams as in k_proj_bw2; a lane keeps the current cell's 8 corners x 8 channels in registers,
// prescaled, and converts them once per cell (round to nearest).
constexpr int kScatterThreadsI32 = 512;  // 256 VGPRs per lane: the cell sums and an 8-sample block in registers

__global__ void __launch_bounds__(kScatterThreadsI32) k_proj_bw2_i32(ProjGeom g, Plan p, const float* dfeat,
                                                                  const float4* sgc, const unsigned int* dmax,
                                                                  float* grad_grid, float* partial_grids)
{
    extern __shared__ float sgf[];
    int32_t* sg        = reinterpret_cast<int32_t*>(sgf);  // planar [channel][vertex]
    const int64_t c    = blockIdx.x;
    if (c >= p.chunk_off2[g.N]) return;
    const int32_t slot = p.chunk2_slot[c];
    const int32_t q0 = p.chunk2_q0[c], q1 = p.chunk2_q1[c];
    const int32_t G = g.G, nv = G * G * G;
    for (int32_t i = threadIdx.x * 4; i < 8 * nv; i += blockDim.x * 4)
        *reinterpret_cast<int4*>(sg + i) = make_int4(0, 0, 0, 0);
    __syncthreads();
    const int64_t base = p.nm_soff[q0];
    const int64_t ns   = p.nm_soff[q1] - base;
    // the chunk's scale, from its slot's largest |dL/dfeature|; F stays inside float's exponent range
    int e;
    frexpf(__uint_as_float(dmax[1 + slot]), &e);
    int lg = 0;
    while (((int64_t)1 << lg) < max<int64_t>(ns, 1)) ++lg;
    const int F       = min(100, max(-100, 30 - lg - e));
    const float scale = ldexpf(1.0f, F);
    const int64_t L2 = (ns + blockDim.x - 1) / blockDim.x;
    const int64_t s2 = (int64_t)threadIdx.x * L2;
    const int64_t e2 = min(ns, s2 + L2);
    const uint2* pc = reinterpret_cast<const uint2*>(sgc) + base;  // packed cell coordinates
    const uint4* dq = reinterpret_cast<const uint4*>(dfeat) + base;  // dL/dfeature, 8 BF16 channels
    if (s2 < e2)
    {
        int cell = -1;
        int cx0 = 0, cx1 = 0, cy0 = 0, cy1 = 0, cz0 = 0, cz1 = 0;
        float acc[8][8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
            for (int q = 0; q < 8; ++q) acc[k][q] = 0.f;
        auto flush = [&]() {
            const int vz[2] = {cz0 * G * G, cz1 * G * G}, vy[2] = {cy0 * G, cy1 * G}, vx[2] = {cx0, cx1};
#pragma unroll
            for (int k = 0; k < 8; ++k)
            {
                const int v = vz[k >> 2] + vy[(k >> 1) & 1] + vx[k & 1];
#pragma unroll
                for (int q = 0; q < 8; ++q)
                {
                    atomicAdd(sg + q * nv + v, (int32_t)__builtin_rintf(acc[k][q]));
                    acc[k][q] = 0.f;
                }
            }
        };
        // blocks of 8 samples: 64 B of cell coordinates (half a cache line, the other half is the stream's next
        // block) and 128 B of dL/dfeature per stream and block, once the stream is aligned to them (its first block
        // runs to the next 8-sample boundary)
        constexpr int kSB  = 8;
        constexpr float kQ = 1.f / 65536.f;
        for (int64_t s0 = s2; s0 < e2;)
        {
            const int64_t s1 = min(e2, ((base + s0) / kSB + 1) * kSB - base);
            uint2 cb[kSB];
            uint4 db[kSB];
#pragma unroll
            for (int b = 0; b < kSB; ++b)
            {
                if (s0 + b < s1)
                {
                    cb[b] = pc[s0 + b];
                    db[b] = dq[s0 + b];
                }
            }
#pragma unroll
            for (int b = 0; b < kSB; ++b)
            {
                if (s0 + b >= s1) break;
                const uint2 u = cb[b];
                const int cc  = (int)(u.y >> 16);
                if (cc != cell)
                {
                    if (cell >= 0) flush();
                    cell = cc;
                    const int x0 = (cc & 31) - 1, y0 = ((cc >> 5) & 31) - 1, z0 = (cc >> 10) - 1;
                    cx0 = clamp_g(x0, G);
                    cx1 = clamp_g(x0 + 1, G);
                    cy0 = clamp_g(y0, G);
                    cy1 = clamp_g(y0 + 1, G);
                    cz0 = clamp_g(z0, G);
                    cz1 = clamp_g(z0 + 1, G);
                }
                const float fx1 = (float)(u.x & 0xFFFFu) * kQ, fx0 = 1.f - fx1;
                const float fy1 = (float)(u.x >> 16) * kQ, fy0 = 1.f - fy1;
                const float fz1 = (float)(u.y & 0xFFFFu) * kQ, fz0 = 1.f - fz1;
                const float wxy[4] = {fx0 * fy0, fx1 * fy0, fx0 * fy1, fx1 * fy1};
                float dv[8];
                bf16x8_floats(db[b], dv);
#pragma unroll
                for (int q = 0; q < 8; ++q) dv[q] *= scale;
#pragma unroll
                for (int k = 0; k < 8; ++k)
                {
                    const float wk = wxy[k & 3] * ((k & 4) ? fz1 : fz0);
#pragma unroll
                    for (int q = 0; q < 8; ++q) acc[k][q] = fmaf(wk, dv[q], acc[k][q]);
                }
            }
            s0 = s1;
        }
        if (cell >= 0) flush();
    }
    __syncthreads();
    const bool split = p.chunk_off2[slot + 1] - p.chunk_off2[slot] > 1;
    float* dst       = split ? partial_grids + c * (int64_t)nv * 8 : grad_grid + (int64_t)slot * nv * 8;
    const float inv  = ldexpf(1.0f, -F);
    for (int32_t v = threadIdx.x; v < nv; v += blockDim.x)
    {
        float4 a, b;
        a.x = (float)sg[0 * nv + v] * inv;
        a.y = (float)sg[1 * nv + v] * inv;
        a.z = (float)sg[2 * nv + v] * inv;
        a.w = (float)sg[3 * nv + v] * inv;
        b.x = (float)sg[4 * nv + v] * inv;
        b.y = (float)sg[5 * nv + v] * inv;
        b.z = (float)sg[6 * nv + v] * inv;
        b.w = (float)sg[7 * nv + v] * inv;
        *reinterpret_cast<float4*>(dst + (int64_t)v * 8)     = a;
        *reinterpret_cast<float4*>(dst + (int64_t)v * 8 + 4) = b;
    }
}

// Slots cut into several scatter chunks: sum the chunk grids in chunk order; slots without samples: zero. The plan
// lists exactly these slots (redo_slots, in any order: each output element is one thread's sum in chunk order, so the
// result does not depend on it), and a fixed number of workgroups walks the list's (slot, 1024-element tile) items.
constexpr int kReduceWG = 1024;
__global__ void __launch_bounds__(256) k_proj_bw2_reduce(ProjGeom g, Plan p, int32_t gsize, const float* partial_grids,
                                                         float* grad_grid)
{
    const int32_t tiles = (gsize + 1023) / 1024;
    const int64_t items = (int64_t)*p.nredo * tiles;
    for (int64_t item = blockIdx.x; item < items; item += gridDim.x)
    {
        const int64_t slot = p.redo_slots[item / tiles];
        const int32_t i    = (int32_t)(item % tiles) * 1024 + 4 * (int32_t)threadIdx.x;
        if (i >= gsize) continue;
        const int64_t c0 = p.chunk_off2[slot], c1 = p.chunk_off2[slot + 1];
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int64_t c = c0; c < c1; ++c)
        {
            const float4 v = *reinterpret_cast<const float4*>(partial_grids + c * gsize + i);
            acc.x += v.x;
            acc.y += v.y;
            acc.z += v.z;
            acc.w += v.w;
        }
        *reinterpret_cast<float4*>(grad_grid + slot * gsize + i) = acc;
    }
}

// The slots the reduction visits: those with no scatter chunk or with several.
__global__ void __launch_bounds__(256) k_plan_redo(ProjGeom g, Plan p)
{
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= g.N) return;
    if (p.chunk_off2[slot + 1] - p.chunk_off2[slot] != 1) p.redo_slots[atomicAdd(p.nredo, 1)] = (int32_t)slot;
}

__global__ void __launch_bounds__(256) k_fill_f32(float* p, int64_t n, float v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Forward: persistent workgroups of 4 waves over the chunks; tiles of 32 samples per wave on the matrix cores;
// per-segment sums sequential in j as in k_proj_fwd.
template <typename P, int H, int ACT>
__global__ void __launch_bounds__(256, 2) k_mfma_fwd(ProjGeom g, Plan p, const float* __restrict__ params,
                                                  const void* grid_lowp, float* seg_partial, float* sample_out)
{
    __shared__ WeightLDS<P, H> w;
    __shared__ ChunkCtx cx;
    __shared__ float s_contrib[kFwdChunk];
    __shared__ float s_seg[kSegCap];
    load_weights<P, H, 8>(params, w);
    const int l = threadIdx.x & 63, wave = threadIdx.x >> 6, n = l & 31, hg = l >> 5;
    const int64_t nchunks = p.chunk_off1[g.N];
    for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x)
    {
        const int32_t slot = p.chunk1_slot[c];
        const int32_t q0 = p.chunk1_q0[c], q1 = p.chunk1_q1[c], nseg = q1 - q0;
        if (nseg == 0) continue;
        const int64_t base = p.nm_soff[q0];
        __syncthreads();  // the previous chunk's readers of cx / s_seg are done
        load_chunk_ctx(g, p, q0, q1, base, nullptr, cx);
        for (int32_t t = threadIdx.x; t < nseg; t += blockDim.x) s_seg[t] = 0.f;
        NodeBox nb;
        node_bounds(g, g.seg_node[p.perm[q0]], nb);
        const void* ngrid = node_grid_ptr(g, grid_lowp, P::kBF16, slot);
        __syncthreads();
        const int32_t nsamp = cx.beg[nseg];
        for (int32_t w0 = 0; w0 < nsamp; w0 += kFwdChunk)
        {
            const int32_t wn = min(kFwdChunk, nsamp - w0);
            expand_window(cx, nseg, w0, wn);
            __syncthreads();
            // pairs of tiles per wave (fetch_pair)
            auto tile = [&](int32_t i, const float (&x4)[4], float wgt) {
                Fwd<H / 32> f;
                tile_forward_mfma<P, H, ACT>(w, x4, g.beta, f);
                if (i < wn && hg == 0)
                {
                    const float y = act_f(NEAT_ACT_SOFTPLUS, f.z, g.beta);
                    s_contrib[i]  = y * wgt;
                    if (sample_out) sample_out[base + w0 + i] = y;
                }
            };
            for (int32_t p0 = wave * 64; p0 < wn; p0 += 256)
            {
                PairIn cur;
                fetch_pair<P::kBF16, false>(g, cx, p0, w0, wn, base, nb, ngrid, nullptr, cur);
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

// Backward: persistent workgroups of 4 waves; per wave and tile of 32 samples, recompute the forward, back-propagate
// on the matrix cores, write dL/dfeature per sample (for the LDS scatter), and accumulate the parameter gradient
// with products over the tile's samples (staged through LDS as [row][sample]). Each wave writes its own slice of the
// partials (summed later in a fixed order).
template <typename P, int H, int ACT>
__global__ void __launch_bounds__(256) k_mfma_bwd(ProjGeom g, Plan p, const float* __restrict__ params,
                                                  const void* grid_lowp, const float* grad_proj, float* dfeat,
                                                  float4* sgc, unsigned int* dmax, float* partials, int32_t P_total)
{
    constexpr int MT = H / 32, C = 8;
    __shared__ WeightLDS<P, H> w;
    __shared__ ChunkCtx cx;
    typedef Stage<P> ST;
    __shared__ __align__(16) typename ST::E stage[4][ST::elems(H)];
    load_weights<P, H, C>(params, w);
    const int l = threadIdx.x & 63, wave = threadIdx.x >> 6, hg = l >> 5;
    BwdStage<P, H> sb(stage[wave]);
    BwdAcc<MT> acc;
    float amax = 0.f;  // largest |dL/dfeature| this lane wrote
    const int64_t nchunks = p.chunk_off1[g.N];
    // the chunk's largest |dL/dfeature| into its slot's entry (dmax[1 + slot])
    auto slot_max = [&](float m, int32_t slot) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
        if (l == 0) atomicMax(dmax + 1 + slot, __float_as_uint(m));
    };
    for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x)
    {
        const int32_t slot = p.chunk1_slot[c];
        const int32_t q0 = p.chunk1_q0[c], q1 = p.chunk1_q1[c], nseg = q1 - q0;
        if (nseg == 0) continue;
        const int64_t base = p.nm_soff[q0];
        __syncthreads();
        load_chunk_ctx(g, p, q0, q1, base, grad_proj, cx);
        NodeBox nb;
        node_bounds(g, g.seg_node[p.perm[q0]], nb);
        const void* ngrid = node_grid_ptr(g, grid_lowp, P::kBF16, slot);
        __syncthreads();
        const int32_t nsamp = cx.beg[nseg];
        float cmax = 0.f;  // this chunk's
        for (int32_t w0 = 0; w0 < nsamp; w0 += kFwdChunk)
        {
            const int32_t wn = min(kFwdChunk, nsamp - w0);
            expand_window(cx, nseg, w0, wn);
            __syncthreads();
            // pairs of tiles per wave (fetch_pair)
            auto tile = [&](int32_t i, const float (&x4)[4], float gy) {
                const bool valid = i < wn;
                const f32x16 dx = tile_backward_mfma<P, H, ACT>(w, x4, gy, valid, NEAT_ACT_SOFTPLUS, g.beta, sb, acc);
                if (valid)
                {
                    // FP32 form: [2][S][4] (channel halves); BF16 form: [S][8] BF16, this lane's 4 channels
                    if constexpr (P::kBF16)
                        *reinterpret_cast<uint2*>(reinterpret_cast<unsigned short*>(dfeat) + (base + w0 + i) * 8 +
                                                  4 * hg) = bf16x4_bits(dx[0], dx[1], dx[2], dx[3]);
                    else
                        *reinterpret_cast<float4*>(dfeat + ((int64_t)hg * g.S + base + w0 + i) * 4) =
                            make_float4(dx[0], dx[1], dx[2], dx[3]);
                    cmax = fmaxf(cmax, fmaxf(fmaxf(fabsf(dx[0]), fabsf(dx[1])), fmaxf(fabsf(dx[2]), fabsf(dx[3]))));
                }
            };
            for (int32_t p0 = wave * 64; p0 < wn; p0 += 256)
            {
                PairIn cur;
                fetch_pair<P::kBF16, true>(g, cx, p0, w0, wn, base, nb, ngrid, sgc, cur);
                // one copy of the tile code for both halves of the pair (smaller code, fewer live registers)
#pragma unroll 1
                for (int h = 0; h < 2; ++h)
                {
                    if (h == 1 && p0 + 32 >= wn) break;
                    float xs[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) xs[j] = h ? cur.xb[j] : cur.xa[j];
                    tile(p0 + 32 * h + (l & 31), xs, h ? cur.vb : cur.va);
                }
            }
            __syncthreads();
        }
        slot_max(cmax, slot);
        amax = fmaxf(amax, cmax);
    }
    // this wave's slice of the partial parameter gradient (packed layout)
    write_param_partials<H, C>(acc, partials + ((int64_t)blockIdx.x * 4 + wave) * P_total);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) amax = fmaxf(amax, __shfl_xor(amax, off));
    if (l == 0) atomicMax(dmax, __float_as_uint(amax));  // non-negative floats order like their bit patterns
}

// ---- workspace layout ---------------------------------------------------------------------------------------------------
struct Layout
{
    size_t off[24];
    size_t total;
    int n = 0;
    size_t add(size_t bytes)
    {
        size_t o  = total;
        off[n++]  = o;
        total     = o + ((bytes + 255) / 256) * 256;
        return o;
    }
};

struct ProjSizes
{
    int64_t K, S, N;
    int32_t G, C, P;
    int64_t max1, max2;
    int32_t chunk2;
    size_t sort_temp;
    bool prop;  // the segment list carries proposal CDF rows
};

int32_t scatter_chunk(int64_t S)
{
    // aim for >= ~1024 scatter workgroups, but keep chunks between 2048 and 65536 samples
    int64_t t = S / 1024;
    return (int32_t)std::max<int64_t>(2048, std::min<int64_t>(65536, t));
}

ProjSizes proj_sizes(int64_t K, int64_t S, int64_t N, int32_t G, int32_t C, int32_t P, bool prop)
{
    ProjSizes z{K, S, N, G, C, P, 0, 0, scatter_chunk(S), 0, prop};
    z.max1 = (S + (int64_t)kSegCost * K + kFwdChunk - 1) / kFwdChunk + N;
    z.max2 = (S + z.chunk2 - 1) / z.chunk2 + N;
    size_t tmp = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (int32_t*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr,
                                       (int32_t*)nullptr, (int)std::max<int64_t>(K, 1), 0, 32);
    z.sort_temp = tmp;
    return z;
}

// plan arrays, then forward / backward scratch
Layout proj_layout(const ProjSizes& z)
{
    Layout L{};
    L.add(4 * z.K);                       // 0 perm
    L.add(8 * (z.K + 1));                 // 1 nm_soff
    L.add(4 * (z.N + 1));                 // 2 slot_begin
    L.add(8 * (z.N + 1));                 // 3 chunk_off1
    L.add(8 * (z.N + 1));                 // 4 chunk_off2
    L.add(4 * z.max1 * 3);                // 5 chunk1 slot/q0/q1
    L.add(4 * z.max2 * 3);                // 6 chunk2 slot/q0/q1
    L.add(4 * z.K);                       // 7 keys (sort in)
    L.add(4 * z.K);                       // 8 keys (sorted)
    L.add(4 * z.K);                       // 9 vals (sort in)
    L.add(std::max<size_t>(z.sort_temp, scan_workspace_bytes(z.K + 1)));  // 10 sort / scan temp
    L.add(4 * z.K);                       // 11 seg_partial (forward)
    L.add(4 * (size_t)z.P);               // 12 transposed decoder params
    L.add(4 * (size_t)z.S * z.C);         // 13 dfeat ([S][C]; FP32 matrix-core: [C/4][S][4]; BF16: [S][8] BF16)
    L.add(16 * (size_t)z.S);              // 14 grid coordinates per sample, matrix-core backward (BF16: packed, 8 B)
    L.add(4 * (size_t)(z.N + 1));         // 15 max |dL/dfeature| (float bits) of the batch, then of each slot
    L.add(4 * (kNodeWGMax + 1));          // 16 first chunk of each node-resident workgroup
    L.add(64 * (size_t)z.K);              // 17 node-major segment records (ot, dt, i, h)
    L.add(z.prop ? 32 * (size_t)z.K : 16);  // 18 node-major proposal cells / CDF rows
    L.add(24 * (size_t)z.K);              // 19 per-segment (dL/do, dL/dd) of the positional gradient
    L.add(4 * (size_t)(z.N + 1));         // 20 slots the scatter's reduction visits, and their count
    return L;
}

// Node-resident workgroups: kNodeWGPerCU per CU of the current device (at most kNodeWGMax).
int node_workgroups()
{
    int dev = 0, cu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                hipSuccess || cu <= 0)
        cu = 256;
    return std::min(cu * kNodeWGPerCU, kNodeWGMax);
}

int bw1_workgroups(const ProjSizes& z) { return (int)std::max<int64_t>(1, std::min<int64_t>(z.max1, 2048)); }
// matrix-core backward: persistent workgroups of 4 waves, one partial slice per wave
int mfma_bwd_workgroups(const ProjSizes& z) { return (int)std::max<int64_t>(1, std::min<int64_t>(z.max1, 256)); }
int mfma_fwd_workgroups(const ProjSizes& z) { return (int)std::max<int64_t>(1, std::min<int64_t>(z.max1, 1024)); }

size_t partials_bytes(const ProjSizes& z)
{
    size_t slices = std::max<size_t>({(size_t)bw1_workgroups(z), (size_t)mfma_bwd_workgroups(z) * 4,
                                      (size_t)kNodeWGMax * kNodeWaves});
    return (slices * z.P * 4 + 255) / 256 * 256;
}

size_t proj_bytes(const ProjSizes& z)
{
    Layout L     = proj_layout(z);
    size_t grids = (size_t)z.max2 * z.G * z.G * z.G * z.C * 4;
    return L.total + partials_bytes(z) + grids + 256;
}

Plan plan_view(void* ws, const ProjSizes& z)
{
    Layout L = proj_layout(z);
    char* b  = static_cast<char*>(ws);
    Plan p;
    p.perm        = reinterpret_cast<int32_t*>(b + L.off[0]);
    p.nm_soff     = reinterpret_cast<int64_t*>(b + L.off[1]);
    p.slot_begin  = reinterpret_cast<int32_t*>(b + L.off[2]);
    p.chunk_off1  = reinterpret_cast<int64_t*>(b + L.off[3]);
    p.chunk_off2  = reinterpret_cast<int64_t*>(b + L.off[4]);
    int32_t* c1   = reinterpret_cast<int32_t*>(b + L.off[5]);
    int32_t* c2   = reinterpret_cast<int32_t*>(b + L.off[6]);
    p.chunk1_slot = c1;
    p.chunk1_q0   = c1 + z.max1;
    p.chunk1_q1   = c1 + 2 * z.max1;
    p.chunk2_slot = c2;
    p.chunk2_q0   = c2 + z.max2;
    p.chunk2_q1   = c2 + 2 * z.max2;
    p.max1        = z.max1;
    p.max2        = z.max2;
    p.chunk2      = z.chunk2;
    p.wg_first    = reinterpret_cast<int32_t*>(b + L.off[16]);
    p.nwg         = node_workgroups();
    char* rec     = b + L.off[17];
    p.rec_ot      = reinterpret_cast<float4*>(rec);
    p.rec_dt      = reinterpret_cast<float4*>(rec + 16 * z.K);
    p.rec_i       = reinterpret_cast<int4*>(rec + 32 * z.K);
    p.rec_h       = reinterpret_cast<uint4*>(rec + 48 * z.K);
    p.rec_c       = z.prop ? reinterpret_cast<float4*>(b + L.off[18]) : nullptr;
    p.redo_slots  = reinterpret_cast<int32_t*>(b + L.off[20]);
    p.nredo       = p.redo_slots + z.N;
    // region 13 holds 32 B per sample with 8 channels: the BF16 dL/dfeature in the first half, the feature cache in
    // the second; region 14 (16 B per sample): the packed cell coordinates, then the weights
    const bool c8 = z.C == 8;
    p.feat_cache  = c8 ? reinterpret_cast<uint4*>(b + L.off[13] + 16 * (size_t)z.S) : nullptr;
    p.w_cache     = c8 ? reinterpret_cast<float*>(b + L.off[14] + 8 * (size_t)z.S) : nullptr;
    p.cell_cache  = c8 ? reinterpret_cast<uint2*>(b + L.off[14]) : nullptr;
    return p;
}

}  // namespace neat

using namespace neat;

namespace
{

struct Args
{
    ProjGeom g;
    MLP m;
    int act;
    const float* params;
    ProjSizes z;
    int precision;
    const void* grid_bf16;
    bool mfma;    // matrix-core kernels (8 channels) rather than the vector-ALU ones
    bool relu16;  // BF16 relu decoder on the node-resident kernels (k_mfma_fwd_relu / k_mfma_bwd_relu)
};

int make_args(Args& a, const neat_projector_args* x)
{
    NEAT_REQUIRE(x, "projector: null argument struct");
    const int64_t R = x->num_rays, K = x->num_segments, S = x->num_samples, N = x->num_slots;
    const int32_t C = x->channels, G = x->grid_size;
    NEAT_REQUIRE(R >= 0 && K >= 0 && S >= 0 && N > 0 && x->num_nodes > 0 && G >= 2 && C > 0,
                 "projector: bad sizes");
    NEAT_REQUIRE(x->node_pos_min && x->node_pos_max && x->node_active_prefix_sum && x->grid && x->params,
                 "projector: null pointer");
    NEAT_REQUIRE(R == 0 || (x->ray_origin && x->ray_direction && x->ray_seg_offset), "projector: null ray arrays");
    NEAT_REQUIRE(K == 0 || (x->seg_node && x->seg_ray && x->seg_count && x->seg_tmin && x->seg_tmax),
                 "projector: null segment arrays");
    NEAT_REQUIRE(K < INT32_MAX && S < ((int64_t)1 << 40), "projector: too many segments");
    NEAT_REQUIRE(x->precision == NEAT_PREC_FP32 || x->precision == NEAT_PREC_BF16 ||
                     x->precision == NEAT_PREC_FP32_VALU,
                 "projector: unknown precision %d", x->precision);
    const int64_t* st = x->grid_strides;
    const int64_t gs  = (int64_t)G * G * G * C;
    if (!(st[1] == 1 && st[4] == C && st[3] == (int64_t)G * C && st[2] == (int64_t)G * G * C && st[0] == gs))
        return fail(NEAT_E_UNSUPPORTED, "projector: the grid must be cubic and channels-last contiguous");
    if (!(x->in_features == C && (C == 4 || C == 8) && (x->hidden == 32 || x->hidden == 64) &&
          x->hidden_layers == 1 && x->out_features == 1 && (x->act == NEAT_ACT_RELU || x->act == NEAT_ACT_SILU) &&
          x->last_act == NEAT_ACT_SOFTPLUS))
        return fail(NEAT_E_UNSUPPORTED,
                    "projector: fused kernels cover in=C in {4,8}, hidden {32,64}, one hidden layer, out 1, "
                    "relu/silu hidden activations and a softplus output (got in=%d C=%d hidden=%d layers=%d out=%d "
                    "act=%d last=%d)",
                    x->in_features, C, x->hidden, x->hidden_layers, x->out_features, x->act, x->last_act);
    if ((int64_t)G * G * G * C * 4 > 160 * 1024)
        return fail(NEAT_E_UNSUPPORTED, "projector: a node grid (%d^3 x %d FP32) exceeds the 160 KiB of LDS the "
                    "gradient scatter holds it in", G, C);
    if (x->precision == NEAT_PREC_BF16)
    {
        if (C != 8) return fail(NEAT_E_UNSUPPORTED, "projector: the BF16 kernels need 8 channels (got %d)", C);
        NEAT_REQUIRE(x->grid_bf16 && (reinterpret_cast<uintptr_t>(x->grid_bf16) % 16) == 0,
                     "projector: NEAT_PREC_BF16 needs a 16-byte aligned grid_bf16");
    }
    NEAT_REQUIRE((reinterpret_cast<uintptr_t>(x->grid) % 16) == 0, "projector: grid must be 16-byte aligned");
    NEAT_REQUIRE((reinterpret_cast<uintptr_t>(x->seg_cdf) % 16) == 0, "projector: seg_cdf must be 16-byte aligned");
    a.g = ProjGeom{x->node_pos_min, x->node_pos_max, x->node_active_prefix_sum, x->num_nodes, x->ray_origin,
                   x->ray_direction, R, x->ray_seg_offset, K, x->seg_node, x->seg_ray, x->seg_count, x->seg_tmin,
                   x->seg_tmax, x->jitter, x->seed, x->grid, N, G, x->softplus_beta, x->seg_cdf, S};
    a.m         = MLP{x->in_features, x->hidden, x->hidden_layers, x->out_features};
    a.act       = x->act;
    a.params    = x->params;
    a.z         = proj_sizes(K, S, N, G, C, a.m.nparams(), x->seg_cdf != nullptr);
    a.precision = x->precision;
    a.grid_bf16 = x->grid_bf16;
    a.mfma      = x->precision != NEAT_PREC_FP32_VALU && C == 8;
    // the node-resident relu kernels stage a node's BF16 grid in LDS (up to 17^3 vertices)
    a.relu16    = a.mfma && x->precision == NEAT_PREC_BF16 && x->act == NEAT_ACT_RELU &&
               (int64_t)G * G * G <= kStagedVerts;
    return NEAT_OK;
}

int check_ws(const Args& a, void* ws, size_t bytes)
{
    size_t need = proj_bytes(a.z);
    if (!ws || bytes < need) return fail(NEAT_E_OVERFLOW, "projector workspace too small: %zu < %zu", bytes, need);
    if (reinterpret_cast<uintptr_t>(ws) % 256) return fail(NEAT_E_ARG, "projector workspace must be 256-byte aligned");
    return NEAT_OK;
}

char* region(void* ws, const ProjSizes& z, int i) { return static_cast<char*>(ws) + proj_layout(z).off[i]; }

template <int C, int H>
int launch_fwd_valu(hipStream_t s, const Args& a, const Plan& p, float* seg_partial)
{
    unsigned grid = (unsigned)std::max<int64_t>(1, a.z.max1);
    if (a.act == NEAT_ACT_RELU)
        k_proj_fwd<C, H, NEAT_ACT_RELU><<<grid, kFwdThreads, 0, s>>>(a.g, p, a.params, seg_partial);
    else
        k_proj_fwd<C, H, NEAT_ACT_SILU><<<grid, kFwdThreads, 0, s>>>(a.g, p, a.params, seg_partial);
    NEAT_LAUNCHED();
    return NEAT_OK;
}

template <typename P, int H>
int launch_fwd_mfma(hipStream_t s, const Args& a, const Plan& p, float* seg_partial, float* sample_out = nullptr)
{
    const unsigned grid = (unsigned)mfma_fwd_workgroups(a.z);
    if (P::kBF16 && a.relu16)
    {
        NEAT_REQUIRE(launch_mfma_fwd_relu(s, H, a.g, p, a.params, a.grid_bf16, seg_partial, sample_out) == 0,
                     "k_mfma_fwd_relu launch failed");
        return NEAT_OK;
    }
    if (a.act == NEAT_ACT_RELU)
        k_mfma_fwd<P, H, NEAT_ACT_RELU><<<grid, 256, 0, s>>>(a.g, p, a.params, a.grid_bf16, seg_partial, sample_out);
    else
        k_mfma_fwd<P, H, NEAT_ACT_SILU><<<grid, 256, 0, s>>>(a.g, p, a.params, a.grid_bf16, seg_partial, sample_out);
    NEAT_LAUNCHED();
    return NEAT_OK;
}

// Scatter of dL/dfeature into the node gradient grids (pass 2), shared by both backward variants.
template <int C>
int launch_scatter(hipStream_t s, const Args& a, const Plan& p, const float* dfeat, const float4* sgc,
                   const unsigned int* dmax, float* grid_partials, float* grad_grid)
{
    const int32_t gsize = a.g.G * a.g.G * a.g.G * C;
    const size_t smem   = (size_t)gsize * sizeof(float);
    if (C == 8 && sgc && a.precision == NEAT_PREC_BF16)
    {
        // the BF16 form: one pass over all 8 channels in 32-bit fixed point
        NEAT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_proj_bw2_i32),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        k_proj_bw2_i32<<<(unsigned)std::max<int64_t>(1, a.z.max2), kScatterThreadsI32, smem, s>>>(
            a.g, p, dfeat, sgc, dmax, grad_grid, grid_partials);
        NEAT_LAUNCHED();
        k_proj_bw2_reduce<<<kReduceWG, 256, 0, s>>>(a.g, p, gsize, grid_partials, grad_grid);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }
    NEAT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_proj_bw2<C>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    // the fixed-point form holds half the channels per workgroup as int64 (the same bytes as all of them in FP32)
    const dim3 grid((unsigned)std::max<int64_t>(1, a.z.max2), sgc ? C / 4 : 1);
    k_proj_bw2<C><<<grid, kScatterThreads, smem, s>>>(a.g, p, dfeat, sgc, dmax, grad_grid, grid_partials);
    NEAT_LAUNCHED();
    k_proj_bw2_reduce<<<kReduceWG, 256, 0, s>>>(a.g, p, gsize, grid_partials, grad_grid);
    NEAT_LAUNCHED();
    return NEAT_OK;
}

// The backward's stages: 1 the decoder backward (dL/dfeature per sample, per-workgroup parameter partials),
// 2 the parameter gradient (the partials summed in a fixed order), 3 the grid gradient (the scatter).
struct BwdBufs
{
    const float* grad_proj;
    float* wt;
    float* dfeat;
    float4* sgc;
    unsigned int* dmax;
    float* partials;
    float* grid_partials;
    float* grad_grid;
    float* grad_params;
};

template <int C, int H>
int launch_bw_valu(hipStream_t s, const Args& a, const Plan& p, const BwdBufs& b, int stage)
{
    const int nwg = bw1_workgroups(a.z);
    const int P   = a.m.nparams();
    if (stage == 1)
    {
        k_transpose_params<<<grid_for(P, 256), 256, 0, s>>>(a.params, a.m, b.wt);
        NEAT_LAUNCHED();
        NEAT_HIP(hipMemsetAsync(b.partials, 0, (size_t)nwg * P * sizeof(float), s));
        k_proj_bw1<C, H><<<nwg, kTile, 0, s>>>(a.g, p, a.m, b.wt, a.act, b.grad_proj, b.dfeat, b.partials);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }
    if (stage == 2)
    {
        k_dec_reduce_partials<<<reduce_partials_blocks(P), 256, 0, s>>>(b.partials, nwg, P, b.grad_params);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }
    return launch_scatter<C>(s, a, p, b.dfeat, nullptr, nullptr, b.grid_partials, b.grad_grid);
}

template <typename P, int H>
int launch_bw_mfma(hipStream_t s, const Args& a, const Plan& p, const BwdBufs& b, int stage)
{
    // BF16 with relu hidden layers runs k_mfma_bwd_relu (node-resident: p.nwg workgroups of kNodeWaves waves)
    const bool relu16 = P::kBF16 && a.relu16;
    const int slices  = relu16 ? p.nwg * kNodeWaves : mfma_bwd_workgroups(a.z) * 4;
    const int Ptot    = a.m.nparams();
    if (stage == 1)
    {
        k_fill_f32<<<grid_for(a.g.N + 1, 256), 256, 0, s>>>(reinterpret_cast<float*>(b.dmax), a.g.N + 1, 0.f);
        NEAT_LAUNCHED();
        if (relu16)
        {
            // reads the per-sample caches neat_projector_fwd left in this workspace
            const int rc = launch_mfma_bwd_relu(s, H, a.g, p, a.params, b.grad_proj, b.dfeat, b.dmax, b.partials, Ptot);
            NEAT_REQUIRE(rc == 0, "k_mfma_bwd_relu launch failed");
            return NEAT_OK;
        }
        const int nwg = mfma_bwd_workgroups(a.z);
        if (a.act == NEAT_ACT_RELU)
            k_mfma_bwd<P, H, NEAT_ACT_RELU><<<nwg, 256, 0, s>>>(a.g, p, a.params, a.grid_bf16, b.grad_proj, b.dfeat,
                                                                 b.sgc, b.dmax, b.partials, Ptot);
        else
            k_mfma_bwd<P, H, NEAT_ACT_SILU><<<nwg, 256, 0, s>>>(a.g, p, a.params, a.grid_bf16, b.grad_proj, b.dfeat,
                                                                 b.sgc, b.dmax, b.partials, Ptot);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }
    if (stage == 2)
    {
        k_dec_reduce_partials<<<reduce_partials_blocks(Ptot), 256, 0, s>>>(b.partials, slices, Ptot, b.grad_params);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }
    return launch_scatter<8>(s, a, p, b.dfeat, b.sgc, b.dmax, b.grid_partials, b.grad_grid);
}

int dispatch_fwd(hipStream_t s, const Args& a, const Plan& p, float* seg_partial)
{
    const int C = a.m.in, H = a.m.hidden;
    if (a.mfma)
    {
        if (a.precision == NEAT_PREC_BF16)
            return H == 64 ? launch_fwd_mfma<PrecBF16, 64>(s, a, p, seg_partial)
                           : launch_fwd_mfma<PrecBF16, 32>(s, a, p, seg_partial);
        return H == 64 ? launch_fwd_mfma<PrecF32, 64>(s, a, p, seg_partial)
                       : launch_fwd_mfma<PrecF32, 32>(s, a, p, seg_partial);
    }
    if (C == 8) return H == 64 ? launch_fwd_valu<8, 64>(s, a, p, seg_partial) : launch_fwd_valu<8, 32>(s, a, p, seg_partial);
    return H == 64 ? launch_fwd_valu<4, 64>(s, a, p, seg_partial) : launch_fwd_valu<4, 32>(s, a, p, seg_partial);
}

int dispatch_bwd(hipStream_t s, const Args& a, const Plan& p, const BwdBufs& b, int stage)
{
    const int C = a.m.in, H = a.m.hidden;
    if (a.mfma)
    {
        if (a.precision == NEAT_PREC_BF16)
            return H == 64 ? launch_bw_mfma<PrecBF16, 64>(s, a, p, b, stage) : launch_bw_mfma<PrecBF16, 32>(s, a, p, b, stage);
        return H == 64 ? launch_bw_mfma<PrecF32, 64>(s, a, p, b, stage) : launch_bw_mfma<PrecF32, 32>(s, a, p, b, stage);
    }
    if (C == 8)
        return H == 64 ? launch_bw_valu<8, 64>(s, a, p, b, stage) : launch_bw_valu<8, 32>(s, a, p, b, stage);
    return H == 64 ? launch_bw_valu<4, 64>(s, a, p, b, stage) : launch_bw_valu<4, 32>(s, a, p, b, stage);
}

// ---- positional gradient (rays that require grad) -------------------------------------------------------------------
// Sample j of the segment at node-major position q, from the plan's record: t and the local coordinate with the
// operations of ctx_sample, so the same bits.
__device__ __forceinline__ void record_sample(const ProjGeom& g, const Plan& p, int64_t q, int32_t j,
                                              const NodeBox& nb, float& tt, float (&uv)[3])
{
    const float4 ot = p.rec_ot[q], dt = p.rec_dt[q];
    const int4 ri   = p.rec_i[q];
    const uint4 rh  = p.rec_h[q];
    const float tmin = ot.w, tmax = dt.w, step = __uint_as_float(rh.w);
    const int32_t rank = ri.z, n = ri.w;
    const uint64_t hr  = ((uint64_t)rh.y << 32) | rh.x;
    auto a_of          = [&](int32_t jj) { return g.jitter ? jitter_a_h(hr, rank, jj) : 0.5f; };
    if (rh.z != 0)
    {
        float cell[8];
        const float4 u = p.rec_c[2 * q], v = p.rec_c[2 * q + 1];
        cell[0] = u.x; cell[1] = u.y; cell[2] = u.z; cell[3] = u.w;
        cell[4] = v.x; cell[5] = v.y; cell[6] = v.z; cell[7] = v.w;
        float t1, t2;
        if (rh.z == 1)
        {
            t1 = j == 0 ? tmin : cell[j - 1];
            t2 = j + 1 >= n ? tmax : cell[j];
        }
        else
        {
            t1 = proposal_cell(cell, tmin, tmax, n, j);
            t2 = proposal_cell(cell, tmin, tmax, n, j + 1);
        }
        const float a = a_of(j);
        tt            = t1 * (1.0f - a) + t2 * a;
    }
    else
    {
        tt = sample_t(tmin, step, j, a_of(j));
    }
    const float o[3] = {ot.x, ot.y, ot.z}, d[3] = {dt.x, dt.y, dt.z};
#pragma unroll
    for (int e = 0; e < 3; ++e)
    {
        const float pd  = o[e] + d[e] * tt;
        const float rel = nb.rinv[e] != 0.f ? (pd - nb.mn[e]) * nb.rinv[e] : (pd - nb.mn[e]) / (nb.mx[e] - nb.mn[e]);
        uv[e]           = rel * 2 - 1;
    }
}

// One thread per node-major segment: for every sample, the uv gradient of the trilinear gather contracted with the
// backward's dL/dfeature (IndirectGridSample3D.cu:752-790, the terms in the reference's order), through the local
// coordinates' backward (grad * 2) / (max - min) onto the position o + d t, summed over the segment's samples in j
// order into (dL/do, dL/dd); written at the canonical segment index. Segments without a grid slot get zeros.
// dL/dfeature: the BF16 backward's [S][8] BF16 (kBF16); otherwise [C/4][S][4] (kHalfMajor: the FP32 matrix-core
// backward) or [S][C].
template <int C, bool kBF16, bool kHalfMajor>
__global__ void __launch_bounds__(256) k_proj_ray_grad(ProjGeom g, Plan p, const void* grid_lowp, const float* dfeat,
                                                       float* seg_grad)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.K) return;
    const int4 ri     = p.rec_i[q];
    const int32_t k   = ri.x, n = ri.w;
    const int32_t node = g.seg_node[k];
    const int64_t slot = (node >= 0 && node < g.num_nodes) ? g.prefix[node] : -1;
    float go[3] = {0.f, 0.f, 0.f}, gd[3] = {0.f, 0.f, 0.f};
    if (slot >= 0 && slot < g.N)
    {
        NodeBox nb;
        node_bounds(g, node, nb);
        const int32_t G     = g.G;
        const int64_t nvox  = (int64_t)G * G * G * C;
        const float* gf     = g.grid + slot * nvox;
        const unsigned short* gh = static_cast<const unsigned short*>(grid_lowp) + slot * nvox;
        const int64_t s0    = p.nm_soff[q];
        for (int32_t j = 0; j < n; ++j)
        {
            float tt, uv[3];
            record_sample(g, p, q, j, nb, tt, uv);
            const float ix = grid_coord(uv[0], G), iy = grid_coord(uv[1], G), iz = grid_coord(uv[2], G);
            const Corners cr = corners_ix(ix, iy, iz, G, C);
            const float x0 = floorf(ix), y0 = floorf(iy), z0 = floorf(iz);
            const float x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
            float gix = 0.f, giy = 0.f, giz = 0.f;
            const int64_t smp = s0 + j;
            // the 8 corner vertices, all channels of each in one (BF16) or C / 4 (FP32) vector loads
            static_assert(!kBF16 || C == 8, "the BF16 grid has 8 channels");
            float vv[8][C];
#pragma unroll
            for (int c8 = 0; c8 < 8; ++c8)
            {
                if constexpr (kBF16)
                {
                    const uint4 v = *reinterpret_cast<const uint4*>(gh + cr.off[c8]);
                    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                    for (int h = 0; h < 4; ++h)
                    {
                        vv[c8][2 * h]     = __uint_as_float(w[h] << 16);
                        vv[c8][2 * h + 1] = __uint_as_float(w[h] & 0xFFFF0000u);
                    }
                }
                else
                {
#pragma unroll
                    for (int h = 0; h < C / 4; ++h)
                    {
                        const float4 v = *reinterpret_cast<const float4*>(gf + cr.off[c8] + 4 * h);
                        vv[c8][4 * h]     = v.x;
                        vv[c8][4 * h + 1] = v.y;
                        vv[c8][4 * h + 2] = v.z;
                        vv[c8][4 * h + 3] = v.w;
                    }
                }
            }
            float gq[C];
            if constexpr (kBF16)
                bf16x8_floats(reinterpret_cast<const uint4*>(dfeat)[smp], gq);
            else
            {
#pragma unroll
                for (int ch = 0; ch < C; ++ch)
                    gq[ch] = kHalfMajor ? dfeat[((int64_t)(ch >> 2) * g.S + smp) * 4 + (ch & 3)] : dfeat[smp * C + ch];
            }
#pragma unroll
            for (int ch = 0; ch < C; ++ch)
            {
                const float gout = gq[ch];
                float val[8];
#pragma unroll
                for (int c8 = 0; c8 < 8; ++c8) val[c8] = vv[c8][ch];
                const float m1 = -1.f;
                gix += m1 * val[0] * (y1 - iy) * (z1 - iz) * gout;
                gix += val[1] * (y1 - iy) * (z1 - iz) * gout;
                gix += m1 * val[2] * (iy - y0) * (z1 - iz) * gout;
                gix += val[3] * (iy - y0) * (z1 - iz) * gout;
                gix += m1 * val[4] * (y1 - iy) * (iz - z0) * gout;
                gix += val[5] * (y1 - iy) * (iz - z0) * gout;
                gix += m1 * val[6] * (iy - y0) * (iz - z0) * gout;
                gix += val[7] * (iy - y0) * (iz - z0) * gout;

                giy += m1 * val[0] * (x1 - ix) * (z1 - iz) * gout;
                giy += m1 * val[1] * (ix - x0) * (z1 - iz) * gout;
                giy += val[2] * (x1 - ix) * (z1 - iz) * gout;
                giy += val[3] * (ix - x0) * (z1 - iz) * gout;
                giy += m1 * val[4] * (x1 - ix) * (iz - z0) * gout;
                giy += m1 * val[5] * (ix - x0) * (iz - z0) * gout;
                giy += val[6] * (x1 - ix) * (iz - z0) * gout;
                giy += val[7] * (ix - x0) * (iz - z0) * gout;

                giz += m1 * val[0] * (x1 - ix) * (y1 - iy) * gout;
                giz += m1 * val[1] * (ix - x0) * (y1 - iy) * gout;
                giz += m1 * val[2] * (x1 - ix) * (iy - y0) * gout;
                giz += m1 * val[3] * (ix - x0) * (iy - y0) * gout;
                giz += val[4] * (x1 - ix) * (y1 - iy) * gout;
                giz += val[5] * (ix - x0) * (y1 - iy) * gout;
                giz += val[6] * (x1 - ix) * (iy - y0) * gout;
                giz += val[7] * (ix - x0) * (iy - y0) * gout;
            }
            const float guv[3] = {gix * (float)(G - 1) / 2, giy * (float)(G - 1) / 2, giz * (float)(G - 1) / 2};
#pragma unroll
            for (int e = 0; e < 3; ++e)
            {
                const float gpos = (guv[e] * 2) / (nb.mx[e] - nb.mn[e]);
                go[e] += gpos;
                gd[e] += gpos * tt;
            }
        }
    }
#pragma unroll
    for (int e = 0; e < 3; ++e)
    {
        seg_grad[6 * (int64_t)k + e]     = go[e];
        seg_grad[6 * (int64_t)k + 3 + e] = gd[e];
    }
}

// dL/do, dL/dd of every ray: its segments' sums in canonical order.
__global__ void __launch_bounds__(256) k_ray_grad_sum(const int64_t* ray_seg_offset, int64_t R, const float* seg_grad,
                                                      float* grad_origin, float* grad_direction)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int64_t k = ray_seg_offset[r]; k < ray_seg_offset[r + 1]; ++k)
    {
#pragma unroll
        for (int e = 0; e < 6; ++e) acc[e] += seg_grad[6 * k + e];
    }
#pragma unroll
    for (int e = 0; e < 3; ++e)
    {
        if (grad_origin) grad_origin[3 * r + e] = acc[e];
        if (grad_direction) grad_direction[3 * r + e] = acc[3 + e];
    }
}

// first node-major sample of every canonical segment
__global__ void __launch_bounds__(256) k_seg_value_offset(const int32_t* perm, const int64_t* nm_soff, int64_t K,
                                                          int64_t* off)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < K) off[perm[q]] = nm_soff[q];
}

__global__ void __launch_bounds__(256) k_to_bf16(const float* src, int64_t n, unsigned short* dst)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (unsigned short)bf16_bits(src[i]);
}

}  // namespace

extern "C"
{
    size_t neat_projector_workspace_bytes(const neat_projector_args* x)
    {
        Args a;
        if (make_args(a, x) != NEAT_OK) return 0;
        return proj_bytes(a.z);
    }

    int neat_grid_to_bf16(neat_stream_t stream, const float* src, int64_t n, uint16_t* dst)
    {
        NEAT_REQUIRE(n >= 0 && (n == 0 || (src && dst)), "neat_grid_to_bf16: bad arguments");
        if (n == 0) return NEAT_OK;
        k_to_bf16<<<grid_for(n, 256), 256, 0, as_stream(stream)>>>(src, n, reinterpret_cast<unsigned short*>(dst));
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_projector_plan(neat_stream_t stream, const neat_projector_args* x, void* ws, size_t bytes)
    {
        Args a;
        int st = make_args(a, x);
        if (st) return st;
        if ((st = check_ws(a, ws, bytes))) return st;
        hipStream_t s   = as_stream(stream);
        const auto& z   = a.z;
        Plan p          = plan_view(ws, z);
        int32_t* keys   = reinterpret_cast<int32_t*>(region(ws, z, 7));
        int32_t* skeys  = reinterpret_cast<int32_t*>(region(ws, z, 8));
        int32_t* vals   = reinterpret_cast<int32_t*>(region(ws, z, 9));
        void* temp      = region(ws, z, 10);
        size_t temp_sz  = proj_layout(z).off[11] - proj_layout(z).off[10];
        if (z.K > 0)
        {
            k_plan_keys<<<grid_for(z.K, 256), 256, 0, s>>>(a.g, keys, vals);
            NEAT_LAUNCHED();
            int end_bit = 1;
            while (((int64_t)1 << end_bit) <= z.N) ++end_bit;
            size_t tmp = z.sort_temp;
            NEAT_HIP(hipcub::DeviceRadixSort::SortPairs(temp, tmp, keys, skeys, vals, p.perm, (int)z.K, 0, end_bit,
                                                        s));
            k_plan_records<<<grid_for(z.K, 256), 256, 0, s>>>(a.g, p);
            NEAT_LAUNCHED();
            k_plan_counts<<<grid_for(z.K, 256), 256, 0, s>>>(a.g, p.perm, p.nm_soff);
            NEAT_LAUNCHED();
        }
        if ((st = exclusive_scan_i64(s, p.nm_soff, z.K, temp, temp_sz))) return st;
        if (z.K > 0)
        {
            k_plan_slots<<<grid_for(z.N + 1, 256), 256, 0, s>>>(a.g, skeys, p);
            NEAT_LAUNCHED();
        }
        else
        {
            NEAT_HIP(hipMemsetAsync(p.slot_begin, 0, (size_t)(z.N + 1) * sizeof(int32_t), s));
        }
        k_plan_chunk_counts<<<grid_for(z.N, 256), 256, 0, s>>>(a.g, p);
        NEAT_LAUNCHED();
        if ((st = exclusive_scan_i64(s, p.chunk_off1, z.N, temp, temp_sz))) return st;
        if ((st = exclusive_scan_i64(s, p.chunk_off2, z.N, temp, temp_sz))) return st;
        k_plan_chunks<<<grid_for(z.max1 + z.max2, 256), 256, 0, s>>>(a.g, p);
        NEAT_LAUNCHED();
        k_plan_wg<<<grid_for(p.nwg + 1, 256), 256, 0, s>>>(a.g, p);
        NEAT_LAUNCHED();
        k_fill_f32<<<1, 64, 0, s>>>(reinterpret_cast<float*>(p.nredo), 1, 0.f);  // zero bits
        NEAT_LAUNCHED();
        k_plan_redo<<<grid_for(z.N, 256), 256, 0, s>>>(a.g, p);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_projector_fwd(neat_stream_t stream, const neat_projector_args* x, int32_t compensated, float* proj,
                           void* ws, size_t bytes)
    {
        Args a;
        int st = make_args(a, x);
        if (st) return st;
        if ((st = check_ws(a, ws, bytes))) return st;
        NEAT_REQUIRE(proj || (a.z.K == 0 && x->num_rays == 0), "projector: null output");
        hipStream_t s      = as_stream(stream);
        Plan p             = plan_view(ws, a.z);
        float* seg_partial = reinterpret_cast<float*>(region(ws, a.z, 11));
        if (a.z.K > 0)
        {
            // segments whose node has no valid slot are in no chunk: they contribute zero
            k_fill_f32<<<grid_for(a.z.K, 256), 256, 0, s>>>(seg_partial, a.z.K, 0.f);
            NEAT_LAUNCHED();
            if ((st = dispatch_fwd(s, a, p, seg_partial))) return st;
        }
        if (x->num_rays > 0)
        {
            k_proj_raysum<<<grid_for(x->num_rays, 256), 256, 0, s>>>(x->ray_seg_offset, x->num_rays, seg_partial,
                                                                      compensated, proj);
            NEAT_LAUNCHED();
        }
        return NEAT_OK;
    }

    int neat_projector_sample_values(neat_stream_t stream, const neat_projector_args* x, float* values,
                                     int64_t* seg_value_offset, void* ws, size_t bytes)
    {
        Args a;
        int st = make_args(a, x);
        if (st) return st;
        if ((st = check_ws(a, ws, bytes))) return st;
        if (!a.mfma)
            return fail(NEAT_E_UNSUPPORTED, "neat_projector_sample_values: needs the matrix-core kernels (8 channels, "
                                            "precision fp32 or bf16)");
        NEAT_REQUIRE(a.z.K == 0 || (values && seg_value_offset), "neat_projector_sample_values: null output");
        if (a.z.K == 0) return NEAT_OK;
        hipStream_t s      = as_stream(stream);
        Plan p             = plan_view(ws, a.z);
        float* seg_partial = reinterpret_cast<float*>(region(ws, a.z, 11));
        k_fill_f32<<<grid_for(a.z.K, 256), 256, 0, s>>>(seg_partial, a.z.K, 0.f);
        NEAT_LAUNCHED();
        if (a.precision == NEAT_PREC_BF16)
            st = a.m.hidden == 64 ? launch_fwd_mfma<PrecBF16, 64>(s, a, p, seg_partial, values)
                                  : launch_fwd_mfma<PrecBF16, 32>(s, a, p, seg_partial, values);
        else
            st = a.m.hidden == 64 ? launch_fwd_mfma<PrecF32, 64>(s, a, p, seg_partial, values)
                                  : launch_fwd_mfma<PrecF32, 32>(s, a, p, seg_partial, values);
        if (st) return st;
        k_seg_value_offset<<<grid_for(a.z.K, 256), 256, 0, s>>>(p.perm, p.nm_soff, a.z.K, seg_value_offset);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_projector_bwd_stage(neat_stream_t stream, const neat_projector_args* x, int32_t stage,
                                 const float* grad_proj, float* grad_grid, float* grad_params, void* ws, size_t bytes)
    {
        Args a;
        int st = make_args(a, x);
        if (st) return st;
        if ((st = check_ws(a, ws, bytes))) return st;
        NEAT_REQUIRE(stage >= 1 && stage <= 3, "neat_projector_bwd_stage: stage %d not in 1..3", stage);
        NEAT_REQUIRE((stage != 1 || grad_proj || x->num_rays == 0) && (stage != 2 || grad_params) &&
                         (stage != 3 || grad_grid),
                     "projector: null gradient buffer");
        hipStream_t s   = as_stream(stream);
        const auto& z   = a.z;
        Plan p          = plan_view(ws, z);
        const Layout L  = proj_layout(z);
        if (z.K == 0)
        {
            const int64_t ng = z.N * z.G * z.G * z.G * z.C;
            if (stage == 3)
            {
                k_fill_f32<<<grid_for(ng, 256), 256, 0, s>>>(grad_grid, ng, 0.f);
                NEAT_LAUNCHED();
            }
            if (stage == 2)
            {
                k_fill_f32<<<grid_for(z.P, 256), 256, 0, s>>>(grad_params, z.P, 0.f);
                NEAT_LAUNCHED();
            }
            return NEAT_OK;
        }
        BwdBufs b;
        b.grad_proj     = grad_proj;
        b.wt            = reinterpret_cast<float*>(region(ws, z, 12));
        b.dfeat         = reinterpret_cast<float*>(region(ws, z, 13));
        b.sgc           = reinterpret_cast<float4*>(region(ws, z, 14));
        b.dmax          = reinterpret_cast<unsigned int*>(region(ws, z, 15));
        b.partials      = reinterpret_cast<float*>(static_cast<char*>(ws) + L.total);
        b.grid_partials = reinterpret_cast<float*>(static_cast<char*>(ws) + L.total + partials_bytes(z));
        b.grad_grid     = grad_grid;
        b.grad_params   = grad_params;
        return dispatch_bwd(s, a, p, b, stage);
    }

    int neat_projector_ray_grad(neat_stream_t stream, const neat_projector_args* x, float* grad_origin,
                                float* grad_direction, void* ws, size_t bytes)
    {
        Args a;
        int st = make_args(a, x);
        if (st) return st;
        if ((st = check_ws(a, ws, bytes))) return st;
        const int64_t R = x->num_rays;
        NEAT_REQUIRE(R == 0 || grad_origin || grad_direction, "neat_projector_ray_grad: no output");
        if (R == 0) return NEAT_OK;
        hipStream_t s = as_stream(stream);
        const auto& z = a.z;
        float* seg_grad = reinterpret_cast<float*>(region(ws, z, 19));
        if (z.K > 0)
        {
            Plan p             = plan_view(ws, z);
            const float* dfeat = reinterpret_cast<const float*>(region(ws, z, 13));
            const bool bf16    = a.precision == NEAT_PREC_BF16;
            const unsigned nb  = grid_for(z.K, 256);
            if (z.C == 8)
            {
                if (a.mfma && bf16)
                    k_proj_ray_grad<8, true, true><<<nb, 256, 0, s>>>(a.g, p, a.grid_bf16, dfeat, seg_grad);
                else if (a.mfma)
                    k_proj_ray_grad<8, false, true><<<nb, 256, 0, s>>>(a.g, p, nullptr, dfeat, seg_grad);
                else
                    k_proj_ray_grad<8, false, false><<<nb, 256, 0, s>>>(a.g, p, nullptr, dfeat, seg_grad);
            }
            else
            {
                k_proj_ray_grad<4, false, false><<<nb, 256, 0, s>>>(a.g, p, nullptr, dfeat, seg_grad);
            }
            NEAT_LAUNCHED();
        }
        k_ray_grad_sum<<<grid_for(R, 256), 256, 0, s>>>(x->ray_seg_offset, R, seg_grad, grad_origin,
                                                         grad_direction);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_projector_bwd(neat_stream_t stream, const neat_projector_args* x, const float* grad_proj,
                           float* grad_grid, float* grad_params, void* ws, size_t bytes)
    {
        NEAT_REQUIRE(grad_grid && grad_params && (grad_proj || (x && x->num_rays == 0)),
                     "projector: null gradient buffer");
        for (int stage = 1; stage <= 3; ++stage)
        {
            int st = neat_projector_bwd_stage(stream, x, stage, grad_proj, grad_grid, grad_params, ws, bytes);
            if (st) return st;
        }
        return NEAT_OK;
    }
}
