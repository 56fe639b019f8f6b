// Volume regulariser of the reference's training step (GeometryExEx::VolumeRegularizer, geometry_ex_ex.cpp:86-160):
// the edge loss between the grids of neighbouring active nodes along their shared faces, and the total variation of
// the active grids.
//
//   neighbour pairs : NodeNeighborSamples (HyperTree.cu:830-940), one thread per candidate face point, compacted in
//                     the canonical order (active node, side, front, face point) by a scan (the reference's order
//                     depends on an atomic counter).
//   edge loss       : per pair mean_c |f_a - f_b| times a per-pair random factor (geometry_ex_ex.cpp:121-134); the
//                     factor is a counter-based hash of (seed, pair) where the reference draws torch::rand_like.
//   total variation : TVLoss::forward (geometry_ex_ex.h:18-61) over the active grids; its gradient is formed per
//                     grid entry from the entry's neighbours, so no atomics are needed.
// Sums are FP64 and run in a fixed order (per-block partials, then one block), so losses are deterministic.
#include <type_traits>

#include "gridsample_dev.hpp"
#include "tree_dev.hpp"

namespace neat
{

struct Strides5
{
    int64_t s[5];  // element strides of (N, C, Z, Y, X)
};

// Face point sid of side `side`, front `front` of node `node` (NodeNeighborSamples' local / global coordinates).
__device__ __forceinline__ void face_point(const float* pmin, const float* pmax, int32_t node, int side, int front,
                                           int32_t sid, int32_t G, float p[3])
{
    const int sa = (side + 1) % 3, sb = (side + 2) % 3;
    float local[3];
    local[side] = front ? 1.f : 0.f;
    local[sa]   = (float)(sid % G) / ((float)G - 1.f);
    local[sb]   = (float)(sid / G) / ((float)G - 1.f);
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        const float mn = pmin[3 * node + d], mx = pmax[3 * node + d];
        p[d]           = local[d] * (mx - mn) + mn;
    }
}

// candidate c = ((a * 3 + side) * 2 + front) * G^2 + sid -> neighbour node id, or -1 (no pair). I is the index type:
// 32-bit whenever the candidate count allows (cheaper divisions), 64-bit otherwise.
struct Candidate
{
    int32_t aidx, side, front, sid;
};

template <typename I>
__device__ __forceinline__ Candidate split_candidate(I c, int32_t G)
{
    const I G2 = (I)G * (I)G;
    const I q  = c / G2;
    return Candidate{(int32_t)(q / 6), (int32_t)((q / 2) % 3), (int32_t)(q % 2), (int32_t)(c - q * G2)};
}

template <typename I>
__global__ void __launch_bounds__(256) k_neighbor_candidates(const float* pmin, const float* pmax, const uint8_t* flags,
                                                             int32_t first_leaf, const int64_t* active_ids,
                                                             int64_t num_candidates, int32_t G, float eps,
                                                             int32_t* other, int64_t* is_pair)
{
    const I c = (I)blockIdx.x * (I)blockDim.x + (I)threadIdx.x;
    if ((int64_t)c >= num_candidates) return;
    const Candidate k  = split_candidate<I>(c, G);
    const int32_t node = (int32_t)active_ids[k.aidx];
    float p[3];
    face_point(pmin, pmax, node, k.side, k.front, k.sid, G, p);
    float q[3] = {p[0], p[1], p[2]};
    q[k.side]  = p[k.side] + (k.front ? eps : -eps);
    const int32_t o = active_node_at(pmin, pmax, flags, first_leaf, q);
    const bool pair = o != -1 && o != node;
    other[c]        = pair ? o : -1;
    is_pair[c]      = pair ? 1 : 0;
}

// pairs in candidate order; is_pair holds the exclusive scan (the output row of each pair)
__global__ void __launch_bounds__(256) k_neighbor_emit(const float* pmin, const float* pmax, const int64_t* active_ids,
                                                       int64_t num_candidates, int32_t G, const int32_t* other,
                                                       const int64_t* row, float* out_coord, int64_t* out_node,
                                                       int64_t* out_ray, int64_t* out_local, float* out_weight)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= num_candidates || other[c] < 0) return;
    const int64_t G2   = (int64_t)G * G;
    const int32_t sid  = (int32_t)(c % G2);
    const int front    = (int)((c / G2) % 2);
    const int side     = (int)((c / (2 * G2)) % 3);
    const int32_t node = (int32_t)active_ids[c / (6 * G2)];
    float p[3];
    face_point(pmin, pmax, node, side, front, sid, G, p);
    const int64_t r = 2 * row[c];
    for (int k = 0; k < 2; ++k)
    {
#pragma unroll
        for (int d = 0; d < 3; ++d) out_coord[3 * (r + k) + d] = p[d];
        out_node[r + k] = k == 0 ? node : other[c];
        if (out_ray) out_ray[r + k] = r + k;
        if (out_local) out_local[r + k] = k;
        if (out_weight) out_weight[r + k] = 1.f;
    }
}

// per-pair random factor in [0, 1): 24 bits of a splitmix64 hash of (seed, pair)
__device__ __forceinline__ float pair_factor(uint64_t seed, int64_t pair)
{
    return (float)(mix64(seed ^ ((uint64_t)pair * 0x9E3779B97F4A7C15ull)) >> 40) * (1.0f / 16777216.0f);
}

constexpr int kRedThreads = 256;
constexpr int64_t kEdgeDirectBlocks = 4096;  // grid-stride blocks of the fused edge loss (partials per call)

// block total of one double per thread, valid in thread 0, in a fixed order: a butterfly within each wave, then the
// waves' sums in wave order (blockDim.x = kRedThreads)
__device__ __forceinline__ double block_total(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __shared__ double s[kRedThreads / 64];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
    {
#pragma unroll
        for (int w = 0; w < kRedThreads / 64; ++w) t += s[w];
    }
    return t;
}

__device__ __forceinline__ void block_sum_to(double v, double* out)
{
    const double t = block_total(v);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// this thread's share of partials[0, n) (strided by the block), four independent accumulators so the loads overlap
__device__ __forceinline__ double thread_share(const double* partials, int64_t n)
{
    const int64_t step = blockDim.x;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int64_t i = threadIdx.x;
    for (; i + 3 * step < n; i += 4 * step)
    {
        a0 += partials[i];
        a1 += partials[i + step];
        a2 += partials[i + 2 * step];
        a3 += partials[i + 3 * step];
    }
    for (; i < n; i += step) a0 += partials[i];
    return (a0 + a1) + (a2 + a3);
}

__global__ void __launch_bounds__(kRedThreads) k_sum_partials(const double* partials, int64_t n, double scale,
                                                              double* out)
{
    const double t = block_total(thread_share(partials, n));
    if (threadIdx.x == 0) *out = t * scale;
}

// edge_error_p = mean_c |f[2p, c] - f[2p + 1, c]| * factor_p, summed per block
__global__ void __launch_bounds__(kRedThreads) k_edge_fwd(const float* f, int64_t P, int32_t C, uint64_t seed,
                                                          double* partials)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v        = 0.0;
    if (p < P)
    {
        float e = 0.f;
        for (int c = 0; c < C; ++c) e += fabsf(f[(2 * p) * C + c] - f[(2 * p + 1) * C + c]);
        v = (double)(e / (float)C * pair_factor(seed, p));
    }
    block_sum_to(v, partials);
}

// d loss / d f for loss = mean_p(edge_error_p) * weight * C, times the incoming scalar gradient
__global__ void __launch_bounds__(256) k_edge_bwd(const float* f, int64_t P, int32_t C, uint64_t seed, double weight,
                                                  const float* grad_loss, float* grad_f)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P * C) return;
    const int64_t p = i / C;
    const int32_t c = (int32_t)(i % C);
    const float d   = f[(2 * p) * C + c] - f[(2 * p + 1) * C + c];
    const float s   = d > 0.f ? 1.f : (d < 0.f ? -1.f : 0.f);
    // d/dd of |d| / C * r, averaged over P, times weight * C
    const float g = (float)((double)*grad_loss * weight / (double)P) * pair_factor(seed, p) * s;
    grad_f[(2 * p) * C + c]     = g;
    grad_f[(2 * p + 1) * C + c] = -g;
}

__device__ __forceinline__ float sgn(float d) { return d > 0.f ? 1.f : (d < 0.f ? -1.f : 0.f); }

// The edge loss straight from the candidate list, without materialising the neighbour rows, their local coordinates
// or features: lane (candidate, channel quad), so the lanes of a pair read (and, for the gradient, add to) each
// corner's contiguous channel run of a channels-last grid together (16-byte loads when kVec4). Per pair k (row[c],
// the canonical pair index) the lane samples its channels of both nodes' grids at the shared face point exactly as
// the grid sample does (same local coordinates, same trilinear weights and corner order), so the features, the
// random factor r_k and the per-feature gradient sign(d) r_k gl weight / P equal the unfused path's bit for bit;
// corners whose weight is exactly zero add nothing and are skipped (face points sit on grid vertices of same-size
// neighbours). The loss runs per channel as |d| r_k in FP64 (the 1/C of the channel mean cancels against the
// reference's factor C), per-block partials over a fixed grid-stride assignment, so it is deterministic.
template <typename I, bool kVec4, bool kLoss, bool kGrad>
__global__ void __launch_bounds__(kRedThreads) k_edge_direct(GridArgs g, const float* pmin, const float* pmax,
                                                             const int64_t* active_ids, const int64_t* node_slot,
                                                             int64_t num_candidates, int32_t G, const int32_t* other,
                                                             const int64_t* row, uint64_t seed, double weight,
                                                             const float* grad_loss, float* grad_grid,
                                                             double* partials)
{
    const I quads = (I)((g.C + 3) / 4);
    const I total = (I)num_candidates * quads;
    const int64_t P = row[num_candidates];
    float gscale    = 0.f;
    if constexpr (kGrad)
        if (P > 0) gscale = (float)((double)*grad_loss * weight / (double)P);
    double acc = 0.0;
    for (I i = (I)blockIdx.x * (I)blockDim.x + (I)threadIdx.x; i < total; i += (I)gridDim.x * (I)blockDim.x)
    {
        const I c       = i / quads;
        const int32_t o = other[c];
        if (o < 0) continue;
        const int32_t ch0  = 4 * (int32_t)(i - c * quads);
        const Candidate cd = split_candidate<I>(c, G);
        const int32_t node = (int32_t)active_ids[cd.aidx];
        const int64_t sa = node_slot[node], so = node_slot[o];
        if (sa < 0 || sa >= g.N || so < 0 || so >= g.N) continue;  // the grid sample reads zeros there
        float p[3];
        face_point(pmin, pmax, node, cd.side, cd.front, cd.sid, G, p);
        float ua[3], uo[3];
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            const float mna = pmin[3 * node + d], mxa = pmax[3 * node + d];
            const float mno = pmin[3 * o + d], mxo = pmax[3 * o + d];
            ua[d] = (p[d] - mna) / (mxa - mna) * 2 - 1;
            uo[d] = (p[d] - mno) / (mxo - mno) * 2 - 1;
        }
        const TriLocal ta = trilinear_local(g, ua[0], ua[1], ua[2]);
        const TriLocal to = trilinear_local(g, uo[0], uo[1], uo[2]);
        const float* ga   = g.grid + sa * g.s0;
        const float* go   = g.grid + so * g.s0;
        float fa[4] = {0.f, 0.f, 0.f, 0.f}, fo[4] = {0.f, 0.f, 0.f, 0.f};
        if constexpr (kVec4)
        {
#pragma unroll
            for (int k = 0; k < 8; ++k)
            {
                const float4 va = *reinterpret_cast<const float4*>(ga + ta.off[k] + ch0);
                fa[0] += va.x * ta.w[k];
                fa[1] += va.y * ta.w[k];
                fa[2] += va.z * ta.w[k];
                fa[3] += va.w * ta.w[k];
            }
#pragma unroll
            for (int k = 0; k < 8; ++k)
            {
                const float4 vo = *reinterpret_cast<const float4*>(go + to.off[k] + ch0);
                fo[0] += vo.x * to.w[k];
                fo[1] += vo.y * to.w[k];
                fo[2] += vo.z * to.w[k];
                fo[3] += vo.w * to.w[k];
            }
        }
        else
        {
            for (int j = 0; j < 4 && ch0 + j < g.C; ++j)
            {
                const int32_t co = (ch0 + j) * (int32_t)g.s1;
#pragma unroll
                for (int k = 0; k < 8; ++k) fa[j] += ga[ta.off[k] + co] * ta.w[k];
#pragma unroll
                for (int k = 0; k < 8; ++k) fo[j] += go[to.off[k] + co] * to.w[k];
            }
        }
        const float r = pair_factor(seed, row[c]);
        const int nch = min(4, g.C - ch0);
        float gv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            const float d = fa[j] - fo[j];
            if (j < nch)
            {
                if constexpr (kLoss) acc += (double)fabsf(d) * (double)r;
            }
            gv[j] = j < nch ? gscale * r * sgn(d) : 0.f;
        }
        if constexpr (kGrad)
        {
            float* da = grad_grid + sa * g.s0;
            float* dq = grad_grid + so * g.s0;
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (ta.w[k] != 0.f)
                    for (int j = 0; j < nch; ++j)
                        if (gv[j] != 0.f) atomicAdd(da + ta.off[k] + (ch0 + j) * (int32_t)g.s1, ta.w[k] * gv[j]);
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (to.w[k] != 0.f)
                    for (int j = 0; j < nch; ++j)
                        if (gv[j] != 0.f) atomicAdd(dq + to.off[k] + (ch0 + j) * (int32_t)g.s1, to.w[k] * -gv[j]);
        }
    }
    if constexpr (kLoss) block_sum_to(acc, partials);
}

// loss = sum(partials) * weight / P with P = row[num_candidates] read on the device (0 when there are no pairs)
__global__ void __launch_bounds__(kRedThreads) k_sum_partials_per_pair(const double* partials, int64_t n,
                                                                       const int64_t* row, int64_t num_candidates,
                                                                       double weight, double* out)
{
    const double t  = block_total(thread_share(partials, n));
    const int64_t P = row[num_candidates];
    if (threadIdx.x == 0) *out = P > 0 ? t * weight / (double)P : 0.0;
}

// Total variation, kTvPerThread grid entries per thread (strided by the block size), the
// entries of a node numbered in memory order (channel fastest when the grid is channels-last, x fastest otherwise) so
// the entry and its three forward neighbours are coalesced loads. Blocks are laid out node-major (blocks_per_node per
// node), so the within-node entry index stays 32-bit and is split by multiply-shift division (divmod_small)
// (exact for indices below 2^24 and divisors below 256, which the host checks). When the node's entries are dense in
// memory order (channels-last or plain contiguous), an entry's offset is its index and only the boundary tests need
// the coordinates.
constexpr int kTvPerThread = 8;
constexpr uint32_t kXcds   = 8;

// Logical block of physical block b in a grid of nb blocks (nb a multiple of kXcds). The hardware hands consecutive
// blocks to the XCDs in turn; this gives XCD x the logical range [x nb / kXcds, (x + 1) nb / kXcds) in order, so the
// y and z neighbours of a block's entries (other blocks' entries, a few KB away) are read through the same XCD's L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) { return (b % kXcds) * (nb / kXcds) + b / kXcds; }

struct TvDims
{
    int32_t C, G;
    uint32_t MC, MG;          // ceil(2^32 / C), ceil(2^32 / G) (0 for a divisor of 1)
    int32_t per_node;         // entries per node, C * G^3
    int32_t blocks_per_node;  // blocks of one node
    uint32_t blocks;          // num_active * blocks_per_node; the grid is rounded up to a multiple of kXcds
    bool channels_last;
    bool dense;  // offset of entry e within a slot is e
    int64_t N;   // grid slots: a slot outside [0, N) reads and writes nothing
};

// q = e / d, rem = e % d for 0 <= e < 2^24 and 1 <= d < 256 with M = ceil(2^32 / d): the round-up multiply of
// Granlund and Montgomery, "Division by invariant integers using multiplication" (PLDI 1994). e * M / 2^32 exceeds
// e / d by less than e / 2^32 < 2^-8 < 1 / d, so its floor is exact; d = 1 passes through.
__device__ __forceinline__ int32_t divmod_small(int32_t e, int32_t d, uint32_t M, int32_t& rem)
{
    const int32_t q = d == 1 ? e : (int32_t)__umulhi((uint32_t)e, M);
    rem             = e - q * d;
    return q;
}

struct TvEntry
{
    int32_t c, z, y, x;
};

__device__ __forceinline__ TvEntry tv_entry(int32_t e, const TvDims& D)
{
    TvEntry t;
    if (D.channels_last)
    {
        e   = divmod_small(e, D.C, D.MC, t.c);
        e   = divmod_small(e, D.G, D.MG, t.x);
        t.z = divmod_small(e, D.G, D.MG, t.y);
    }
    else
    {
        e   = divmod_small(e, D.G, D.MG, t.x);
        e   = divmod_small(e, D.G, D.MG, t.y);
        t.c = divmod_small(e, D.G, D.MG, t.z);
    }
    return t;
}

// sum over the node's entries of |x+1 - v| + |y+1 - v| + |z+1 - v| (where the neighbour exists), FP64 per thread,
// one partial per block; scale per axis = 1 / (A * C * G * G * (G - 1)) (the means), times C * weight, applied by
// the final sum
__global__ void __launch_bounds__(kRedThreads) k_tv_fwd(const float* __restrict__ grid, TvDims D, Strides5 S,
                                                        const int64_t* __restrict__ slots, double* __restrict__ partials)
{
    const int64_t* st = S.s;
    const uint32_t L  = xcd_block(blockIdx.x, gridDim.x);
    if (L >= D.blocks)
    {
        if (threadIdx.x == 0) partials[L] = 0.0;  // a padding block
        return;
    }
    const int64_t a  = L / D.blocks_per_node;
    const int32_t e0 = (int32_t)(L - a * D.blocks_per_node) * (kRedThreads * kTvPerThread) + threadIdx.x;
    const int64_t sl = slots[a];
    if (sl < 0 || sl >= D.N)
    {
        if (threadIdx.x == 0) partials[L] = 0.0;  // no grid row: nothing to read
        return;
    }
    const float* g0  = grid + sl * st[0];
    const int32_t G  = D.G;
    // every load first (a missing neighbour reads the entry itself: difference 0), then the sums
    float v[kTvPerThread][4];
#pragma unroll
    for (int k = 0; k < kTvPerThread; ++k)
    {
        const int32_t e = min(e0 + k * kRedThreads, D.per_node - 1);
        const TvEntry t = tv_entry(e, D);
        const int64_t o = D.dense ? (int64_t)e : t.c * st[1] + t.z * st[2] + t.y * st[3] + t.x * st[4];
        v[k][0]         = g0[o];
        v[k][1]         = g0[o + (t.x + 1 < G ? st[4] : 0)];
        v[k][2]         = g0[o + (t.y + 1 < G ? st[3] : 0)];
        v[k][3]         = g0[o + (t.z + 1 < G ? st[2] : 0)];
    }
    double sum = 0.0;
#pragma unroll
    for (int k = 0; k < kTvPerThread; ++k)
    {
        float acc = 0.f;
        if (e0 + k * kRedThreads < D.per_node)
        {
#pragma unroll
            for (int j = 1; j <= 3; ++j) acc += fabsf(v[k][j] - v[k][0]);
        }
        sum += (double)acc;
    }
    const double tot = block_total(sum);
    if (threadIdx.x == 0) partials[L] = tot;
}

// Forward and backward of the total variation in one pass (kTvPerThread entries per thread, as the forward): the
// loss partials as k_tv_fwd's (kLoss), and the gradient of each entry, sign(v - prev) - sign(next - v) per axis times
// (*grad_loss) * scale, written (overwrite) or added (accumulate); without kLoss it is the backward alone,
// so the grid is read once and, when overwriting, the gradient buffer is neither cleared nor read. All loads of a
// thread's entries are issued before any arithmetic, so a wave keeps 7 * kTvPerThread loads in flight; a missing
// neighbour (at the grid's border) reads the entry itself, whose difference 0 adds nothing to either sum. kDense: the
// entry's offset within its slot is its index (32-bit offsets).
template <bool kOverwrite, bool kDense, bool kLoss = true>
__global__ void __launch_bounds__(kRedThreads) k_tv_fwd_bwd(const float* __restrict__ grid, TvDims D, Strides5 S,
                                                            const int64_t* __restrict__ slots, double scale,
                                                            const float* __restrict__ grad_loss,
                                                            float* __restrict__ grad_grid, double* __restrict__ partials)
{
    using Off = typename std::conditional<kDense, int32_t, int64_t>::type;
    const int64_t* st = S.s;
    const uint32_t L  = xcd_block(blockIdx.x, gridDim.x);
    if (L >= D.blocks)
    {
        if (kLoss && threadIdx.x == 0) partials[L] = 0.0;  // a padding block
        return;
    }
    const int64_t a  = L / D.blocks_per_node;
    const int32_t e0 = (int32_t)(L - a * D.blocks_per_node) * (kRedThreads * kTvPerThread) + threadIdx.x;
    const int64_t sl = slots[a];
    if (sl < 0 || sl >= D.N)
    {
        if (kLoss && threadIdx.x == 0) partials[L] = 0.0;  // no grid row: nothing to read or write
        return;
    }
    const int64_t b0 = sl * st[0];
    const float* g0  = grid + b0;
    float* d0        = grad_grid + b0;
    const int32_t G  = D.G;
    const Off sx = (Off)st[4], sy = (Off)st[3], sz = (Off)st[2];
    Off off[kTvPerThread];
    float v[kTvPerThread][7];  // the entry, its +x +y +z and -x -y -z neighbours
#pragma unroll
    for (int k = 0; k < kTvPerThread; ++k)
    {
        const int32_t e = min(e0 + k * kRedThreads, D.per_node - 1);
        const TvEntry t = tv_entry(e, D);
        const Off o = kDense ? (Off)e : (Off)(t.c * st[1] + t.z * st[2] + t.y * st[3] + t.x * st[4]);
        off[k]      = o;
        v[k][0]     = g0[o];
        v[k][1]     = g0[o + (t.x + 1 < G ? sx : 0)];
        v[k][2]     = g0[o + (t.y + 1 < G ? sy : 0)];
        v[k][3]     = g0[o + (t.z + 1 < G ? sz : 0)];
        v[k][4]     = g0[o - (t.x > 0 ? sx : 0)];
        v[k][5]     = g0[o - (t.y > 0 ? sy : 0)];
        v[k][6]     = g0[o - (t.z > 0 ? sz : 0)];
    }
    const double gs = (double)*grad_loss * scale;
    double sum      = 0.0;
#pragma unroll
    for (int k = 0; k < kTvPerThread; ++k)
    {
        if (e0 + k * kRedThreads >= D.per_node) break;
        const float v0 = v[k][0];
        float acc = 0.f, s = 0.f;
#pragma unroll
        for (int j = 1; j <= 3; ++j)
        {
            const float d = v[k][j] - v0;
            acc += fabsf(d);
            s -= sgn(d);
        }
#pragma unroll
        for (int j = 4; j <= 6; ++j) s += sgn(v0 - v[k][j]);
        sum += (double)acc;
        const float gv = (float)(gs * s);
        if constexpr (kOverwrite)
            d0[off[k]] = gv;
        else
            d0[off[k]] += gv;
    }
    if constexpr (kLoss)
    {
        const double tot = block_total(sum);
        if (threadIdx.x == 0) partials[L] = tot;
    }
}

}  // namespace neat

using namespace neat;

static int64_t blocks_for(int64_t n, int t) { return (n + t - 1) / t; }
static size_t align256(size_t b) { return (b + 255) / 256 * 256; }
static Strides5 strides5(const int64_t* st) { return Strides5{{st[0], st[1], st[2], st[3], st[4]}}; }

static int64_t xcd_grid(int64_t blocks) { return (blocks + kXcds - 1) / kXcds * kXcds; }

static TvDims tv_dims(int32_t C, int32_t G, const int64_t* strides, int entries_per_block, int64_t num_active,
                      int64_t N)
{
    TvDims D;
    D.N               = N;
    D.C               = C;
    D.G               = G;
    D.MC              = C >= 2 ? (uint32_t)((((uint64_t)1 << 32) + C - 1) / C) : 0u;
    D.MG              = G >= 2 ? (uint32_t)((((uint64_t)1 << 32) + G - 1) / G) : 0u;
    D.per_node        = C * G * G * G;
    D.blocks_per_node = (D.per_node + entries_per_block - 1) / entries_per_block;
    D.blocks          = (uint32_t)(num_active * D.blocks_per_node);
    D.channels_last   = strides[1] == 1;
    const int64_t G2 = (int64_t)G * G, G3 = G2 * G;
    D.dense = D.channels_last ? (strides[4] == C && strides[3] == G * (int64_t)C && strides[2] == G2 * C)
                              : (strides[4] == 1 && strides[3] == G && strides[2] == G2 && strides[1] == G3);
    return D;
}

static void launch_candidates(hipStream_t s, const float* pmin, const float* pmax, const uint8_t* flags,
                              int32_t num_nodes, const int64_t* ids, int64_t n, int32_t G, float eps, int32_t* other,
                              int64_t* is_pair)
{
    const unsigned nb = (unsigned)blocks_for(n, 256);
    const int32_t first_leaf = (num_nodes - 1) / 8;
    if (n < ((int64_t)1 << 31))
        k_neighbor_candidates<uint32_t><<<nb, 256, 0, s>>>(pmin, pmax, flags, first_leaf, ids, n, G, eps, other,
                                                            is_pair);
    else
        k_neighbor_candidates<int64_t><<<nb, 256, 0, s>>>(pmin, pmax, flags, first_leaf, ids, n, G, eps, other,
                                                           is_pair);
}

template <typename I, bool V>
static void launch_edge_direct(unsigned nb, hipStream_t s, bool want_loss, bool want_grad, const GridArgs& g,
                               const float* pmin, const float* pmax, const int64_t* ids, const int64_t* slot,
                               int64_t n, int32_t G, const int32_t* other, const int64_t* row, uint64_t seed,
                               double weight, const float* gl, float* gg, double* part)
{
    if (want_loss && want_grad)
        k_edge_direct<I, V, true, true><<<nb, kRedThreads, 0, s>>>(g, pmin, pmax, ids, slot, n, G, other, row, seed,
                                                                   weight, gl, gg, part);
    else if (want_loss)
        k_edge_direct<I, V, true, false><<<nb, kRedThreads, 0, s>>>(g, pmin, pmax, ids, slot, n, G, other, row, seed,
                                                                    weight, nullptr, nullptr, part);
    else
        k_edge_direct<I, V, false, true><<<nb, kRedThreads, 0, s>>>(g, pmin, pmax, ids, slot, n, G, other, row, seed,
                                                                    weight, gl, gg, part);
}

extern "C"
{
    size_t neat_neighbor_pairs_workspace_bytes(int64_t num_active, int32_t grid_size)
    {
        if (num_active < 0 || grid_size < 2) return 0;
        const int64_t n = num_active * 6 * (int64_t)grid_size * grid_size;
        return align256((size_t)n * 4) + align256((size_t)(n + 1) * 8) + scan_workspace_bytes(n + 1);
    }

    int neat_neighbor_pairs(neat_stream_t stream, const float* node_pos_min, const float* node_pos_max,
                            const uint8_t* node_flags, int32_t num_nodes, const int64_t* active_node_ids,
                            int64_t num_active, int32_t grid_size, float epsilon, int64_t max_pairs,
                            float* out_coord, int64_t* out_node, int64_t* out_ray, int64_t* out_local,
                            float* out_weight, int64_t* num_pairs, void* workspace, size_t workspace_bytes)
    {
        NEAT_REQUIRE(node_pos_min && node_pos_max && node_flags && num_pairs, "neat_neighbor_pairs: null pointer");
        NEAT_REQUIRE(num_active >= 0 && grid_size >= 2 && num_nodes > 0, "neat_neighbor_pairs: bad sizes");
        const int64_t n = num_active * 6 * (int64_t)grid_size * grid_size;
        NEAT_REQUIRE(max_pairs >= n, "neat_neighbor_pairs: max_pairs %lld < %lld candidates", (long long)max_pairs,
                     (long long)n);
        const size_t need = neat_neighbor_pairs_workspace_bytes(num_active, grid_size);
        NEAT_REQUIRE(workspace && workspace_bytes >= need, "neat_neighbor_pairs: workspace too small (%zu < %zu)",
                     workspace_bytes, need);
        hipStream_t s = as_stream(stream);
        char* w       = static_cast<char*>(workspace);
        int32_t* other = reinterpret_cast<int32_t*>(w);
        int64_t* row   = reinterpret_cast<int64_t*>(w + align256((size_t)n * 4));
        void* scan_ws  = w + align256((size_t)n * 4) + align256((size_t)(n + 1) * 8);
        const size_t scan_bytes = scan_workspace_bytes(n + 1);
        if (n == 0)
        {
            NEAT_HIP(hipMemsetAsync(num_pairs, 0, sizeof(int64_t), s));
            return NEAT_OK;
        }
        NEAT_REQUIRE(active_node_ids && out_coord && out_node, "neat_neighbor_pairs: null pointer");
        launch_candidates(s, node_pos_min, node_pos_max, node_flags, num_nodes, active_node_ids, n, grid_size, epsilon,
                          other, row);
        NEAT_LAUNCHED();
        int st = exclusive_scan_i64(s, row, n, scan_ws, scan_bytes);  // row[n] = total pairs
        if (st) return st;
        k_neighbor_emit<<<(unsigned)blocks_for(n, 256), 256, 0, s>>>(node_pos_min, node_pos_max, active_node_ids, n,
                                                                      grid_size, other, row, out_coord, out_node,
                                                                      out_ray, out_local, out_weight);
        NEAT_LAUNCHED();
        NEAT_HIP(hipMemcpyAsync(num_pairs, row + n, sizeof(int64_t), hipMemcpyDeviceToDevice, s));
        return NEAT_OK;
    }

    size_t neat_reg_workspace_bytes(int64_t num_items) { return align256((size_t)(blocks_for(num_items, kRedThreads) + 1) * 8); }

    size_t neat_tv_workspace_bytes(int64_t num_active, int32_t C, int32_t G)
    {
        if (num_active < 0 || C <= 0 || G < 2) return 0;
        const int64_t per_node = C * (int64_t)G * G * G;
        const int64_t bpn      = (per_node + kRedThreads * kTvPerThread - 1) / (kRedThreads * kTvPerThread);
        return align256((size_t)(num_active * bpn + kXcds + 1) * 8);
    }

    int neat_edge_loss_fwd(neat_stream_t stream, const float* features, int64_t num_pairs, int32_t C, uint64_t seed,
                           double weight, double* loss, void* workspace, size_t workspace_bytes)
    {
        NEAT_REQUIRE(loss && num_pairs >= 0 && C > 0, "neat_edge_loss_fwd: bad arguments");
        hipStream_t s = as_stream(stream);
        if (num_pairs == 0)
        {
            NEAT_HIP(hipMemsetAsync(loss, 0, sizeof(double), s));
            return NEAT_OK;
        }
        NEAT_REQUIRE(features, "neat_edge_loss_fwd: null features");
        const int64_t nb = blocks_for(num_pairs, kRedThreads);
        NEAT_REQUIRE(workspace && workspace_bytes >= neat_reg_workspace_bytes(num_pairs),
                     "neat_edge_loss_fwd: workspace too small");
        double* part = static_cast<double*>(workspace);
        k_edge_fwd<<<(unsigned)nb, kRedThreads, 0, s>>>(features, num_pairs, C, seed, part);
        NEAT_LAUNCHED();
        // loss = mean_p(edge_error_p) * weight * C
        k_sum_partials<<<1, kRedThreads, 0, s>>>(part, nb, weight * C / (double)num_pairs, loss);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_edge_loss_bwd(neat_stream_t stream, const float* features, int64_t num_pairs, int32_t C, uint64_t seed,
                           double weight, const float* grad_loss, float* grad_features)
    {
        NEAT_REQUIRE(num_pairs >= 0 && C > 0, "neat_edge_loss_bwd: bad arguments");
        if (num_pairs == 0) return NEAT_OK;
        NEAT_REQUIRE(features && grad_loss && grad_features, "neat_edge_loss_bwd: null pointer");
        const int64_t n = num_pairs * C;
        k_edge_bwd<<<(unsigned)blocks_for(n, 256), 256, 0, as_stream(stream)>>>(features, num_pairs, C, seed, weight,
                                                                                 grad_loss, grad_features);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_tv_loss_fwd(neat_stream_t stream, const float* grid, int64_t N, int32_t C, int32_t G,
                         const int64_t* strides, const int64_t* slots, int64_t num_active, double weight, double* loss,
                         void* workspace, size_t workspace_bytes)
    {
        NEAT_REQUIRE(loss && strides && C > 0 && G >= 2 && num_active >= 0 && N >= 0, "neat_tv_loss_fwd: bad arguments");
        hipStream_t s = as_stream(stream);
        if (num_active == 0)
        {
            NEAT_HIP(hipMemsetAsync(loss, 0, sizeof(double), s));
            return NEAT_OK;
        }
        NEAT_REQUIRE(grid && slots, "neat_tv_loss_fwd: null pointer");
        NEAT_REQUIRE(num_active <= N, "neat_tv_loss_fwd: %lld active nodes for %lld grid rows", (long long)num_active,
                     (long long)N);
        NEAT_REQUIRE(C * (int64_t)G * G * G < (1 << 24) && C < 256 && G < 256,
                     "neat_tv_loss_fwd: C * G^3 must be below 2^24, C and G below 256");
        const TvDims D   = tv_dims(C, G, strides, kRedThreads * kTvPerThread, num_active, N);
        const int64_t nb = xcd_grid(D.blocks);
        NEAT_REQUIRE(workspace && workspace_bytes >= neat_tv_workspace_bytes(num_active, C, G),
                     "neat_tv_loss_fwd: workspace too small");
        double* part = static_cast<double*>(workspace);
        k_tv_fwd<<<(unsigned)nb, kRedThreads, 0, s>>>(grid, D, strides5(strides), slots, part);
        NEAT_LAUNCHED();
        // each axis: mean over (A, C, G, G, G - 1) differences; the three axes have equal counts; times C * weight
        k_sum_partials<<<1, kRedThreads, 0, s>>>(part, nb, weight * C / ((double)num_active * C * G * G * (G - 1)),
                                                 loss);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_tv_loss_bwd(neat_stream_t stream, const float* grid, int64_t N, int32_t C, int32_t G,
                         const int64_t* strides, const int64_t* slots, int64_t num_active, double weight,
                         const float* grad_loss, float* grad_grid)
    {
        NEAT_REQUIRE(strides && C > 0 && G >= 2 && num_active >= 0 && N >= 0, "neat_tv_loss_bwd: bad arguments");
        if (num_active == 0) return NEAT_OK;
        NEAT_REQUIRE(grid && slots && grad_loss && grad_grid, "neat_tv_loss_bwd: null pointer");
        NEAT_REQUIRE(num_active <= N, "neat_tv_loss_bwd: %lld active nodes for %lld grid rows", (long long)num_active,
                     (long long)N);
        NEAT_REQUIRE(C * (int64_t)G * G * G < (1 << 24) && C < 256 && G < 256,
                     "neat_tv_loss_bwd: C * G^3 must be below 2^24, C and G below 256");
        const TvDims D     = tv_dims(C, G, strides, kRedThreads * kTvPerThread, num_active, N);
        const double scale = weight * C / ((double)num_active * C * G * G * (G - 1));
        // the combined pass without its loss: same per-entry gradient, all of a thread's loads issued first
        if (D.dense)
            k_tv_fwd_bwd<false, true, false><<<(unsigned)xcd_grid(D.blocks), kRedThreads, 0, as_stream(stream)>>>(
                grid, D, strides5(strides), slots, scale, grad_loss, grad_grid, nullptr);
        else
            k_tv_fwd_bwd<false, false, false><<<(unsigned)xcd_grid(D.blocks), kRedThreads, 0, as_stream(stream)>>>(
                grid, D, strides5(strides), slots, scale, grad_loss, grad_grid, nullptr);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_tv_loss_fwd_bwd(neat_stream_t stream, const float* grid, int64_t N, int32_t C, int32_t G,
                             const int64_t* strides, const int64_t* slots, int64_t num_active, double weight,
                             double* loss, const float* grad_loss, float* grad_grid, int32_t overwrite,
                             void* workspace, size_t workspace_bytes)
    {
        NEAT_REQUIRE(loss && strides && C > 0 && G >= 2 && num_active >= 0 && N >= 0,
                     "neat_tv_loss_fwd_bwd: bad arguments");
        hipStream_t s = as_stream(stream);
        if (num_active == 0)
        {
            NEAT_HIP(hipMemsetAsync(loss, 0, sizeof(double), s));
            return NEAT_OK;
        }
        NEAT_REQUIRE(grid && slots && grad_loss && grad_grid, "neat_tv_loss_fwd_bwd: null pointer");
        NEAT_REQUIRE(num_active <= N, "neat_tv_loss_fwd_bwd: %lld active nodes for %lld grid rows",
                     (long long)num_active, (long long)N);
        NEAT_REQUIRE(C * (int64_t)G * G * G < (1 << 24) && C < 256 && G < 256,
                     "neat_tv_loss_fwd_bwd: C * G^3 must be below 2^24, C and G below 256");
        NEAT_REQUIRE(workspace && workspace_bytes >= neat_tv_workspace_bytes(num_active, C, G),
                     "neat_tv_loss_fwd_bwd: workspace too small");
        const TvDims D     = tv_dims(C, G, strides, kRedThreads * kTvPerThread, num_active, N);
        const int64_t nb   = xcd_grid(D.blocks);
        const double scale = weight * C / ((double)num_active * C * G * G * (G - 1));
        double* part       = static_cast<double*>(workspace);
        const Strides5 S5 = strides5(strides);
        if (overwrite && D.dense)
            k_tv_fwd_bwd<true, true><<<(unsigned)nb, kRedThreads, 0, s>>>(grid, D, S5, slots, scale, grad_loss, grad_grid,
                                                                           part);
        else if (overwrite)
            k_tv_fwd_bwd<true, false><<<(unsigned)nb, kRedThreads, 0, s>>>(grid, D, S5, slots, scale, grad_loss,
                                                                            grad_grid, part);
        else if (D.dense)
            k_tv_fwd_bwd<false, true><<<(unsigned)nb, kRedThreads, 0, s>>>(grid, D, S5, slots, scale, grad_loss,
                                                                            grad_grid, part);
        else
            k_tv_fwd_bwd<false, false><<<(unsigned)nb, kRedThreads, 0, s>>>(grid, D, S5, slots, scale, grad_loss,
                                                                             grad_grid, part);
        NEAT_LAUNCHED();
        k_sum_partials<<<1, kRedThreads, 0, s>>>(part, nb, scale, loss);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    size_t neat_edge_loss_direct_workspace_bytes(int64_t num_active, int32_t grid_size)
    {
        if (num_active < 0 || grid_size < 2) return 0;
        return align256(neat_neighbor_pairs_workspace_bytes(num_active, grid_size)) + align256((size_t)kEdgeDirectBlocks * 8);
    }

    int neat_edge_loss_direct(neat_stream_t stream, const float* node_pos_min, const float* node_pos_max,
                              const uint8_t* node_flags, int32_t num_nodes, const int64_t* active_node_ids,
                              int64_t num_active, const int64_t* node_slot, const float* grid, int64_t N, int32_t C,
                              int32_t G, const int64_t* strides, float epsilon, uint64_t seed, double weight,
                              double* loss, const float* grad_loss, float* grad_grid, void* workspace,
                              size_t workspace_bytes)
    {
        NEAT_REQUIRE(node_pos_min && node_pos_max && node_flags && strides, "neat_edge_loss_direct: null pointer");
        NEAT_REQUIRE(num_active >= 0 && G >= 2 && C > 0 && N >= 0 && num_nodes > 0, "neat_edge_loss_direct: bad sizes");
        NEAT_REQUIRE(loss || grad_grid, "neat_edge_loss_direct: neither loss nor grad_grid requested");
        NEAT_REQUIRE(!grad_grid || grad_loss, "neat_edge_loss_direct: grad_grid needs grad_loss");
        const size_t need = neat_edge_loss_direct_workspace_bytes(num_active, G);
        NEAT_REQUIRE(workspace && workspace_bytes >= need, "neat_edge_loss_direct: workspace too small (%zu < %zu)",
                     workspace_bytes, need);
        hipStream_t s   = as_stream(stream);
        const int64_t n = num_active * 6 * (int64_t)G * G;
        if (n == 0 || N == 0)
        {
            if (loss) NEAT_HIP(hipMemsetAsync(loss, 0, sizeof(double), s));
            return NEAT_OK;
        }
        NEAT_REQUIRE(active_node_ids && node_slot && grid, "neat_edge_loss_direct: null pointer");
        NEAT_REQUIRE(strides[1] >= 0 && strides[2] >= 0 && strides[3] >= 0 && strides[4] >= 0 &&
                         (int64_t)(G - 1) * (strides[2] + strides[3] + strides[4]) + (int64_t)(C + 3) * strides[1] <
                             ((int64_t)1 << 31),
                     "neat_edge_loss_direct: a slot's extent must fit 32-bit offsets");
        char* w                 = static_cast<char*>(workspace);
        int32_t* other          = reinterpret_cast<int32_t*>(w);
        int64_t* row            = reinterpret_cast<int64_t*>(w + align256((size_t)n * 4));
        void* scan_ws           = w + align256((size_t)n * 4) + align256((size_t)(n + 1) * 8);
        const size_t scan_bytes = scan_workspace_bytes(n + 1);
        double* part = reinterpret_cast<double*>(w + align256(neat_neighbor_pairs_workspace_bytes(num_active, G)));
        launch_candidates(s, node_pos_min, node_pos_max, node_flags, num_nodes, active_node_ids, n, G, epsilon, other,
                          row);
        NEAT_LAUNCHED();
        int st = exclusive_scan_i64(s, row, n, scan_ws, scan_bytes);  // row[n] = number of pairs P
        if (st) return st;
        GridArgs g{grid, strides[0], strides[1], strides[2], strides[3], strides[4], N, C, G, G, G};
        const int64_t total = n * ((C + 3) / 4);  // lanes: (candidate, channel quad)
        const unsigned nb   = (unsigned)std::min<int64_t>(blocks_for(total, kRedThreads), kEdgeDirectBlocks);
        // 16-byte loads: channels contiguous, whole quads, every corner offset a multiple of 4 floats
        const bool vec4 = strides[1] == 1 && C % 4 == 0 && strides[0] % 4 == 0 && strides[2] % 4 == 0 &&
                          strides[3] % 4 == 0 && strides[4] % 4 == 0 && reinterpret_cast<uintptr_t>(grid) % 16 == 0;
        const bool small = total < ((int64_t)1 << 31);
        const bool wl = loss != nullptr, wg = grad_grid != nullptr;
        if (small && vec4)
            launch_edge_direct<uint32_t, true>(nb, s, wl, wg, g, node_pos_min, node_pos_max, active_node_ids, node_slot,
                                               n, G, other, row, seed, weight, grad_loss, grad_grid, part);
        else if (small)
            launch_edge_direct<uint32_t, false>(nb, s, wl, wg, g, node_pos_min, node_pos_max, active_node_ids,
                                                node_slot, n, G, other, row, seed, weight, grad_loss, grad_grid, part);
        else if (vec4)
            launch_edge_direct<int64_t, true>(nb, s, wl, wg, g, node_pos_min, node_pos_max, active_node_ids, node_slot,
                                              n, G, other, row, seed, weight, grad_loss, grad_grid, part);
        else
            launch_edge_direct<int64_t, false>(nb, s, wl, wg, g, node_pos_min, node_pos_max, active_node_ids, node_slot,
                                               n, G, other, row, seed, weight, grad_loss, grad_grid, part);
        NEAT_LAUNCHED();
        if (loss)
        {
            k_sum_partials_per_pair<<<1, kRedThreads, 0, s>>>(part, nb, row, n, weight, loss);
            NEAT_LAUNCHED();
        }
        return NEAT_OK;
    }
}
