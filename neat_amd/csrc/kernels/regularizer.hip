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
#include "common.hpp"

namespace neat
{

struct Strides5
{
    int64_t s[5];  // element strides of (N, C, Z, Y, X)
};

// ActiveNodeIdForGlobalPosition (HyperTree.cu:600-640): from the root down through inactive nodes into the first
// child (in child order) whose closed box contains p (BoxContainsPoint, HyperTree.h:85-95); -1 when p is in no
// active node.
__device__ __forceinline__ int32_t active_node_at(const float* pmin, const float* pmax, const uint8_t* flags,
                                                  int32_t first_leaf, const float p[3])
{
    int32_t n = 0;
    for (int depth = 0; depth < 10; ++depth)
    {
        if (flags[n] & 1) return n;
        if (n >= first_leaf) return -1;  // a leaf of the table that is not active: no children
        int32_t next = -1;
        for (int c = 0; c < 8; ++c)
        {
            const int32_t ch = 8 * n + 1 + c;
            bool in          = true;
#pragma unroll
            for (int d = 0; d < 3; ++d)
                if (pmin[3 * ch + d] > p[d] || pmax[3 * ch + d] < p[d]) in = false;
            if (in)
            {
                next = ch;
                break;
            }
        }
        if (next < 0) return -1;
        n = next;
    }
    return -1;
}

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

// candidate c = ((a * 3 + side) * 2 + front) * G^2 + sid -> neighbour node id, or -1 (no pair)
__global__ void __launch_bounds__(256) k_neighbor_candidates(const float* pmin, const float* pmax, const uint8_t* flags,
                                                             int32_t first_leaf, const int64_t* active_ids,
                                                             int64_t num_candidates, int32_t G, float eps,
                                                             int32_t* other, int64_t* is_pair)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= num_candidates) return;
    const int64_t G2   = (int64_t)G * G;
    const int32_t sid  = (int32_t)(c % G2);
    const int front    = (int)((c / G2) % 2);
    const int side     = (int)((c / (2 * G2)) % 3);
    const int32_t node = (int32_t)active_ids[c / (6 * G2)];
    float p[3];
    face_point(pmin, pmax, node, side, front, sid, G, p);
    float q[3] = {p[0], p[1], p[2]};
    q[side]    = p[side] + (front ? eps : -eps);
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

// block sum of one double per thread into out[blockIdx.x] (fixed tree order)
__device__ __forceinline__ void block_sum_to(double v, double* out)
{
    __shared__ double s[kRedThreads];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int h = kRedThreads / 2; h > 0; h >>= 1)
    {
        if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(kRedThreads) k_sum_partials(const double* partials, int64_t n, double scale,
                                                              double* out)
{
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) acc += partials[i];
    __shared__ double s[kRedThreads];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int h = kRedThreads / 2; h > 0; h >>= 1)
    {
        if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0] * scale;
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

// total variation: per thread one grid entry (a, c, z, y, x) of the active grids, the three forward differences that
// start there; scale per axis = 1 / (A * C * G * G * (G - 1)) (the means), times C * weight
__global__ void __launch_bounds__(kRedThreads) k_tv_fwd(const float* grid, int32_t C, int32_t G, Strides5 S,
                                                        const int64_t* slots, int64_t A, double* partials)
{
    const int64_t* st = S.s;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = A * C * G * G * G;
    double v        = 0.0;
    if (i < n)
    {
        const int32_t x = (int32_t)(i % G), y = (int32_t)((i / G) % G), z = (int32_t)((i / ((int64_t)G * G)) % G);
        const int32_t c = (int32_t)((i / ((int64_t)G * G * G)) % C);
        const int64_t a = i / ((int64_t)G * G * G * C);
        const float* g0 = grid + slots[a] * st[0] + c * st[1];
        const float v0  = g0[z * st[2] + y * st[3] + x * st[4]];
        float acc       = 0.f;
        if (x + 1 < G) acc += fabsf(g0[z * st[2] + y * st[3] + (x + 1) * st[4]] - v0);
        if (y + 1 < G) acc += fabsf(g0[z * st[2] + (y + 1) * st[3] + x * st[4]] - v0);
        if (z + 1 < G) acc += fabsf(g0[(z + 1) * st[2] + y * st[3] + x * st[4]] - v0);
        v = (double)acc;
    }
    block_sum_to(v, partials);
}

__device__ __forceinline__ float sgn(float d) { return d > 0.f ? 1.f : (d < 0.f ? -1.f : 0.f); }

// gradient of the total variation, accumulated into grad_grid: entry v gets sign(v - prev) - sign(next - v) per axis
__global__ void __launch_bounds__(256) k_tv_bwd(const float* grid, int32_t C, int32_t G, Strides5 S,
                                                const int64_t* slots, int64_t A, double weight, const float* grad_loss,
                                                float* grad_grid)
{
    const int64_t* st = S.s;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = A * C * G * G * G;
    if (i >= n) return;
    const int32_t x = (int32_t)(i % G), y = (int32_t)((i / G) % G), z = (int32_t)((i / ((int64_t)G * G)) % G);
    const int32_t c = (int32_t)((i / ((int64_t)G * G * G)) % C);
    const int64_t a = i / ((int64_t)G * G * G * C);
    const int64_t b = slots[a] * st[0] + c * st[1];
    const float* g0 = grid + b;
    const int64_t o = z * st[2] + y * st[3] + x * st[4];
    const float v0  = g0[o];
    float s         = 0.f;
    if (x > 0) s += sgn(v0 - g0[o - st[4]]);
    if (x + 1 < G) s -= sgn(g0[o + st[4]] - v0);
    if (y > 0) s += sgn(v0 - g0[o - st[3]]);
    if (y + 1 < G) s -= sgn(g0[o + st[3]] - v0);
    if (z > 0) s += sgn(v0 - g0[o - st[2]]);
    if (z + 1 < G) s -= sgn(g0[o + st[2]] - v0);
    const double scale = (double)*grad_loss * weight * C / ((double)A * C * G * G * (G - 1));
    grad_grid[b + o] += (float)(scale * s);
}

}  // namespace neat

using namespace neat;

static int64_t blocks_for(int64_t n, int t) { return (n + t - 1) / t; }
static size_t align256(size_t b) { return (b + 255) / 256 * 256; }
static Strides5 strides5(const int64_t* st) { return Strides5{{st[0], st[1], st[2], st[3], st[4]}}; }

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
        const int32_t first_leaf = (num_nodes - 1) / 8;
        if (n == 0)
        {
            NEAT_HIP(hipMemsetAsync(num_pairs, 0, sizeof(int64_t), s));
            return NEAT_OK;
        }
        NEAT_REQUIRE(active_node_ids && out_coord && out_node, "neat_neighbor_pairs: null pointer");
        k_neighbor_candidates<<<(unsigned)blocks_for(n, 256), 256, 0, s>>>(
            node_pos_min, node_pos_max, node_flags, first_leaf, active_node_ids, n, grid_size, epsilon, other, row);
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
        const int64_t n  = num_active * C * (int64_t)G * G * G;
        const int64_t nb = blocks_for(n, kRedThreads);
        NEAT_REQUIRE(workspace && workspace_bytes >= neat_reg_workspace_bytes(n), "neat_tv_loss_fwd: workspace too small");
        double* part = static_cast<double*>(workspace);
        k_tv_fwd<<<(unsigned)nb, kRedThreads, 0, s>>>(grid, C, G, strides5(strides), slots, num_active, part);
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
        const int64_t n = num_active * C * (int64_t)G * G * G;
        k_tv_bwd<<<(unsigned)blocks_for(n, 256), 256, 0, as_stream(stream)>>>(grid, C, G, strides5(strides), slots,
                                                                               num_active, weight, grad_loss, grad_grid);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }
}
