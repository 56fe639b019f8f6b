// Error plumbing, ABI version, host-side octree helpers and the int64 exclusive scan.
#include <cstdarg>
#include <cstring>

#include "common.hpp"

namespace neat
{

static thread_local char g_last_error[1024] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

int fail(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

// ---- exclusive scan -----------------------------------------------------------------------------------------------
// Three passes over tiles of 1024 elements (256 threads x 4): tile sums, a scan of the tile sums by one workgroup,
// then each tile rescans itself with its offset. Deterministic (integer) and in place.
constexpr int kScanThreads = 256;
constexpr int kScanItems   = 4;
constexpr int kScanTile    = kScanThreads * kScanItems;

__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t* lds, int64_t* total)
{
    int tid = threadIdx.x;
    lds[tid] = v;
    __syncthreads();
    // Hillis-Steele inclusive scan over 256 entries
    for (int off = 1; off < kScanThreads; off <<= 1)
    {
        int64_t add = tid >= off ? lds[tid - off] : 0;
        __syncthreads();
        lds[tid] += add;
        __syncthreads();
    }
    int64_t incl = lds[tid];
    *total       = lds[kScanThreads - 1];
    __syncthreads();
    return incl - v;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_tile_sums(const int64_t* a, int64_t n, int64_t* tile_sums)
{
    __shared__ int64_t lds[kScanThreads];
    int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int64_t s    = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) s += a[base + k];
    int64_t total;
    block_exclusive_scan(s, lds, &total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// One workgroup scans the tile sums (loops over chunks of 1024) into exclusive tile offsets.
__global__ void __launch_bounds__(kScanThreads) k_scan_tile_offsets(int64_t* tile_sums, int64_t num_tiles)
{
    __shared__ int64_t lds[kScanThreads];
    int64_t carry = 0;
    for (int64_t base = 0; base < num_tiles; base += kScanTile)
    {
        int64_t idx = base + threadIdx.x * kScanItems;
        int64_t v[kScanItems];
        int64_t s = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
        {
            v[k] = (idx + k < num_tiles) ? tile_sums[idx + k] : 0;
            s += v[k];
        }
        int64_t total;
        int64_t ex = block_exclusive_scan(s, lds, &total) + carry;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
        {
            if (idx + k < num_tiles) tile_sums[idx + k] = ex;
            ex += v[k];
        }
        carry += total;
    }
}

__global__ void __launch_bounds__(kScanThreads) k_scan_apply(int64_t* a, int64_t n, const int64_t* tile_offsets)
{
    __shared__ int64_t lds[kScanThreads];
    int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int64_t v[kScanItems];
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
    {
        v[k] = (base + k < n) ? a[base + k] : 0;
        s += v[k];
    }
    int64_t total;
    int64_t ex = block_exclusive_scan(s, lds, &total) + tile_offsets[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
    {
        if (base + k < n) a[base + k] = ex;
        ex += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) a[n] = tile_offsets[blockIdx.x] + total;
}

__global__ void k_fill_zero_i64(int64_t* a) { a[0] = 0; }

size_t scan_workspace_bytes(int64_t n) { return (size_t)((n + kScanTile - 1) / kScanTile + 1) * sizeof(int64_t); }

int exclusive_scan_i64(hipStream_t stream, int64_t* a, int64_t n, void* workspace, size_t workspace_bytes)
{
    if (n == 0)
    {
        k_fill_zero_i64<<<1, 1, 0, stream>>>(a);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }
    int64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (workspace_bytes < scan_workspace_bytes(n) || workspace == nullptr)
        return fail(NEAT_E_OVERFLOW, "scan workspace too small: %zu < %zu", workspace_bytes, scan_workspace_bytes(n));
    int64_t* tile = static_cast<int64_t*>(workspace);
    k_scan_tile_sums<<<(unsigned)tiles, kScanThreads, 0, stream>>>(a, n, tile);
    NEAT_LAUNCHED();
    k_scan_tile_offsets<<<1, kScanThreads, 0, stream>>>(tile, tiles);
    NEAT_LAUNCHED();
    k_scan_apply<<<(unsigned)tiles, kScanThreads, 0, stream>>>(a, n, tile);
    NEAT_LAUNCHED();
    return NEAT_OK;
}

}  // namespace neat

using namespace neat;

extern "C"
{
    const char* neat_last_error(void) { return g_last_error; }

    int neat_abi_version(void) { return 1; }

// the build passes the source hash as a bare token (id + 16 hex digits), stringified here
#ifndef NEAT_BUILD_ID
#define NEAT_BUILD_ID idunknown
#endif
#define NEAT_STR2_(x) #x
#define NEAT_STR_(x) NEAT_STR2_(x)
    const char* neat_build_id(void) { return NEAT_STR_(NEAT_BUILD_ID) + 2; }

    int64_t neat_tree_num_nodes(int32_t max_depth)
    {
        if (max_depth < 0 || max_depth > 9) return -1;
        int64_t n = 0, p = 1;
        for (int d = 0; d <= max_depth; ++d, p *= 8) n += p;
        return n;
    }

    // Breadth-first dyadic octree over [-1,1]^3: node n's child i is 8n+1+i and takes the upper half on axis k iff
    // bit k of i is set; scale = depth / max_depth * 2 - 1 (HyperTreeBaseImpl::HyperTreeBaseImpl,
    // HyperTree.cpp:11-132). Children are -1 on the deepest level.
    int neat_tree_build_host(int32_t max_depth, int32_t* parent, int32_t* children, float* pos_min, float* pos_max,
                             float* scale, int32_t* depth)
    {
        NEAT_REQUIRE(max_depth >= 0 && max_depth <= 9, "neat_tree_build_host: max_depth %d outside [0, 9]",
                     max_depth);
        NEAT_REQUIRE(parent && children && pos_min && pos_max && scale && depth, "neat_tree_build_host: null pointer");
        const int64_t num_nodes = neat_tree_num_nodes(max_depth);
        for (int64_t i = 0; i < num_nodes * 8; ++i) children[i] = -1;
        auto set_node = [&](int64_t id, int32_t par, const float* mn, const float* mx, int32_t d)
        {
            float s = max_depth == 0 ? 0.f : ((float)d / (float)max_depth) * 2.f - 1.f;
            for (int k = 0; k < 3; ++k)
            {
                pos_min[3 * id + k] = mn[k];
                pos_max[3 * id + k] = mx[k];
            }
            parent[id] = par;
            depth[id]  = d;
            scale[id]  = s;
        };
        const float rmin[3] = {-1.f, -1.f, -1.f}, rmax[3] = {1.f, 1.f, 1.f};
        set_node(0, -1, rmin, rmax, 0);
        int64_t level_begin = 0, level_size = 1;
        for (int32_t d = 1; d <= max_depth; ++d)
        {
            for (int64_t n = level_begin; n < level_begin + level_size; ++n)
            {
                for (int i = 0; i < 8; ++i)
                {
                    float mn[3], mx[3];
                    for (int k = 0; k < 3; ++k)
                    {
                        float lo = pos_min[3 * n + k], hi = pos_max[3 * n + k], c = (lo + hi) * 0.5f;
                        mn[k] = ((i >> k) & 1) ? c : lo;
                        mx[k] = ((i >> k) & 1) ? hi : c;
                    }
                    int64_t child         = 8 * n + 1 + i;
                    children[8 * n + i]   = (int32_t)child;
                    set_node(child, (int32_t)n, mn, mx, d);
                }
            }
            level_begin += level_size;
            level_size *= 8;
        }
        return NEAT_OK;
    }

    // active_node_ids (ascending) and node_active_prefix_sum (slot or -1) from node_active (UpdateActive,
    // HyperTree.cpp:169-206). Returns the number of active nodes, or -1 on a null pointer.
    int64_t neat_tree_update_active_host(const int32_t* node_active, int32_t num_nodes, int64_t* active_ids,
                                         int64_t* prefix)
    {
        if (!node_active || !active_ids || !prefix || num_nodes < 0)
        {
            fail(NEAT_E_ARG, "neat_tree_update_active_host: bad arguments");
            return -1;
        }
        int64_t count = 0;
        for (int32_t i = 0; i < num_nodes; ++i)
        {
            if (node_active[i] == 1)
            {
                active_ids[count] = i;
                prefix[i]         = count++;
            }
            else
            {
                prefix[i] = -1;
            }
        }
        return count;
    }

    int neat_tree_flags_host(const int32_t* node_active, int32_t num_nodes, int32_t max_depth, uint8_t* node_flags)
    {
        NEAT_REQUIRE(node_active && node_flags, "neat_tree_flags_host: null pointer");
        NEAT_REQUIRE(max_depth >= 0 && max_depth <= 9, "neat_tree_flags_host: max_depth %d outside [0, 9]", max_depth);
        int64_t expect = 0, p = 1;
        for (int d = 0; d <= max_depth; ++d, p *= 8) expect += p;
        NEAT_REQUIRE(num_nodes == expect, "neat_tree_flags_host: num_nodes %d != %lld for depth %d", num_nodes,
                     (long long)expect, max_depth);
        for (int32_t n = 0; n < num_nodes; ++n) node_flags[n] = node_active[n] == 1 ? 1 : 0;
        // bottom-up: a node descends if any child is active or descends (children of n are 8n+1..8n+8)
        int64_t first_leaf = (expect - p / 8);  // first node of the deepest level
        for (int64_t n = first_leaf - 1; n >= 0; --n)
        {
            bool any = false;
            for (int i = 0; i < 8; ++i) any |= node_flags[8 * n + 1 + i] != 0;
            if (any) node_flags[n] |= 2;
        }
        return NEAT_OK;
    }

    int neat_tree_validate_host(const float* pmin, const float* pmax, const int32_t* children, int32_t num_nodes,
                                int32_t max_depth)
    {
        NEAT_REQUIRE(pmin && pmax && children, "neat_tree_validate_host: null pointer");
        NEAT_REQUIRE(max_depth >= 0 && max_depth <= 9, "neat_tree_validate_host: max_depth %d outside [0, 9]",
                     max_depth);
        int64_t expect = 0, p = 1;
        for (int d = 0; d <= max_depth; ++d, p *= 8) expect += p;
        NEAT_REQUIRE(num_nodes == expect, "neat_tree_validate_host: num_nodes %d != %lld", num_nodes,
                     (long long)expect);
        for (int k = 0; k < 3; ++k)
            NEAT_REQUIRE(pmin[k] == -1.f && pmax[k] == 1.f, "neat_tree_validate_host: root box is not [-1,1]^3");
        int64_t first_leaf = expect - p / 8;
        for (int64_t n = 0; n < num_nodes; ++n)
        {
            for (int i = 0; i < 8; ++i)
            {
                int32_t c = children[8 * n + i];
                if (n >= first_leaf)
                {
                    NEAT_REQUIRE(c == -1, "neat_tree_validate_host: leaf %lld has child %d", (long long)n, c);
                    continue;
                }
                NEAT_REQUIRE(c == 8 * n + 1 + i, "neat_tree_validate_host: child %d of node %lld is %d, not %lld", i,
                             (long long)n, c, (long long)(8 * n + 1 + i));
                for (int k = 0; k < 3; ++k)
                {
                    float center = (pmin[3 * n + k] + pmax[3 * n + k]) * 0.5f;
                    float emn    = ((i >> k) & 1) ? center : pmin[3 * n + k];
                    float emx    = ((i >> k) & 1) ? pmax[3 * n + k] : center;
                    NEAT_REQUIRE(pmin[3 * c + k] == emn && pmax[3 * c + k] == emx,
                                 "neat_tree_validate_host: node %d box differs from its parent's split", c);
                }
            }
        }
        return NEAT_OK;
    }
}
