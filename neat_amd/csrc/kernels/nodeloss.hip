// Per-node loss accumulation of the reference's evaluation pass (HierarchicalNeuralGeometry::AccumulateSampleLossPerNode,
// geometry.cpp:38-58, called from reconstruct.cpp:589-590): every sample's per-ray loss times its weight, and the
// weight itself, summed per octree node. Here over the canonical segment list of a uniformly sampled batch, with the
// samples' weights recomputed exactly as the sampler emits them (sample_w): per segment the two sums run in FP32 in
// j order, then per node in FP64 over the node's segments in canonical order (a stable sort by node), so the result
// is deterministic where the reference's scatter_add follows atomic order.
#include <hipcub/hipcub.hpp>

#include "common.hpp"

namespace neat
{

__global__ void __launch_bounds__(256) k_seg_loss(const int64_t* ray_seg_offset, const int32_t* seg_node,
                                                  const int32_t* seg_ray, const int32_t* seg_count,
                                                  const float* seg_tmin, const float* seg_tmax, int64_t K,
                                                  int32_t jitter, uint64_t seed, const float* per_ray_loss,
                                                  float* seg_lw, float* seg_w, int32_t* keys, int32_t* vals,
                                                  int64_t* node_count)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int64_t ray  = seg_ray[k];
    const int32_t rank = (int32_t)(k - ray_seg_offset[ray]);
    const int32_t n    = seg_count[k];
    const float tmin = seg_tmin[k], tmax = seg_tmax[k];
    const float step = (tmax - tmin) / (float)n;
    const float loss = per_ray_loss[ray];
    float lw = 0.f, ws = 0.f;
    for (int32_t j = 0; j < n; ++j)
    {
        const float t = sample_t(tmin, step, j, seg_a(jitter, seed, ray, rank, j));
        const float w = sample_w(tmin, tmax, step, n, j, t, jitter, seed, ray, rank);
        lw += loss * w;
        ws += w;
    }
    seg_lw[k] = lw;
    seg_w[k]  = ws;
    keys[k]   = seg_node[k];
    vals[k]   = (int32_t)k;
    atomicAdd(reinterpret_cast<unsigned long long*>(node_count) + seg_node[k], 1ull);  // counts: order-free
}

__global__ void __launch_bounds__(256) k_node_sum(const int64_t* node_off, int32_t num_nodes, const int32_t* sorted_k,
                                                  const float* seg_lw, const float* seg_w, float* loss_sum,
                                                  float* weight_sum)
{
    const int32_t nd = blockIdx.x * blockDim.x + threadIdx.x;
    if (nd >= num_nodes) return;
    double l = 0.0, w = 0.0;
    for (int64_t i = node_off[nd]; i < node_off[nd + 1]; ++i)
    {
        const int32_t k = sorted_k[i];
        l += (double)seg_lw[k];
        w += (double)seg_w[k];
    }
    loss_sum[nd]   = (float)l;
    weight_sum[nd] = (float)w;
}

}  // namespace neat

using namespace neat;

namespace
{
size_t a256(size_t b) { return (b + 255) / 256 * 256; }

size_t sort_temp_bytes(int64_t K)
{
    size_t tmp = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (int32_t*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr,
                                       (int32_t*)nullptr, (int)std::max<int64_t>(K, 1), 0, 32);
    return tmp;
}
}  // namespace

extern "C"
{
    size_t neat_node_loss_workspace_bytes(int64_t num_segments, int32_t num_nodes)
    {
        if (num_segments < 0 || num_nodes <= 0) return 0;
        const size_t K = (size_t)std::max<int64_t>(num_segments, 1);
        return 6 * a256(K * 4) + a256((size_t)(num_nodes + 1) * 8) + a256(sort_temp_bytes(num_segments)) +
               a256(scan_workspace_bytes(num_nodes));
    }

    int neat_node_loss_accumulate(neat_stream_t stream, const int64_t* ray_seg_offset, int64_t num_rays,
                                  const int32_t* seg_node, const int32_t* seg_ray, const int32_t* seg_count,
                                  const float* seg_tmin, const float* seg_tmax, int64_t num_segments, int32_t jitter,
                                  uint64_t seed, const float* per_ray_loss, int32_t num_nodes, float* loss_sum,
                                  float* weight_sum, void* workspace, size_t workspace_bytes)
    {
        NEAT_REQUIRE(num_rays >= 0 && num_segments >= 0 && num_nodes > 0, "neat_node_loss_accumulate: bad sizes");
        NEAT_REQUIRE(loss_sum && weight_sum, "neat_node_loss_accumulate: null output");
        const size_t need = neat_node_loss_workspace_bytes(num_segments, num_nodes);
        NEAT_REQUIRE(workspace && workspace_bytes >= need, "neat_node_loss_accumulate: workspace too small (%zu < %zu)",
                     workspace_bytes, need);
        hipStream_t s = as_stream(stream);
        const int64_t K = num_segments;
        const size_t KB = a256((size_t)std::max<int64_t>(K, 1) * 4);
        char* w         = static_cast<char*>(workspace);
        float* seg_lw   = reinterpret_cast<float*>(w);
        float* seg_w    = reinterpret_cast<float*>(w + KB);
        int32_t* keys   = reinterpret_cast<int32_t*>(w + 2 * KB);
        int32_t* vals   = reinterpret_cast<int32_t*>(w + 3 * KB);
        int32_t* skeys  = reinterpret_cast<int32_t*>(w + 4 * KB);
        int32_t* svals  = reinterpret_cast<int32_t*>(w + 5 * KB);
        int64_t* off    = reinterpret_cast<int64_t*>(w + 6 * KB);
        char* sort_tmp  = w + 6 * KB + a256((size_t)(num_nodes + 1) * 8);
        size_t sort_sz  = sort_temp_bytes(K);
        void* scan_tmp  = sort_tmp + a256(sort_sz);
        NEAT_HIP(hipMemsetAsync(off, 0, (size_t)(num_nodes + 1) * 8, s));
        if (K > 0)
        {
            NEAT_REQUIRE(ray_seg_offset && seg_node && seg_ray && seg_count && seg_tmin && seg_tmax && per_ray_loss,
                         "neat_node_loss_accumulate: null pointer");
            k_seg_loss<<<(unsigned)((K + 255) / 256), 256, 0, s>>>(ray_seg_offset, seg_node, seg_ray, seg_count,
                                                                   seg_tmin, seg_tmax, K, jitter, seed, per_ray_loss,
                                                                   seg_lw, seg_w, keys, vals, off);
            NEAT_LAUNCHED();
            int end_bit = 1;
            while (((int64_t)1 << end_bit) <= num_nodes) ++end_bit;
            NEAT_HIP(hipcub::DeviceRadixSort::SortPairs(sort_tmp, sort_sz, keys, skeys, vals, svals, (int)K, 0, end_bit,
                                                        s));
        }
        int st = exclusive_scan_i64(s, off, num_nodes, scan_tmp, scan_workspace_bytes(num_nodes));
        if (st) return st;
        k_node_sum<<<(unsigned)((num_nodes + 255) / 256), 256, 0, s>>>(off, num_nodes, svals, seg_lw, seg_w, loss_sum,
                                                                      weight_sum);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }
}
