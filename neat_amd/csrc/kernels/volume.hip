// Uniform volume samples of the reference's volume export (HyperTreeBaseImpl::UniformPhantomSamplesGPU,
// HyperTree.cu:547-597, with NodeIdForPositionGPU, :643-686): one thread per voxel of a size[0] x size[1] x size[2]
// lattice spanning [-1, 1]^3 corner to corner, dimension 0 fastest (the last fastest with swap_xy), its coordinate
// c / (size - 1) * 2 - 1, and the active node containing it; a voxel in no active node gets the first active node
// and weight 0, as there.
#include "tree_dev.hpp"

namespace neat
{

__global__ void __launch_bounds__(256) k_volume_samples(const float* pmin, const float* pmax, const uint8_t* flags,
                                                        int32_t first_leaf, int64_t first_active, int32_t s0,
                                                        int32_t s1, int32_t s2, int32_t swap_xy, int64_t n,
                                                        float* out_coord, int64_t* out_index, int64_t* out_node,
                                                        float* out_weight)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n) return;
    const int32_t size[3] = {s0, s1, s2};
    int32_t c[3];
    int64_t tmp = tid;
    if (swap_xy)
    {
        for (int d = 2; d >= 0; --d)
        {
            c[d] = (int32_t)(tmp % size[d]);
            tmp /= size[d];
        }
    }
    else
    {
        for (int d = 0; d < 3; ++d)
        {
            c[d] = (int32_t)(tmp % size[d]);
            tmp /= size[d];
        }
    }
    float p[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        const float pos = (float)c[d] / ((float)size[d] - 1.f);
        p[d]            = pos * 2 - 1.f;
        out_coord[3 * tid + d] = p[d];
    }
    out_index[tid]  = tid;
    const int32_t o = active_node_at(pmin, pmax, flags, first_leaf, p);
    out_node[tid]   = o < 0 ? first_active : (int64_t)o;
    out_weight[tid] = o < 0 ? 0.f : 1.f;
}

}  // namespace neat

using namespace neat;

extern "C"
{
    int neat_uniform_volume_samples(neat_stream_t stream, const float* node_pos_min, const float* node_pos_max,
                                    const uint8_t* node_flags, int32_t num_nodes, int64_t first_active_node,
                                    const int32_t* size, int32_t swap_xy, float* out_coord, int64_t* out_index,
                                    int64_t* out_node, float* out_weight)
    {
        NEAT_REQUIRE(size && size[0] >= 2 && size[1] >= 2 && size[2] >= 2 && num_nodes > 0,
                     "neat_uniform_volume_samples: every size must be >= 2");
        NEAT_REQUIRE(node_pos_min && node_pos_max && node_flags && out_coord && out_index && out_node && out_weight,
                     "neat_uniform_volume_samples: null pointer");
        const int64_t n = (int64_t)size[0] * size[1] * size[2];
        k_volume_samples<<<grid_for(n, 256), 256, 0, as_stream(stream)>>>(
            node_pos_min, node_pos_max, node_flags, (num_nodes - 1) / 8, first_active_node, size[0], size[1], size[2],
            swap_xy, n, out_coord, out_index, out_node, out_weight);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }
}
