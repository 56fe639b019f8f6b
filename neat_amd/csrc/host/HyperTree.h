// HyperTreeBase: the octree module of the projector path (HyperTree.h:98-243 in the reference) with the same
// buffer names, dtypes and method signatures, on the C ABI's sampler and coordinate kernels.
#pragma once

#include "HelperStructs.h"
#include "neat_host.h"

class HyperTreeBaseImpl : public torch::nn::Module
{
   public:
    // Complete octree of depth max_depth over [-1,1]^d (only d = 3), all of depth 0 active
    // (HyperTreeBaseImpl::HyperTreeBaseImpl, HyperTree.cpp:11-132).
    HyperTreeBaseImpl(int d, int max_depth);

    // Sets all nodes of layer `depth` active, all others inactive (HyperTree.cpp:163-168).
    void SetActive(int depth);
    // Rebuilds active_node_ids / node_active_prefix_sum from node_active (HyperTree.cpp:169-206) and the
    // traversal flags the sampler walks.
    void UpdateActive();

    // Ray x active-leaf intersection and per-segment samples (HyperTree.h:157, HyperTree.cu:393-438), in canonical
    // order. interval_jitter draws each sample's in-cell position from a counter hash seeded from torch's default
    // generator; without it the cell midpoints are used.
    SampleList CreateSamplesForRays(const RayList& rays, int max_samples_per_node, bool interval_jitter);

    // Only the canonical segment list (what the fused projector consumes); SampleList::size() is the sample count.
    SampleList CreateSegmentsForRays(const RayList& rays, int max_samples_per_node, bool interval_jitter);

    // Face-point pairs between neighbouring active nodes (HyperTree.cu:830-940) as a SampleList of two rows per pair,
    // in canonical order (active node, side, front, face point). `size` holds the per-axis face resolution (the
    // grid size; cubic only), epsilon the offset that finds the neighbour.
    SampleList NodeNeighborSamples(const std::vector<int>& size, double epsilon);
    // voxels of a size[0] x size[1] x size[2] lattice over [-1, 1]^3 with their active node (HyperTree.cu:547-597)
    SampleList UniformPhantomSamplesGPU(const std::vector<int>& size, bool swap_xy);

    // u = (g - min)/(max - min) * 2 - 1 for float [N, D] or [groups, group_size, D] samples, differentiable with
    // respect to global_samples (HyperTree.h:134, HyperTree.cu:116-199).
    torch::Tensor ComputeLocalSamples(torch::Tensor global_samples, torch::Tensor node_indices);

    // node_error update from a per-node error (HyperTree.cpp:225-265): "override" replaces the active nodes' errors,
    // "min" keeps the smaller of the old (where already set) and the new one; inactive nodes keep theirs.
    void SetErrorForActiveNodes(torch::Tensor error, std::string strategy = "override");

    // node id -> grid slot (HyperTree.h:183-187).
    torch::Tensor GlobalNodeIdToLocalActiveId(torch::Tensor node_id);

    torch::Tensor ActiveNodeTensor() { return active_node_ids; }
    torch::Tensor InactiveNodeIds();
    torch::Tensor GetNodePositionScaleForId(torch::Tensor node_ids);

    torch::Device device() { return node_parent.device(); }
    int NumNodes() { return node_parent.size(0); }
    int NumActiveNodes() { return active_node_ids.size(0); }
    int D() { return node_position_min.size(1); }
    int NS() { return node_children.size(1); }
    int MaxDepth() { return max_depth_; }

    // Moves buffers, including the traversal flags.
    using torch::nn::Module::to;
    void to(torch::Device device, bool non_blocking = false) override;

    // int [num_nodes], -1 for the root
    torch::Tensor node_parent;
    // int [num_nodes, NS], -1 for leaves
    torch::Tensor node_children;
    // float [num_nodes, D]
    torch::Tensor node_position_min, node_position_max;
    // float [num_nodes]
    torch::Tensor node_scale;
    // float [num_nodes] (-1 = not computed)
    torch::Tensor node_error;
    torch::Tensor node_max_density;
    // int [num_nodes]
    torch::Tensor node_depth;
    // int [num_nodes] 0/1
    torch::Tensor node_active;
    torch::Tensor node_culled;
    // long [num_nodes]: slot of active nodes, -1 otherwise
    torch::Tensor node_active_prefix_sum;
    // long [num_active_nodes]
    torch::Tensor active_node_ids;
    // uint8 [num_nodes]: bit 0 active, bit 1 active strict descendant (not a registered buffer)
    torch::Tensor node_flags;

   private:
    int max_depth_;
    SampleList create_samples(const RayList& rays, int max_samples_per_node, bool interval_jitter, bool emit);
};

TORCH_MODULE(HyperTreeBase);
