#include "HyperTree.h"

#include <torch/csrc/autograd/custom_function.h>

using namespace neat_host;
using torch::autograd::AutogradContext;
using torch::autograd::tensor_list;

namespace
{

// (g - min)/(max - min) * 2 - 1 with its backward (grad * 2) / (max - min), on the C ABI.
struct LocalSamplesFn : public torch::autograd::Function<LocalSamplesFn>
{
    static torch::Tensor forward(AutogradContext* ctx, torch::Tensor global, torch::Tensor node, torch::Tensor pmin,
                                 torch::Tensor pmax)
    {
        auto g   = global.contiguous();
        auto n   = node.to(torch::kLong).contiguous();
        auto out = torch::empty_like(g);
        NEAT_CALL(neat_local_samples(current_stream(), dptr<float>(g), dptr<int64_t>(n), n.numel(), dptr<float>(pmin),
                                     dptr<float>(pmax), (int32_t)pmin.size(0), dptr<float>(out)));
        ctx->save_for_backward({n, pmin, pmax});
        return out;
    }

    static tensor_list backward(AutogradContext* ctx, tensor_list grads)
    {
        auto saved = ctx->get_saved_variables();
        auto gl    = grads[0].contiguous();
        auto gg    = torch::empty_like(gl);
        NEAT_CALL(neat_local_samples_bwd(current_stream(), dptr<float>(gl), dptr<int64_t>(saved[0]), saved[0].numel(),
                                         dptr<float>(saved[1]), dptr<float>(saved[2]), (int32_t)saved[1].size(0),
                                         dptr<float>(gg)));
        return {gg, torch::Tensor(), torch::Tensor(), torch::Tensor()};
    }
};

}  // namespace

HyperTreeBaseImpl::HyperTreeBaseImpl(int d, int max_depth) : max_depth_(max_depth)
{
    TORCH_CHECK(d == 3, "HyperTreeBase: only 3-D octrees are supported (got d = ", d, ")");
    const int64_t n = neat_tree_num_nodes(max_depth);
    TORCH_CHECK(n > 0, "HyperTreeBase: max_depth ", max_depth, " outside [0, 9]");
    node_parent            = torch::empty({n}, torch::kInt32);
    node_children          = torch::empty({n, 8}, torch::kInt32);
    node_position_min      = torch::empty({n, 3}, torch::kFloat32);
    node_position_max      = torch::empty({n, 3}, torch::kFloat32);
    node_scale             = torch::empty({n}, torch::kFloat32);
    node_depth             = torch::empty({n}, torch::kInt32);
    node_error             = -torch::ones({n}, torch::kFloat32);
    node_max_density       = -torch::ones({n}, torch::kFloat32);
    node_active            = torch::zeros({n}, torch::kInt32);
    node_culled            = torch::zeros({n}, torch::kInt32);
    node_active_prefix_sum = torch::zeros({n}, torch::kLong);
    active_node_ids        = torch::zeros({1}, torch::kLong);
    NEAT_CALL(neat_tree_build_host(max_depth, dptr<int32_t>(node_parent), dptr<int32_t>(node_children),
                                   dptr<float>(node_position_min), dptr<float>(node_position_max),
                                   dptr<float>(node_scale), dptr<int32_t>(node_depth)));
    NEAT_CALL(neat_tree_validate_host(dptr<float>(node_position_min), dptr<float>(node_position_max),
                                      dptr<int32_t>(node_children), (int32_t)n, max_depth));
    // buffer names as registered by the reference (HyperTree.cpp:43-57), so state dicts line up
    register_buffer("node_parent", node_parent);
    register_buffer("node_children", node_children);
    register_buffer("node_position_min", node_position_min);
    register_buffer("node_position_max", node_position_max);
    register_buffer("node_scale", node_scale);
    register_buffer("node_error", node_error);
    register_buffer("node_max_density", node_max_density);
    register_buffer("node_active", node_active);
    register_buffer("node_culled", node_culled);
    register_buffer("node_depth", node_depth);
    register_buffer("node_active_id", node_active_prefix_sum);
    SetActive(0);
    register_buffer("active_node_ids", active_node_ids);
}

void HyperTreeBaseImpl::SetActive(int depth)
{
    node_active.set_data((node_depth == depth).to(torch::kInt32));
    UpdateActive();
}

void HyperTreeBaseImpl::UpdateActive()
{
    auto active = node_active.to(torch::kCPU).to(torch::kInt32).contiguous();
    auto culled = node_culled.to(torch::kCPU).to(torch::kInt32);
    TORCH_CHECK((active * culled).sum().item<int64_t>() == 0, "a culled node is not allowed to be active");
    const int32_t n = (int32_t)active.size(0);
    auto ids        = torch::empty({n}, torch::kLong);
    auto prefix     = torch::empty({n}, torch::kLong);
    int64_t count   = neat_tree_update_active_host(dptr<int32_t>(active), n, dptr<int64_t>(ids), dptr<int64_t>(prefix));
    TORCH_CHECK(count >= 0, neat_last_error());
    auto flags = torch::empty({n}, torch::kUInt8);
    NEAT_CALL(neat_tree_flags_host(dptr<int32_t>(active), n, max_depth_, flags.data_ptr<uint8_t>()));
    const auto dev = node_parent.device();
    // set_data keeps the registered buffers valid
    active_node_ids.set_data(ids.slice(0, 0, count).clone().to(dev));
    node_active_prefix_sum.set_data(prefix.to(dev));
    node_flags = flags.to(dev);
}

void HyperTreeBaseImpl::to(torch::Device device, bool non_blocking)
{
    torch::nn::Module::to(device, non_blocking);
    if (node_flags.defined()) node_flags = node_flags.to(device);
}

torch::Tensor HyperTreeBaseImpl::InactiveNodeIds()
{
    auto a = node_active.to(torch::kCPU);
    return torch::nonzero(a == 0).reshape({-1}).to(torch::kLong).to(device());
}

torch::Tensor HyperTreeBaseImpl::GetNodePositionScaleForId(torch::Tensor node_ids)
{
    auto pos = torch::index_select(node_position_min, 0, node_ids);
    auto sca = torch::index_select(node_scale, 0, node_ids).unsqueeze(1);
    return torch::cat({pos, sca}, 1);
}

torch::Tensor HyperTreeBaseImpl::GlobalNodeIdToLocalActiveId(torch::Tensor node_id)
{
    check_cuda(node_id, "node_id");
    auto n   = node_id.reshape({-1}).to(torch::kLong).contiguous();
    auto out = torch::empty_like(n);
    NEAT_CALL(neat_node_to_slot(current_stream(), dptr<int64_t>(n), n.numel(), dptr<int64_t>(node_active_prefix_sum),
                                (int32_t)NumNodes(), dptr<int64_t>(out)));
    return out.reshape(node_id.sizes());
}

torch::Tensor HyperTreeBaseImpl::ComputeLocalSamples(torch::Tensor global_samples, torch::Tensor node_indices)
{
    check_cuda(global_samples, "global_samples");
    check_cuda(node_position_min, "the octree");
    torch::Tensor node = node_indices;
    if (global_samples.dim() == 3 && node_indices.dim() == 1)
    {
        // samples batched by node: [groups, group_size, D] with one node id per group
        node = node_indices.unsqueeze(1).expand({global_samples.size(0), global_samples.size(1)}).reshape({-1});
    }
    auto flat = global_samples.reshape({-1, global_samples.size(-1)});
    TORCH_CHECK(flat.size(1) == 3 && node.numel() == flat.size(0), "ComputeLocalSamples: shape mismatch");
    auto out = LocalSamplesFn::apply(flat, node.reshape({-1}), node_position_min, node_position_max);
    return out.reshape(global_samples.sizes());
}

SampleList HyperTreeBaseImpl::CreateSamplesForRays(const RayList& rays, int max_samples_per_node, bool interval_jitter)
{
    return create_samples(rays, max_samples_per_node, interval_jitter, true);
}

SampleList HyperTreeBaseImpl::CreateSegmentsForRays(const RayList& rays, int max_samples_per_node, bool interval_jitter)
{
    return create_samples(rays, max_samples_per_node, interval_jitter, false);
}

SampleList HyperTreeBaseImpl::create_samples(const RayList& rays, int max_spn, bool jitter, bool emit)
{
    check_cuda(rays.direction, "rays.direction");
    check_cuda(node_position_min, "the octree");
    TORCH_CHECK(rays.origin.size(1) == 3 && rays.direction.size(1) == 3, "rays must be [R, 3]");
    const auto dev   = node_position_min.device();
    const auto L     = torch::TensorOptions().dtype(torch::kLong).device(dev);
    const auto I     = torch::TensorOptions().dtype(torch::kInt32).device(dev);
    const auto F     = torch::TensorOptions().dtype(torch::kFloat32).device(dev);
    auto o           = rays.origin.detach().to(torch::kFloat32).contiguous();
    auto d           = rays.direction.detach().to(torch::kFloat32).contiguous();
    const int64_t R  = o.size(0);
    auto s           = current_stream();
    SegmentList seg;
    seg.ray_seg_offset    = torch::empty({R + 1}, L);
    seg.ray_sample_offset = torch::empty({R + 1}, L);
    auto maxr             = torch::empty({1}, I);
    const size_t wsb      = neat_raysample_workspace_bytes(R);
    const auto U8         = torch::TensorOptions().dtype(torch::kUInt8).device(dev);
    auto ws               = torch::empty({(int64_t)std::max<size_t>(wsb, 8)}, U8);
    // one walk per ray: the segments wait in a per-ray scratch until the totals size the list
    const size_t scb      = neat_raysample_scratch_bytes(R, max_depth_);
    auto scratch          = torch::empty({(int64_t)std::max<size_t>(scb, 16)}, U8);
    NEAT_CALL(neat_raysample_walk(s, dptr<float>(node_position_min), dptr<float>(node_position_max),
                                  node_flags.data_ptr<uint8_t>(), NumNodes(), max_depth_, dptr<float>(o),
                                  dptr<float>(d), R, max_spn, dptr<int64_t>(seg.ray_seg_offset),
                                  dptr<int64_t>(seg.ray_sample_offset), dptr<int32_t>(maxr), scratch.data_ptr(), scb,
                                  ws.data_ptr(), wsb));
    // one device -> host read of the totals per batch (the reference syncs on its counters too, HyperTree.cu:420)
    auto totals = torch::stack({seg.ray_seg_offset[R], seg.ray_sample_offset[R], maxr[0].to(torch::kLong)}).cpu();
    const int64_t K = totals[0].item<int64_t>();
    const int64_t S = totals[1].item<int64_t>();
    seg.node_id       = torch::empty({K}, I);
    seg.ray           = torch::empty({K}, I);
    seg.count         = torch::empty({K}, I);
    seg.tmin          = torch::empty({K}, F);
    seg.tmax          = torch::empty({K}, F);
    seg.sample_offset = torch::empty({K}, L);
    NEAT_CALL(neat_raysample_compact(s, dptr<float>(node_position_min), dptr<float>(node_position_max),
                                     node_flags.data_ptr<uint8_t>(), NumNodes(), max_depth_, dptr<float>(o),
                                     dptr<float>(d), R, max_spn, dptr<int64_t>(seg.ray_seg_offset),
                                     dptr<int64_t>(seg.ray_sample_offset), scratch.data_ptr(), scb, K,
                                     dptr<int32_t>(seg.node_id), dptr<int32_t>(seg.ray), dptr<int32_t>(seg.count),
                                     dptr<float>(seg.tmin), dptr<float>(seg.tmax), dptr<int64_t>(seg.sample_offset)));
    seg.num_samples   = S;
    seg.jitter        = jitter;
    seg.seed          = jitter ? (uint64_t)torch::randint(0, (int64_t)1 << 62, {1}, torch::kLong).item<int64_t>() : 0;
    seg.ray_origin       = o;
    seg.ray_direction    = d;
    seg.ray_origin_in    = rays.origin;
    seg.ray_direction_in = rays.direction;

    SampleList list;
    list.max_samples_per_ray = (int)totals[2].item<int64_t>();
    list.segments            = seg;
    if (!emit) return list;
    list.ray_t              = torch::empty({S}, F);
    list.weight             = torch::empty({S}, F);
    list.global_coordinate  = torch::empty({S, 3}, F);
    list.ray_index          = torch::empty({S}, L);
    list.local_index_in_ray = torch::empty({S}, L);
    list.node_id            = torch::empty({S}, L);
    NEAT_CALL(neat_raysample_emit(s, dptr<float>(o), dptr<float>(d), R, dptr<int64_t>(seg.ray_seg_offset),
                                  dptr<int64_t>(seg.ray_sample_offset), K, dptr<int32_t>(seg.node_id),
                                  dptr<int32_t>(seg.ray), dptr<int32_t>(seg.count), dptr<float>(seg.tmin),
                                  dptr<float>(seg.tmax), dptr<int64_t>(seg.sample_offset), S, jitter ? 1 : 0, seg.seed,
                                  dptr<float>(list.ray_t), dptr<float>(list.weight),
                                  dptr<float>(list.global_coordinate), dptr<int64_t>(list.ray_index),
                                  dptr<int64_t>(list.local_index_in_ray), dptr<int64_t>(list.node_id)));
    if (rays.origin.requires_grad() || rays.direction.requires_grad())
    {
        // positions through torch ops so the positional gradient reaches the rays (HyperTree.cu:426-432)
        auto origin2           = torch::index_select(rays.origin, 0, list.ray_index);
        auto dir2              = torch::index_select(rays.direction, 0, list.ray_index);
        list.global_coordinate = origin2 + dir2 * list.ray_t.unsqueeze(1);
    }
    return list;
}

SampleList HyperTreeBaseImpl::NodeNeighborSamples(const std::vector<int>& size, double epsilon)
{
    check_cuda(node_position_min, "the octree");
    TORCH_CHECK(size.size() == 3 && size[0] == size[1] && size[1] == size[2] && size[0] >= 2,
                "NodeNeighborSamples: a cubic face resolution >= 2 is expected");
    const auto dev     = node_position_min.device();
    const auto L       = torch::TensorOptions().dtype(torch::kLong).device(dev);
    const auto F       = torch::TensorOptions().dtype(torch::kFloat32).device(dev);
    const int G        = size[0];
    const int64_t A    = NumActiveNodes();
    const int64_t maxp = A * 6 * (int64_t)G * G;  // the reference's prediction, in pairs
    SampleList list;
    list.global_coordinate  = torch::empty({2 * maxp, 3}, F);
    list.node_id            = torch::empty({2 * maxp}, L);
    list.ray_index          = torch::empty({2 * maxp}, L);
    list.local_index_in_ray = torch::empty({2 * maxp}, L);
    list.weight             = torch::empty({2 * maxp}, F);
    list.ray_t              = torch::zeros({2 * maxp}, F);
    auto npairs             = torch::zeros({1}, L);
    auto ids                = active_node_ids.to(torch::kLong).contiguous();
    const size_t wsb        = neat_neighbor_pairs_workspace_bytes(A, G);
    auto ws = torch::empty({(int64_t)std::max<size_t>(wsb, 16)}, torch::TensorOptions().dtype(torch::kUInt8).device(dev));
    NEAT_CALL(neat_neighbor_pairs(current_stream(), dptr<float>(node_position_min), dptr<float>(node_position_max),
                                  node_flags.data_ptr<uint8_t>(), NumNodes(), dptr<int64_t>(ids), A, G,
                                  (float)epsilon, maxp, dptr<float>(list.global_coordinate),
                                  dptr<int64_t>(list.node_id), dptr<int64_t>(list.ray_index),
                                  dptr<int64_t>(list.local_index_in_ray), dptr<float>(list.weight),
                                  dptr<int64_t>(npairs), ws.data_ptr(), wsb));
    // shrink to the pairs found (the reference reads its counter back too, HyperTree.cu:928-931)
    const int64_t n = 2 * npairs.cpu().item<int64_t>();
    list.global_coordinate  = list.global_coordinate.slice(0, 0, n);
    list.node_id            = list.node_id.slice(0, 0, n);
    list.ray_index          = list.ray_index.slice(0, 0, n);
    list.local_index_in_ray = list.local_index_in_ray.slice(0, 0, n);
    list.weight             = list.weight.slice(0, 0, n);
    list.ray_t              = list.ray_t.slice(0, 0, n);
    return list;
}

SampleList HyperTreeBaseImpl::UniformPhantomSamplesGPU(const std::vector<int>& size, bool swap_xy)
{
    check_cuda(node_position_min, "the octree");
    TORCH_CHECK(size.size() == 3, "UniformPhantomSamplesGPU: a 3-d size is expected");
    TORCH_CHECK(NumActiveNodes() > 0, "UniformPhantomSamplesGPU: the tree has no active node");
    const auto dev  = node_position_min.device();
    const int64_t n = (int64_t)size[0] * size[1] * size[2];
    SampleList list;
    list.global_coordinate  = torch::empty({n, 3}, torch::TensorOptions().dtype(torch::kFloat32).device(dev));
    list.ray_index          = torch::empty({n}, torch::TensorOptions().dtype(torch::kLong).device(dev));
    list.node_id            = torch::empty({n}, torch::TensorOptions().dtype(torch::kLong).device(dev));
    list.weight             = torch::empty({n}, torch::TensorOptions().dtype(torch::kFloat32).device(dev));
    list.local_index_in_ray = torch::zeros({n}, torch::TensorOptions().dtype(torch::kLong).device(dev));
    list.ray_t              = torch::zeros({n}, torch::TensorOptions().dtype(torch::kFloat32).device(dev));
    const int32_t sz[3]     = {size[0], size[1], size[2]};
    const int64_t first     = active_node_ids.to(torch::kLong).cpu()[0].item<int64_t>();
    NEAT_CALL(neat_uniform_volume_samples(current_stream(), dptr<float>(node_position_min),
                                          dptr<float>(node_position_max), node_flags.data_ptr<uint8_t>(), NumNodes(),
                                          first, sz, swap_xy ? 1 : 0, dptr<float>(list.global_coordinate),
                                          dptr<int64_t>(list.ray_index), dptr<int64_t>(list.node_id),
                                          dptr<float>(list.weight)));
    return list;
}

void HyperTreeBaseImpl::SetErrorForActiveNodes(torch::Tensor error, std::string strategy)
{
    // elementwise over [num_nodes], the reference's tensor expressions (control-plane bookkeeping, not a hot path)
    TORCH_CHECK(error.sizes() == node_error.sizes(), "SetErrorForActiveNodes: error must be [num_nodes]");
    TORCH_CHECK(error.device() == node_error.device(), "SetErrorForActiveNodes: device mismatch");
    auto active_float   = node_active.to(torch::kFloat32);
    auto inactive       = 1 - active_float;
    auto invalid_errors = (node_error < 0).to(torch::kFloat32);
    error               = active_float * error;
    if (strategy == "override")
    {
        node_error.set_data(node_error * inactive + error);
    }
    else if (strategy == "min")
    {
        error               = error + inactive * 1000000;
        auto filtered_error = node_error + (invalid_errors * active_float) * 1000000;
        node_error.set_data(torch::min(filtered_error, error));
    }
    else
    {
        TORCH_CHECK(false, "SetErrorForActiveNodes: unknown strategy '", strategy, "' (override, min)");
    }
}
