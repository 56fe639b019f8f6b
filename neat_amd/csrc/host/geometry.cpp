#include "geometry.h"

#include <torch/csrc/autograd/custom_function.h>

using namespace neat_host;
using torch::autograd::AutogradContext;
using torch::autograd::tensor_list;

namespace
{

struct IntegrateFn : public torch::autograd::Function<IntegrateFn>
{
    static torch::Tensor forward(AutogradContext* ctx, torch::Tensor values, torch::Tensor weight,
                                 torch::Tensor ray_index, int64_t num_rays, torch::Tensor offsets)
    {
        auto v  = values.to(torch::kFloat32).contiguous();
        auto w  = weight.to(torch::kFloat32).reshape({-1}).contiguous();
        auto ri = ray_index.to(torch::kLong).reshape({-1}).contiguous();
        const int64_t S = ri.numel();
        const int C     = (int)v.size(1);
        auto out        = torch::empty({C, num_rays}, v.options());
        NEAT_CALL(neat_integrate_fwd(current_stream(), dptr<float>(v), dptr<float>(w), dptr<int64_t>(ri), S, C,
                                     num_rays, offsets.defined() ? dptr<int64_t>(offsets) : nullptr,
                                     dptr<float>(out)));
        ctx->save_for_backward({v, w, ri});
        ctx->saved_data["R"] = num_rays;
        return out;
    }

    static tensor_list backward(AutogradContext* ctx, tensor_list grads)
    {
        auto saved = ctx->get_saved_variables();
        auto v = saved[0], w = saved[1], ri = saved[2];
        const int64_t R = ctx->saved_data["R"].toInt();
        auto g          = grads[0].contiguous();
        auto gv         = torch::empty_like(v);
        NEAT_CALL(neat_integrate_bwd(current_stream(), dptr<float>(g), dptr<float>(w), dptr<int64_t>(ri), ri.numel(),
                                     (int)v.size(1), R, dptr<float>(gv)));
        torch::Tensor gw;
        if (ctx->needs_input_grad(1))
        {
            // d/dw: sum_c grad[c, ray] * v[i, c]
            gw = (torch::index_select(g, 1, ri).t() * v).sum(1, true);
        }
        return {gv, gw, torch::Tensor(), torch::Tensor(), torch::Tensor()};
    }
};

// The fused projector over a canonical segment list: differentiable in the grid and the packed decoder parameters.
struct ProjectorFn : public torch::autograd::Function<ProjectorFn>
{
    static neat_projector_args args(const std::vector<torch::Tensor>& t, const std::vector<int64_t>& iv, double beta)
    {
        // t: grid, params, pmin, pmax, prefix, origin, direction, ray_seg_offset, seg_node, seg_ray, seg_count,
        //    seg_tmin, seg_tmax
        // iv: num_samples, jitter, seed, in, hidden, layers, out, act, last_act, precision
        // t[13]: BF16 copy of the grid (NEAT_PREC_BF16 only)
        neat_projector_args a{};
        a.node_pos_min           = dptr<float>(t[2]);
        a.node_pos_max           = dptr<float>(t[3]);
        a.node_active_prefix_sum = dptr<int64_t>(t[4]);
        a.num_nodes              = (int32_t)t[2].size(0);
        a.ray_origin             = dptr<float>(t[5]);
        a.ray_direction          = dptr<float>(t[6]);
        a.num_rays               = t[5].size(0);
        a.ray_seg_offset         = dptr<int64_t>(t[7]);
        a.num_segments           = t[8].size(0);
        a.seg_node               = dptr<int32_t>(t[8]);
        a.seg_ray                = dptr<int32_t>(t[9]);
        a.seg_count              = dptr<int32_t>(t[10]);
        a.seg_tmin               = dptr<float>(t[11]);
        a.seg_tmax               = dptr<float>(t[12]);
        a.num_samples            = iv[0];
        a.jitter                 = (int32_t)iv[1];
        a.seed                   = (uint64_t)iv[2];
        a.grid                   = dptr<float>(t[0]);
        a.num_slots              = t[0].size(0);
        a.channels               = (int32_t)t[0].size(1);
        a.grid_size              = (int32_t)t[0].size(2);
        for (int i = 0; i < 5; ++i) a.grid_strides[i] = t[0].stride(i);
        a.in_features   = (int32_t)iv[3];
        a.hidden        = (int32_t)iv[4];
        a.hidden_layers = (int32_t)iv[5];
        a.out_features  = (int32_t)iv[6];
        a.params        = dptr<float>(t[1]);
        a.act           = (int32_t)iv[7];
        a.last_act      = (int32_t)iv[8];
        a.softplus_beta = (float)beta;
        a.precision     = (int32_t)iv[9];
        a.grid_bf16     = a.precision == NEAT_PREC_BF16 ? reinterpret_cast<const uint16_t*>(dptr<int16_t>(t[13])) : nullptr;
        return a;
    }

    // origin_in / direction_in: the caller's rays (gradients flow to them when they require grad); geo holds their
    // FP32 contiguous copies the kernels read
    static torch::Tensor forward(AutogradContext* ctx, torch::Tensor grid, torch::Tensor params, torch::Tensor origin_in,
                                 torch::Tensor direction_in, std::vector<torch::Tensor> geo, std::vector<int64_t> iv,
                                 double beta, bool compensated)
    {
        ctx->saved_data["ray_grad"] = origin_in.requires_grad() || direction_in.requires_grad();
        ctx->saved_data["o_type"]   = (int64_t)origin_in.scalar_type();
        ctx->saved_data["d_type"]   = (int64_t)direction_in.scalar_type();
        std::vector<torch::Tensor> t = {grid, params.contiguous()};
        for (auto& g : geo) t.push_back(g);
        auto s = current_stream();
        if (iv[9] == NEAT_PREC_BF16)
        {
            // the grid changes every optimizer step: round it once per forward
            auto gb = torch::empty_like(grid, grid.options().dtype(torch::kInt16));
            NEAT_CALL(neat_grid_to_bf16(s, dptr<float>(grid), grid.numel(), reinterpret_cast<uint16_t*>(dptr<int16_t>(gb))));
            t.push_back(gb);
        }
        else
        {
            t.push_back(torch::empty({0}, grid.options().dtype(torch::kInt16)));
        }
        neat_projector_args a = args(t, iv, beta);
        const size_t bytes    = neat_projector_workspace_bytes(&a);
        TORCH_CHECK(bytes > 0, "fused projector: ", neat_last_error());
        auto ws   = torch::empty({(int64_t)bytes}, grid.options().dtype(torch::kUInt8));
        auto proj = torch::empty({1, a.num_rays}, grid.options().memory_format(at::MemoryFormat::Contiguous));
        NEAT_CALL(neat_projector_plan(s, &a, ws.data_ptr(), bytes));
        NEAT_CALL(neat_projector_fwd(s, &a, compensated ? 1 : 0, dptr<float>(proj), ws.data_ptr(), bytes));
        t.push_back(ws);
        ctx->save_for_backward(t);
        ctx->saved_data["iv"]   = iv;
        ctx->saved_data["beta"] = beta;
        return proj;
    }

    static tensor_list backward(AutogradContext* ctx, tensor_list grads)
    {
        auto t            = ctx->get_saved_variables();
        auto ws           = t.back();
        auto iv           = ctx->saved_data["iv"].toIntVector();
        double beta       = ctx->saved_data["beta"].toDouble();
        neat_projector_args a = args(t, iv, beta);
        auto gproj        = grads[0].reshape({-1}).contiguous();
        auto gg           = torch::empty_like(t[0]);
        auto gp           = torch::empty_like(t[1]);
        NEAT_CALL(neat_projector_bwd(current_stream(), &a, dptr<float>(gproj), dptr<float>(gg), dptr<float>(gp),
                                     ws.data_ptr(), (size_t)ws.numel()));
        // one slot per forward argument (each tensor of `geo` counts); trailing slots stay undefined
        tensor_list out(4 + 11 + 3);
        out[0] = gg;
        out[1] = gp;
        if (ctx->saved_data["ray_grad"].toBool())
        {
            // the positional gradient onto the rays, from the same backward's dL/dfeature (HyperTree.cu:426-432)
            auto go = torch::empty({a.num_rays, 3}, t[0].options().memory_format(at::MemoryFormat::Contiguous));
            auto gd = torch::empty_like(go);
            NEAT_CALL(neat_projector_ray_grad(current_stream(), &a, dptr<float>(go), dptr<float>(gd), ws.data_ptr(),
                                              (size_t)ws.numel()));
            out[2] = go.to((c10::ScalarType)ctx->saved_data["o_type"].toInt());
            out[3] = gd.to((c10::ScalarType)ctx->saved_data["d_type"].toInt());
        }
        return out;
    }
};

}  // namespace

torch::Tensor IntegrateSamples(torch::Tensor sample_values, torch::Tensor integration_weight, torch::Tensor ray_index,
                               int num_channels, int num_rays, torch::Tensor ray_sample_offset)
{
    if (sample_values.numel() == 0)
        return torch::zeros({num_channels, num_rays}, torch::TensorOptions(ray_index.device()));
    TORCH_CHECK(sample_values.dim() == integration_weight.dim());
    auto v = sample_values.reshape({-1, num_channels});
    return IntegrateFn::apply(v, integration_weight.reshape({-1}), ray_index.reshape({-1}), num_rays,
                              ray_sample_offset);
}

torch::Tensor NeuralGeometry::IntegrateSamplesXRay(torch::Tensor sample_values, torch::Tensor integration_weight,
                                                   torch::Tensor ray_index, int num_channels, int num_rays)
{
    // arbitrary sample order: float atomics like the reference's index_add_
    return IntegrateSamples(sample_values, integration_weight, ray_index, num_channels, num_rays, {});
}

HierarchicalNeuralGeometry::HierarchicalNeuralGeometry(int num_channels, int D, std::shared_ptr<CombinedParams> params,
                                                       HyperTreeBase tree)
    : NeuralGeometry(num_channels, D, params), tree(tree)
{
    TORCH_CHECK(!params->net_params.decoder_skip, "decoder_skip is not supported by this build");
    decoder = FCBlock(params->net_params.grid_features, num_channels, params->net_params.decoder_hidden_layers,
                      params->net_params.decoder_hidden_features, true, params->net_params.decoder_activation);
    register_module("decoder", decoder);
}

torch::Tensor HierarchicalNeuralGeometry::DecodeFeatures(torch::Tensor neural_features)
{
    return decoder->forward_with_last(neural_features, params->net_params.last_activation_function,
                                      params->net_params.softplus_beta);
}

torch::Tensor HierarchicalNeuralGeometry::ComputeImage(SampleList all_samples, int num_channels, int num_pixels)
{
    // rays that require grad (e.g. GetRays with pose refinement) keep the fused projector: its backward also returns
    // the positional gradient onto the rays (neat_projector_ray_grad)
    if (params->train_params.use_fused_projector && all_samples.segments.defined() && num_channels == 1)
    {
        auto img = ComputeImageFused(all_samples, num_pixels);
        if (img.defined()) return img;
    }
    TORCH_CHECK(all_samples.global_coordinate.defined(),
                "ComputeImage: the SampleList holds only segments and the fused projector does not cover this "
                "configuration; create it with CreateSamplesForRays");
    auto sample_output = SampleVolumeIndirect(all_samples.global_coordinate, all_samples.node_id);
    TORCH_CHECK(sample_output.defined(), "SampleVolumeIndirect is not implemented by this geometry");
    auto weight = all_samples.weight.unsqueeze(1);
    // the sampler's output is ray-sorted: segmented, deterministic sums
    torch::Tensor offsets =
        all_samples.segments.defined() ? all_samples.segments.ray_sample_offset : torch::Tensor();
    auto img = IntegrateSamples(sample_output, weight, all_samples.ray_index, num_channels, num_pixels, offsets);
    TORCH_CHECK(img.size(0) == num_channels);
    return img;
}

// TVLoss::forward(grid[slots]) * weight (geometry_ex_ex.h:18-61) on neat_tv_loss_fwd / _bwd.
struct TVLossFn : public torch::autograd::Function<TVLossFn>
{
    static torch::Tensor forward(torch::autograd::AutogradContext* ctx, torch::Tensor grid, torch::Tensor slots,
                                 double weight)
    {
        TORCH_CHECK(grid.dim() == 5 && grid.size(2) == grid.size(3) && grid.size(3) == grid.size(4),
                    "TV loss: a cubic grid [N, C, G, G, G] is expected");
        const int64_t st[5] = {grid.stride(0), grid.stride(1), grid.stride(2), grid.stride(3), grid.stride(4)};
        auto loss = torch::empty({1}, grid.options().dtype(torch::kFloat64));
        const int64_t A = slots.size(0), C = grid.size(1), G = grid.size(2);
        TORCH_CHECK(A <= grid.size(0), "TV loss: ", A, " active nodes for ", grid.size(0), " grid rows");
        const size_t wsb = neat_tv_workspace_bytes(A, (int32_t)C, (int32_t)G);
        auto ws = torch::empty({(int64_t)std::max<size_t>(wsb, 16)}, grid.options().dtype(torch::kUInt8));
        NEAT_CALL(neat_tv_loss_fwd(current_stream(), dptr<float>(grid), grid.size(0), (int32_t)C, (int32_t)G, st,
                                   dptr<int64_t>(slots), A, weight, dptr<double>(loss), ws.data_ptr(), wsb));
        ctx->save_for_backward({grid, slots});
        ctx->saved_data["weight"] = weight;
        return loss.to(torch::kFloat32).reshape({});
    }

    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                   torch::autograd::variable_list grad_out)
    {
        auto saved  = ctx->get_saved_variables();
        auto grid   = saved[0], slots = saved[1];
        auto gl     = grad_out[0].to(torch::kFloat32).reshape({1}).contiguous();
        auto gg     = torch::zeros_like(grid);
        const int64_t st[5] = {grid.stride(0), grid.stride(1), grid.stride(2), grid.stride(3), grid.stride(4)};
        TORCH_CHECK(gg.strides() == grid.strides(), "TV loss: gradient layout");
        NEAT_CALL(neat_tv_loss_bwd(current_stream(), dptr<float>(grid), grid.size(0), (int32_t)grid.size(1),
                                   (int32_t)grid.size(2), st, dptr<int64_t>(slots), slots.size(0),
                                   ctx->saved_data["weight"].toDouble(), dptr<float>(gl), dptr<float>(gg)));
        return {gg, torch::Tensor(), torch::Tensor()};
    }
};

// mean_k(mean_c |f[2k] - f[2k+1]| * r_k) * weight * C (geometry_ex_ex.cpp:121-134) on neat_edge_loss_fwd / _bwd.
struct EdgeLossFn : public torch::autograd::Function<EdgeLossFn>
{
    static torch::Tensor forward(torch::autograd::AutogradContext* ctx, torch::Tensor features, int64_t seed,
                                 double weight)
    {
        auto f          = features.contiguous();
        const int64_t P = f.size(0) / 2;
        auto loss       = torch::empty({1}, f.options().dtype(torch::kFloat64));
        const size_t wsb = neat_reg_workspace_bytes(P);
        auto ws = torch::empty({(int64_t)std::max<size_t>(wsb, 16)}, f.options().dtype(torch::kUInt8));
        NEAT_CALL(neat_edge_loss_fwd(current_stream(), dptr<float>(f), P, (int32_t)f.size(1), (uint64_t)seed, weight,
                                     dptr<double>(loss), ws.data_ptr(), wsb));
        ctx->save_for_backward({f});
        ctx->saved_data["seed"]   = seed;
        ctx->saved_data["weight"] = weight;
        return loss.to(torch::kFloat32).reshape({});
    }

    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                   torch::autograd::variable_list grad_out)
    {
        auto f  = ctx->get_saved_variables()[0];
        auto gl = grad_out[0].to(torch::kFloat32).reshape({1}).contiguous();
        auto gf = torch::empty_like(f);
        NEAT_CALL(neat_edge_loss_bwd(current_stream(), dptr<float>(f), f.size(0) / 2, (int32_t)f.size(1),
                                     (uint64_t)ctx->saved_data["seed"].toInt(), ctx->saved_data["weight"].toDouble(),
                                     dptr<float>(gl), dptr<float>(gf)));
        return {gf, torch::Tensor(), torch::Tensor()};
    }
};

// The same edge loss in one fused pass over the candidate face points (neat_edge_loss_direct): the neighbour rows,
// their local coordinates and features are never materialised; backward recomputes the pairs (they depend on the
// tree only) and adds the grid gradient with atomics. Values and gradients equal NodeNeighborSamples ->
// IndirectGridSample3D -> EdgeLossFn up to the order of the FP64 loss sum and of the float atomics.
struct FaceEdgeLossFn : public torch::autograd::Function<FaceEdgeLossFn>
{
    static void call(const torch::Tensor& grid, const std::vector<torch::Tensor>& t, int64_t seed, double weight,
                     double* loss, const float* grad_loss, float* grad_grid)
    {
        const auto &pmin = t[0], &pmax = t[1], &flags = t[2], &ids = t[3], &slot = t[4];
        const int64_t st[5] = {grid.stride(0), grid.stride(1), grid.stride(2), grid.stride(3), grid.stride(4)};
        const int64_t A = ids.size(0);
        const int32_t G = (int32_t)grid.size(2);
        const size_t wsb = neat_edge_loss_direct_workspace_bytes(A, G);
        auto ws = torch::empty({(int64_t)std::max<size_t>(wsb, 16)}, grid.options().dtype(torch::kUInt8));
        NEAT_CALL(neat_edge_loss_direct(current_stream(), dptr<float>(pmin), dptr<float>(pmax),
                                        flags.data_ptr<uint8_t>(), (int32_t)pmin.size(0), dptr<int64_t>(ids), A,
                                        dptr<int64_t>(slot), dptr<float>(grid), grid.size(0), (int32_t)grid.size(1), G,
                                        st, 0.001f, (uint64_t)seed, weight, loss, grad_loss, grad_grid, ws.data_ptr(),
                                        wsb));
    }

    static torch::Tensor forward(torch::autograd::AutogradContext* ctx, torch::Tensor grid, torch::Tensor pos_min,
                                 torch::Tensor pos_max, torch::Tensor flags, torch::Tensor active_ids,
                                 torch::Tensor node_slot, int64_t seed, double weight)
    {
        TORCH_CHECK(grid.dim() == 5 && grid.size(2) == grid.size(3) && grid.size(3) == grid.size(4),
                    "edge loss: a cubic grid [N, C, G, G, G] is expected");
        auto loss = torch::empty({1}, grid.options().dtype(torch::kFloat64));
        call(grid, {pos_min, pos_max, flags, active_ids, node_slot}, seed, weight, dptr<double>(loss), nullptr, nullptr);
        ctx->save_for_backward({grid, pos_min, pos_max, flags, active_ids, node_slot});
        ctx->saved_data["seed"]   = seed;
        ctx->saved_data["weight"] = weight;
        return loss.to(torch::kFloat32).reshape({});
    }

    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                   torch::autograd::variable_list grad_out)
    {
        auto sv   = ctx->get_saved_variables();
        auto grid = sv[0];
        auto gl   = grad_out[0].to(torch::kFloat32).reshape({1}).contiguous();
        auto gg   = torch::zeros_like(grid);
        TORCH_CHECK(gg.strides() == grid.strides(), "edge loss: gradient layout");
        call(grid, {sv[1], sv[2], sv[3], sv[4], sv[5]}, ctx->saved_data["seed"].toInt(),
             ctx->saved_data["weight"].toDouble(), nullptr, dptr<float>(gl), dptr<float>(gg));
        return {gg, torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor(),
                torch::Tensor(), torch::Tensor()};
    }
};

torch::Tensor GeometryExEx::VolumeRegularizer(int64_t seed)
{
    auto& tp  = params->train_params;
    auto grid = explicit_grid_generator->grid_data;
    // the reference's index_select raises for a slot past the grid's rows; so does this
    TORCH_CHECK(tree->NumActiveNodes() <= grid.size(0), "VolumeRegularizer: ", tree->NumActiveNodes(),
                " active nodes for a grid of ", grid.size(0), " slots (max_active_nodes)");
    torch::Tensor loss;
    if (tp.loss_tv > 0)
    {
        auto slots = tree->GlobalNodeIdToLocalActiveId(tree->ActiveNodeTensor()).to(torch::kLong).contiguous();
        loss       = TVLossFn::apply(grid, slots, tp.loss_tv);
    }
    if (tp.loss_edge > 0)
    {
        if (seed < 0) seed = torch::randint(0, (int64_t)1 << 62, {1}, torch::kLong).item<int64_t>();
        torch::Tensor edge;
        if (fused_edge_loss)
        {
            edge = FaceEdgeLossFn::apply(grid, tree->node_position_min, tree->node_position_max, tree->node_flags,
                                         tree->active_node_ids.to(torch::kLong).contiguous(),
                                         tree->node_active_prefix_sum, seed, tp.loss_edge);
        }
        else
        {
            const int G   = params->net_params.grid_size;
            SampleList nb = tree->NodeNeighborSamples({G, G, G}, 0.001);
            if (nb.size() > 0)
            {
                auto local_samples   = tree->ComputeLocalSamples(nb.global_coordinate, nb.node_id);
                auto local_node_id   = tree->GlobalNodeIdToLocalActiveId(nb.node_id);
                auto neural_features = IndirectGridSample3D(grid, local_node_id, local_samples);
                edge                 = EdgeLossFn::apply(neural_features, seed, tp.loss_edge);
            }
        }
        if (edge.defined()) loss = loss.defined() ? loss + edge : edge;
    }
    return loss;
}

std::tuple<torch::Tensor, torch::Tensor, torch::Tensor> HierarchicalNeuralGeometry::UniformSampledVolume(
    std::vector<long> shape, int num_channels)
{
    TORCH_CHECK(shape.size() == 3, "UniformSampledVolume: a 3-d shape is expected");
    torch::NoGradGuard ng;
    auto all      = tree->UniformPhantomSamplesGPU({(int)shape[0], (int)shape[1], (int)shape[2]}, false);
    const int64_t n = all.global_coordinate.size(0);
    auto values   = torch::zeros({num_channels, n}, all.weight.options());
    const int64_t batch = (int64_t)1 << 20;
    for (int64_t b = 0; b < n; b += batch)
    {
        const int64_t e = std::min(n, b + batch);
        auto out = SampleVolumeIndirect(all.global_coordinate.slice(0, b, e), all.node_id.slice(0, b, e));
        TORCH_CHECK(out.defined() && out.size(1) == num_channels, "UniformSampledVolume: the geometry gives ",
                    out.defined() ? out.size(1) : 0, " channels, ", num_channels, " requested");
        // model output times the integration weight (0 outside the active nodes)
        values.slice(1, b, e).copy_((out * all.weight.slice(0, b, e).unsqueeze(1)).t());
    }
    // a voxel in no active node keeps the first active node as its id, as the reference's does (its mask covers only
    // the group padding, geometry.cpp:228-233); the validity flag tells the two apart
    auto valid   = all.weight.to(torch::kLong);
    auto node_id = all.node_id;
    std::vector<long> vshape = shape;
    vshape.insert(vshape.begin(), num_channels);
    return {values.reshape(vshape).cpu(), node_id.reshape(shape).cpu(), valid.reshape(shape).cpu()};
}

std::pair<torch::Tensor, torch::Tensor> HierarchicalNeuralGeometry::AccumulateSampleLossPerNode(
    const SampleList& combined_samples, torch::Tensor per_ray_loss)
{
    const SegmentList& seg = combined_samples.segments;
    TORCH_CHECK(seg.defined(), "AccumulateSampleLossPerNode: the SampleList has no segment list");
    check_cuda(per_ray_loss, "per_ray_loss");
    // [num_rays] (geometry.cpp:42)
    auto loss = (per_ray_loss.dim() == 2 ? torch::mean(per_ray_loss, 0) : per_ray_loss).detach().to(torch::kFloat32)
                    .contiguous();
    const int64_t R = seg.ray_seg_offset.size(0) - 1;
    TORCH_CHECK(loss.numel() == R, "AccumulateSampleLossPerNode: ", loss.numel(), " losses for ", R, " rays");
    const int32_t nn = (int32_t)tree->NumNodes();
    auto F           = loss.options();
    auto ls = torch::empty({nn}, F), ws_ = torch::empty({nn}, F);
    const size_t wsb = neat_node_loss_workspace_bytes(seg.size(), nn);
    auto ws = torch::empty({(int64_t)std::max<size_t>(wsb, 16)}, F.dtype(torch::kUInt8));
    NEAT_CALL(neat_node_loss_accumulate(current_stream(), dptr<int64_t>(seg.ray_seg_offset), R,
                                        dptr<int32_t>(seg.node_id), dptr<int32_t>(seg.ray), dptr<int32_t>(seg.count),
                                        dptr<float>(seg.tmin), dptr<float>(seg.tmax), seg.size(), seg.jitter ? 1 : 0,
                                        seg.seed, dptr<float>(loss), nn, dptr<float>(ls), dptr<float>(ws_),
                                        ws.data_ptr(), wsb));
    return {ls, ws_};
}

GeometryExEx::GeometryExEx(int num_channels, int D, HyperTreeBase tree, std::shared_ptr<CombinedParams> params)
    : HierarchicalNeuralGeometry(num_channels, D, params, tree)
{
    std::vector<long> shape(D, params->net_params.grid_size);
    explicit_grid_generator =
        ExplicitFeatureGrid(params->octree_params.tree_optimizer_params.max_active_nodes,
                            params->net_params.grid_features, shape, params->train_params.grid_init);
    register_module("explicit_grid_generator", explicit_grid_generator);
}

torch::Tensor GeometryExEx::SampleVolumeIndirect(torch::Tensor global_coordinate, torch::Tensor node_id)
{
    auto local_samples   = tree->ComputeLocalSamples(global_coordinate, node_id);
    auto local_node_id   = tree->GlobalNodeIdToLocalActiveId(node_id);
    auto neural_features = IndirectGridSample3D(explicit_grid_generator->grid_data, local_node_id, local_samples);
    return DecodeFeatures(neural_features);
}

std::string GeometryExEx::FusedUnsupportedReason()
{
    auto& np  = params->net_params;
    auto& g   = explicit_grid_generator->grid_data;
    const int C = (int)g.size(1), G = (int)g.size(2);
    if (!(g.size(3) == G && g.size(4) == G)) return "grid is not cubic";
    if (!g.is_contiguous(at::MemoryFormat::ChannelsLast3d)) return "grid is not channels-last-3d";
    if (!(C == 4 || C == 8)) return "grid_features not in {4, 8}";
    if (!(np.decoder_hidden_features == 32 || np.decoder_hidden_features == 64)) return "hidden width not in {32, 64}";
    if (np.decoder_hidden_layers != 1) return "decoder_hidden_layers != 1";
    if (!(np.decoder_activation == "relu" || np.decoder_activation == "silu")) return "hidden activation not relu/silu";
    if (np.last_activation_function != "softplus") return "last activation is not softplus";
    if (num_channels != 1) return "num_channels != 1";
    const int prec = precision_id(params->train_params.projector_precision);
    if (prec < 0) return "projector_precision not in {fp32, bf16, fp32_valu}";
    if (prec == NEAT_PREC_BF16 && C != 8) return "the bf16 projector needs grid_features == 8";
    return "";
}

torch::Tensor GeometryExEx::ComputeImageFused(const SampleList& samples, int num_pixels)
{
    if (!FusedUnsupportedReason().empty()) return {};
    const auto& s = samples.segments;
    TORCH_CHECK(s.ray_origin.size(0) == num_pixels, "ComputeImage: num_pixels differs from the ray count");
    auto& np = params->net_params;
    std::vector<torch::Tensor> geo = {tree->node_position_min, tree->node_position_max, tree->node_active_prefix_sum,
                                      s.ray_origin, s.ray_direction, s.ray_seg_offset, s.node_id, s.ray, s.count,
                                      s.tmin, s.tmax};
    std::vector<int64_t> iv = {s.num_samples, s.jitter ? 1 : 0, (int64_t)s.seed, np.grid_features,
                               np.decoder_hidden_features, np.decoder_hidden_layers, 1, act_id(np.decoder_activation),
                               NEAT_ACT_SOFTPLUS, precision_id(params->train_params.projector_precision)};
    // the caller's rays carry the positional gradient; without them (a SampleList assembled by hand) the FP32 copies
    auto o_in = s.ray_origin_in.defined() ? s.ray_origin_in : s.ray_origin;
    auto d_in = s.ray_direction_in.defined() ? s.ray_direction_in : s.ray_direction;
    return ProjectorFn::apply(explicit_grid_generator->grid_data, decoder->packed_parameters(), o_in, d_in, geo, iv,
                              (double)np.softplus_beta, params->train_params.compensated_sum);
}

void HierarchicalNeuralGeometry::CreateGeometryOptimizer()
{
    optimizer_adam    = std::make_shared<NeatAdam>();
    optimizer_decoder = std::make_shared<NeatAdam>();
    AddParametersToOptimizer();
}

void HierarchicalNeuralGeometry::AddParametersToOptimizer()
{
    if (decoder) optimizer_decoder->add_param_group(decoder->parameters(), params->net_params.decoder_lr);
}

void HierarchicalNeuralGeometry::OptimizerStep(int epoch_id)
{
    (void)epoch_id;
    for (auto& o : {optimizer_adam, optimizer_decoder})
    {
        if (!o) continue;
        o->step();
        o->zero_grad();
    }
}

void HierarchicalNeuralGeometry::UpdateLearningRate(double factor)
{
    if (optimizer_adam) optimizer_adam->update_lr(factor);
    if (optimizer_decoder) optimizer_decoder->update_lr(factor);
}

void GeometryExEx::AddParametersToOptimizer()
{
    HierarchicalNeuralGeometry::AddParametersToOptimizer();
    optimizer_adam->add_param_group(explicit_grid_generator->parameters(), params->train_params.lr_exex_grid_adam);
}
