#include "ImplicitNet.h"

#include <torch/csrc/autograd/custom_function.h>

using namespace neat_host;
using torch::autograd::AutogradContext;
using torch::autograd::tensor_list;

namespace
{

struct DecoderFn : public torch::autograd::Function<DecoderFn>
{
    static torch::Tensor forward(AutogradContext* ctx, torch::Tensor x, torch::Tensor params, int64_t in,
                                 int64_t hidden, int64_t layers, int64_t out, int64_t act, int64_t last_act,
                                 double beta)
    {
        check_cuda(x, "decoder input");
        auto xx = x.to(torch::kFloat32).contiguous();
        auto pp = params.to(torch::kFloat32).contiguous();
        auto y  = torch::empty({xx.size(0), out}, xx.options());
        NEAT_CALL(neat_decoder_fwd(current_stream(), dptr<float>(xx), xx.size(0), (int32_t)in, (int32_t)hidden,
                                   (int32_t)layers, (int32_t)out, dptr<float>(pp), (int32_t)act, (int32_t)last_act,
                                   (float)beta, dptr<float>(y)));
        ctx->save_for_backward({xx, pp});
        ctx->saved_data["dims"] = std::vector<int64_t>{in, hidden, layers, out, act, last_act};
        ctx->saved_data["beta"] = beta;
        return y;
    }

    static tensor_list backward(AutogradContext* ctx, tensor_list grads)
    {
        auto saved = ctx->get_saved_variables();
        auto xx = saved[0], pp = saved[1];
        auto dims       = ctx->saved_data["dims"].toIntVector();
        double beta     = ctx->saved_data["beta"].toDouble();
        auto gy         = grads[0].contiguous();
        auto gp         = torch::empty_like(pp);
        auto gx         = torch::empty_like(xx);
        const int64_t S = xx.size(0);
        size_t wsb      = neat_decoder_workspace_bytes(S, dims[0], dims[1], dims[2], dims[3]);
        auto ws = torch::empty({(int64_t)std::max<size_t>(wsb, 8)}, xx.options().dtype(torch::kUInt8));
        NEAT_CALL(neat_decoder_bwd(current_stream(), dptr<float>(xx), S, (int32_t)dims[0], (int32_t)dims[1],
                                   (int32_t)dims[2], (int32_t)dims[3], dptr<float>(pp), (int32_t)dims[4],
                                   (int32_t)dims[5], (float)beta, dptr<float>(gy), dptr<float>(gp), dptr<float>(gx),
                                   ws.data_ptr(), wsb));
        return {gx, gp, torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor(),
                torch::Tensor(), torch::Tensor()};
    }
};

torch::nn::AnyModule activation_module(const std::string& name)
{
    if (name == "relu") return torch::nn::AnyModule(torch::nn::ReLU());
    if (name == "silu") return torch::nn::AnyModule(torch::nn::SiLU());
    if (name == "id" || name == "identity") return torch::nn::AnyModule(torch::nn::Identity());
    TORCH_CHECK(false, "FCBlock: unsupported activation '", name, "' (relu, silu, id)");
    return {};
}

}  // namespace

torch::Tensor DecoderApply(torch::Tensor x, torch::Tensor packed_params, int in, int hidden, int hidden_layers,
                           int out, int act, int last_act, float softplus_beta)
{
    auto shape = x.sizes().vec();
    auto flat  = x.reshape({-1, shape.back()});
    auto y     = DecoderFn::apply(flat, packed_params, in, hidden, hidden_layers, out, act, last_act,
                                  (double)softplus_beta);
    shape.back() = out;
    return y.reshape(shape);
}

FCBlockImpl::FCBlockImpl(int in_features, int out_features, int num_hidden_layers, float hidden_features,
                         bool outermost_linear, std::string non_linearity)
    : in_features(in_features),
      out_features(out_features),
      num_hidden_layers(num_hidden_layers),
      hidden_features((int)hidden_features),
      outermost_linear(outermost_linear),
      non_linearity(non_linearity)
{
    auto make_lin = [](int in, int out)
    {
        auto lin = torch::nn::Linear(in, out);
        torch::nn::init::kaiming_normal_(lin->weight, 0, torch::kFanIn, torch::kReLU);  // ImplicitNet.h:28
        return lin;
    };
    seq->push_back(make_lin(in_features, this->hidden_features));
    seq->push_back(activation_module(non_linearity));
    for (int i = 0; i < num_hidden_layers; ++i)
    {
        seq->push_back(make_lin(this->hidden_features, this->hidden_features));
        seq->push_back(activation_module(non_linearity));
    }
    seq->push_back(make_lin(this->hidden_features, out_features));
    if (!outermost_linear) seq->push_back(activation_module(non_linearity));
    register_module("seq", seq);
}

torch::Tensor FCBlockImpl::packed_parameters()
{
    std::vector<torch::Tensor> parts;
    for (auto& m : seq->children())
    {
        if (auto* lin = m->as<torch::nn::Linear>())
        {
            parts.push_back(lin->weight.reshape({-1}));
            parts.push_back(lin->bias.reshape({-1}));
        }
    }
    return torch::cat(parts);
}

void FCBlockImpl::load_packed(torch::Tensor packed)
{
    torch::NoGradGuard ng;
    int64_t off = 0;
    for (auto& m : seq->children())
    {
        if (auto* lin = m->as<torch::nn::Linear>())
        {
            auto nw = lin->weight.numel(), nb = lin->bias.numel();
            lin->weight.copy_(packed.slice(0, off, off + nw).view_as(lin->weight));
            off += nw;
            lin->bias.copy_(packed.slice(0, off, off + nb));
            off += nb;
        }
    }
    TORCH_CHECK(off == packed.numel(), "packed decoder parameters have the wrong size");
}

at::Tensor FCBlockImpl::forward(at::Tensor x)
{
    TORCH_CHECK(in_features == x.size(-1));
    const int last = outermost_linear ? NEAT_ACT_ID : act_id(non_linearity);
    return DecoderApply(x, packed_parameters(), in_features, hidden_features, num_hidden_layers, out_features,
                        act_id(non_linearity), last, 1.f);
}

at::Tensor FCBlockImpl::forward_with_last(at::Tensor x, const std::string& last_activation, float softplus_beta)
{
    TORCH_CHECK(in_features == x.size(-1));
    TORCH_CHECK(outermost_linear, "forward_with_last needs an outermost-linear decoder");
    return DecoderApply(x, packed_parameters(), in_features, hidden_features, num_hidden_layers, out_features,
                        act_id(non_linearity), act_id(last_activation), softplus_beta);
}

ExplicitFeatureGridImpl::ExplicitFeatureGridImpl(int num_nodes, int out_features, std::vector<long> output_grid,
                                                 std::string init)
{
    std::vector<int64_t> sizes = {num_nodes, out_features};
    for (auto g : output_grid) sizes.push_back(g);
    TORCH_CHECK(sizes.size() == 5, "ExplicitFeatureGrid: only 3-D grids");
    if (init == "uniform" || init == "random")
        grid_data = torch::empty(sizes).uniform_(-1, 1);
    else if (init == "minus")
        grid_data = -torch::ones(sizes);
    else if (init == "zero")
        grid_data = torch::zeros(sizes);
    else if (init == "normal")
        grid_data = torch::randn(sizes) * 0.1;
    else
        TORCH_CHECK(false, "Unknown grid init: ", init, ". Expected: uniform, minus, zero, normal");
    grid_data = grid_data.contiguous(at::MemoryFormat::ChannelsLast3d);
    register_parameter("grid_data", grid_data);
}
