// FCBlock decoder and ExplicitFeatureGrid storage (ImplicitNet.h:18-119 in the reference). FCBlock keeps the
// reference's module tree (seq of Linear + activation) and parameter names; its forward packs the Linear parameters
// and runs the C ABI's fused decoder kernels through an autograd Function.
#pragma once

#include "neat_host.h"

// Packed decoder [S, in] -> [S, out] with the given hidden / last activations, differentiable in x and params.
torch::Tensor DecoderApply(torch::Tensor x, torch::Tensor packed_params, int in, int hidden, int hidden_layers,
                           int out, int act, int last_act, float softplus_beta);

class FCBlockImpl : public torch::nn::Module
{
   public:
    FCBlockImpl(int in_features, int out_features, int num_hidden_layers, float hidden_features,
                bool outermost_linear = true, std::string non_linearity = "relu");

    // Raw decoder output (the last Linear, or its activation when !outermost_linear).
    at::Tensor forward(at::Tensor x);
    // Decoder followed by a last activation ("softplus", "relu", "abs", "id"), fused in one kernel.
    at::Tensor forward_with_last(at::Tensor x, const std::string& last_activation, float softplus_beta);

    // All Linear parameters packed as W [out, in] then b [out] per layer (differentiable: a torch::cat).
    torch::Tensor packed_parameters();
    // Overwrites the Linear parameters from a packed vector.
    void load_packed(torch::Tensor packed);

    int in_features, out_features, num_hidden_layers, hidden_features;
    bool outermost_linear;
    std::string non_linearity;
    torch::nn::Sequential seq;
};
TORCH_MODULE(FCBlock);

// Stores a [num_nodes, out_features, G, G, G] feature grid (parameter "grid_data"), kept in channels-last-3d memory
// format so the fused projector reads each voxel's channels as one contiguous run.
class ExplicitFeatureGridImpl : public torch::nn::Module
{
   public:
    ExplicitFeatureGridImpl(int num_nodes, int out_features, std::vector<long> output_grid, std::string init = "zero");
    at::Tensor forward(at::Tensor node_index) { return torch::index_select(grid_data, 0, node_index); }
    torch::Tensor grid_data;
};
TORCH_MODULE(ExplicitFeatureGrid);
