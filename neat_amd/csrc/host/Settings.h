// The projector path's parameters, with the field names and nesting of the reference's CombinedParams
// (Settings.h:26-99, 156-284 there). Only the fields the path reads are kept; INI/CLI loading is out of scope.
#pragma once

#include <string>

struct Netparams
{
    int grid_size                        = 17;
    int grid_features                    = 4;
    std::string last_activation_function = "softplus";
    float softplus_beta                  = 2;
    bool decoder_skip                    = false;
    float decoder_lr                     = 0.00005f;
    std::string decoder_activation       = "silu";
    int decoder_hidden_layers            = 1;
    int decoder_hidden_features          = 64;
};

struct TreeOptimizerParams
{
    int max_active_nodes = 512;
};

struct OctreeParams
{
    int start_layer          = 3;
    int tree_depth           = 4;
    int max_samples_per_node = 32;
    TreeOptimizerParams tree_optimizer_params;
};

struct MyTrainParams
{
    std::string grid_init   = "uniform";
    double lr_exex_grid_adam = 0.04;
    // projector execution: the fused kernels when the configuration is covered, compensated per-ray sums
    bool use_fused_projector = true;
    bool compensated_sum     = false;
    // fused projector arithmetic: "fp32" (matrix cores), "bf16" (BF16 grid copy and hidden layers, FP32
    // accumulation) or "fp32_valu" (vector ALUs)
    std::string projector_precision = "fp32";
    // volume regulariser weights (Settings.h:223-224 in the reference)
    double loss_tv   = 1e-4;
    double loss_edge = 1e-3;
};

struct CombinedParams
{
    MyTrainParams train_params;
    OctreeParams octree_params;
    Netparams net_params;
};
