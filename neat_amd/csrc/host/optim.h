// Adam over parameter groups on neat_adam_step: the optimizer HierarchicalNeuralGeometry::OptimizerStep steps in
// the reference (torch::optim::Adam, geometry.h:46-88; its groups: grid geometry_ex_ex.cpp:40-49, decoder
// geometry.cpp:87-97), with torch's defaults (betas 0.9 / 0.999, eps 1e-8, no weight decay).
#pragma once

#include <torch/torch.h>

#include <vector>

class NeatAdam
{
   public:
    NeatAdam(double beta1 = 0.9, double beta2 = 0.999, double eps = 1e-8, double weight_decay = 0)
        : beta1(beta1), beta2(beta2), eps(eps), weight_decay(weight_decay)
    {
    }

    void add_param_group(const std::vector<torch::Tensor>& params, double lr);
    // one step over every parameter with a defined gradient (FP32, contiguous, on the GPU)
    void step();
    void zero_grad();
    // lr of every group *= factor (the reference's UpdateLearningRate)
    void update_lr(double factor);

    struct State
    {
        torch::Tensor exp_avg, exp_avg_sq;
        int64_t step = 0;
    };
    struct Group
    {
        std::vector<torch::Tensor> params;
        std::vector<State> state;
        double lr;
    };
    std::vector<Group> groups;
    double beta1, beta2, eps, weight_decay;
};
