#include "optim.h"

#include "neat_host.h"

using namespace neat_host;

void NeatAdam::add_param_group(const std::vector<torch::Tensor>& params, double lr)
{
    Group g;
    g.params = params;
    g.state.resize(params.size());
    g.lr = lr;
    groups.push_back(std::move(g));
}

void NeatAdam::step()
{
    torch::NoGradGuard ng;
    for (auto& g : groups)
    {
        for (size_t i = 0; i < g.params.size(); ++i)
        {
            auto& p    = g.params[i];
            auto grad  = p.grad();
            if (!grad.defined()) continue;
            TORCH_CHECK(p.scalar_type() == torch::kFloat32 && p.is_cuda(), "NeatAdam: FP32 GPU parameters only");
            TORCH_CHECK(p.is_contiguous() || p.is_contiguous(at::MemoryFormat::ChannelsLast3d),
                        "NeatAdam: dense parameters only");
            TORCH_CHECK(grad.strides() == p.strides(), "NeatAdam: gradient layout differs from the parameter's");
            auto& st = g.state[i];
            if (!st.exp_avg.defined())
            {
                st.exp_avg    = torch::zeros_like(p);
                st.exp_avg_sq = torch::zeros_like(p);
            }
            // moments made for another shape or layout would pair the wrong elements (torch::optim::Adam raises too)
            TORCH_CHECK(st.exp_avg.sizes() == p.sizes() && st.exp_avg.strides() == p.strides() &&
                            st.exp_avg_sq.strides() == p.strides(),
                        "NeatAdam: the parameter's shape or layout changed after its moments were created");
            st.step += 1;
            NEAT_CALL(neat_adam_step(current_stream(), dptr<float>(p), dptr<float>(grad), dptr<float>(st.exp_avg),
                                     dptr<float>(st.exp_avg_sq), p.numel(), g.lr, beta1, beta2, eps, weight_decay,
                                     st.step));
        }
    }
}

void NeatAdam::zero_grad()
{
    for (auto& g : groups)
        for (auto& p : g.params)
            if (p.grad().defined()) p.mutable_grad().zero_();
}

void NeatAdam::update_lr(double factor)
{
    for (auto& g : groups) g.lr *= factor;
}
