#include "IndirectGridSample3D.h"

#include <torch/csrc/autograd/custom_function.h>

using namespace neat_host;

namespace torch::autograd
{

std::vector<torch::Tensor> IndirectGridSample3D::forward(AutogradContext* ctx, torch::Tensor multi_grid,
                                                         torch::Tensor index, torch::Tensor uv)
{
    check_cuda(multi_grid, "multi_grid");
    TORCH_CHECK(multi_grid.scalar_type() == torch::kFloat32, "multi_grid must be float32");
    auto idx = index.to(torch::kLong).contiguous();
    auto u   = uv.to(torch::kFloat32).contiguous();
    const int64_t S = idx.numel();
    auto out = torch::zeros({S, multi_grid.size(1)}, multi_grid.options());
    int64_t st[5];
    for (int i = 0; i < 5; ++i) st[i] = multi_grid.stride(i);
    NEAT_CALL(neat_gs3d_fwd(current_stream(), dptr<float>(multi_grid), multi_grid.size(0), (int32_t)multi_grid.size(1),
                            (int32_t)multi_grid.size(2), (int32_t)multi_grid.size(3), (int32_t)multi_grid.size(4), st,
                            dptr<int64_t>(idx), dptr<float>(u), S, dptr<float>(out)));
    ctx->save_for_backward({multi_grid, idx, u});
    return {out};
}

std::vector<torch::Tensor> IndirectGridSample3D::backward(AutogradContext* ctx, std::vector<torch::Tensor> grad_output)
{
    auto saved      = ctx->get_saved_variables();
    auto multi_grid = saved[0];
    auto idx        = saved[1];
    auto u          = saved[2];
    TORCH_CHECK(grad_output.size() == 1);
    auto go          = grad_output[0].contiguous();
    auto grad_grid   = torch::zeros_like(multi_grid);  // same memory format, so the same strides
    auto grad_uv     = torch::zeros_like(u);
    const int64_t S  = idx.numel();
    int64_t st[5];
    for (int i = 0; i < 5; ++i) st[i] = multi_grid.stride(i);
    TORCH_CHECK(grad_grid.strides() == multi_grid.strides(), "gradient layout differs from the grid's");
    NEAT_CALL(neat_gs3d_bwd(current_stream(), dptr<float>(multi_grid), multi_grid.size(0), (int32_t)multi_grid.size(1),
                            (int32_t)multi_grid.size(2), (int32_t)multi_grid.size(3), (int32_t)multi_grid.size(4), st,
                            dptr<int64_t>(idx), dptr<float>(u), S, dptr<float>(go), dptr<float>(grad_grid),
                            dptr<float>(grad_uv)));
    return {grad_grid, torch::Tensor(), grad_uv};
}

}  // namespace torch::autograd

torch::Tensor IndirectGridSample3D(torch::Tensor multi_grid, torch::Tensor index, torch::Tensor uv)
{
    TORCH_CHECK(multi_grid.defined() && index.defined() && uv.defined());
    TORCH_CHECK(multi_grid.dim() == 5, "multi_grid must be [nodes, channels, z, y, x]");
    TORCH_CHECK(index.dim() == 1, "index must be [N]");
    TORCH_CHECK(uv.dim() == 2 && uv.size(1) == 3 && uv.size(0) == index.size(0), "uv must be [N, 3]");
    auto result = torch::autograd::IndirectGridSample3D::apply(multi_grid, index, uv);
    TORCH_CHECK(result.size() == 1);
    return result.front();
}
