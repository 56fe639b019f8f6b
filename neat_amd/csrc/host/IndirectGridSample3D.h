// IndirectGridSample3D (IndirectGridSample3D.h:17-30 in the reference): out[i, c] = trilinear(multi_grid[index[i], c],
// uv[i]) with align_corners and border clamp, as a torch::autograd::Function whose forward and backward are the C
// ABI's kernels (GPU only: the reference's CPU loops are the oracle here, not a runtime path).
#pragma once

#include "neat_host.h"

namespace torch::autograd
{
struct IndirectGridSample3D : public Function<IndirectGridSample3D>
{
    static std::vector<torch::Tensor> forward(AutogradContext* ctx, torch::Tensor multi_grid, torch::Tensor index,
                                              torch::Tensor uv);
    static std::vector<torch::Tensor> backward(AutogradContext* ctx, std::vector<torch::Tensor> grad_output);
};
}  // namespace torch::autograd

// Input:
//      multi_grid: [nodes, channels, z, y, x] (any strides; channels-last-3d is the fast layout)
//      index: long [N] slot per sample; uv: float [N, 3]
// Output: float [N, channels]
torch::Tensor IndirectGridSample3D(torch::Tensor multi_grid, torch::Tensor index, torch::Tensor uv);
