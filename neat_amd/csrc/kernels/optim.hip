// Adam step for the grid and decoder parameters: the update of torch::optim::Adam as the reference runs it
// (geometry.h:66-88 OptimizerStep; param groups geometry_ex_ex.cpp:40-49 and geometry.cpp:90-97), in one pass over
// param / grad / exp_avg / exp_avg_sq. Purely HBM-bound: 16 B read + 12 B written per parameter (20 B more read and
// nothing more written with weight decay off).
//   exp_avg    = beta1 exp_avg + (1 - beta1) grad
//   exp_avg_sq = beta2 exp_avg_sq + (1 - beta2) grad^2
//   param     -= lr / (1 - beta1^step) * exp_avg / (sqrt(exp_avg_sq) / sqrt(1 - beta2^step) + eps)
#include <cmath>

#include "common.hpp"

namespace neat
{
namespace
{

struct AdamArgs
{
    float lr_bc1;     // lr / (1 - beta1^step)
    float bc2_sqrt;   // sqrt(1 - beta2^step)
    float beta1, one_minus_beta1;
    float beta2, one_minus_beta2;
    float eps, weight_decay;
};

__device__ __forceinline__ void adam_one(float& p, float g, float& m, float& v, const AdamArgs& a)
{
    if (a.weight_decay != 0.f) g = g + p * a.weight_decay;
    m = m * a.beta1 + g * a.one_minus_beta1;
    v = v * a.beta2 + (g * g) * a.one_minus_beta2;
    const float denom = sqrtf(v) / a.bc2_sqrt + a.eps;
    p = p - a.lr_bc1 * (m / denom);
}

__global__ void __launch_bounds__(256) k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, int64_t n, AdamArgs a)
{
    const int64_t n4     = n / 4;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
    {
        float4 pp = reinterpret_cast<float4*>(p)[i];
        float4 gg = reinterpret_cast<const float4*>(g)[i];
        float4 mm = reinterpret_cast<float4*>(m)[i];
        float4 vv = reinterpret_cast<float4*>(v)[i];
        adam_one(pp.x, gg.x, mm.x, vv.x, a);
        adam_one(pp.y, gg.y, mm.y, vv.y, a);
        adam_one(pp.z, gg.z, mm.z, vv.z, a);
        adam_one(pp.w, gg.w, mm.w, vv.w, a);
        reinterpret_cast<float4*>(p)[i] = pp;
        reinterpret_cast<float4*>(m)[i] = mm;
        reinterpret_cast<float4*>(v)[i] = vv;
    }
    for (int64_t i = 4 * n4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        adam_one(p[i], g[i], m[i], v[i], a);
}

}  // namespace
}  // namespace neat

using namespace neat;

extern "C" int neat_adam_step(neat_stream_t stream, float* param, const float* grad, float* exp_avg,
                              float* exp_avg_sq, int64_t n, double lr, double beta1, double beta2, double eps,
                              double weight_decay, int64_t step)
{
    NEAT_REQUIRE(n >= 0 && step >= 1, "neat_adam_step: bad sizes (n %lld, step %lld)", (long long)n, (long long)step);
    if (n == 0) return NEAT_OK;
    NEAT_REQUIRE(param && grad && exp_avg && exp_avg_sq, "neat_adam_step: null pointer");
    const uintptr_t mis = reinterpret_cast<uintptr_t>(param) | reinterpret_cast<uintptr_t>(grad) |
                          reinterpret_cast<uintptr_t>(exp_avg) | reinterpret_cast<uintptr_t>(exp_avg_sq);
    NEAT_REQUIRE((mis & 15) == 0, "neat_adam_step: buffers must be 16-byte aligned");
    NEAT_REQUIRE(beta1 >= 0 && beta1 < 1 && beta2 >= 0 && beta2 < 1 && eps >= 0 && lr >= 0,
                 "neat_adam_step: bad hyper-parameters");
    AdamArgs a;
    const double bc1 = 1.0 - std::pow(beta1, (double)step);
    const double bc2 = 1.0 - std::pow(beta2, (double)step);
    a.lr_bc1          = (float)(lr / bc1);
    a.bc2_sqrt        = (float)std::sqrt(bc2);
    a.beta1           = (float)beta1;
    a.one_minus_beta1 = (float)(1.0 - beta1);
    a.beta2           = (float)beta2;
    a.one_minus_beta2 = (float)(1.0 - beta2);
    a.eps             = (float)eps;
    a.weight_decay    = (float)weight_decay;
    const int64_t blocks = std::min<int64_t>((n / 4 + 255) / 256 + 1, 8192);
    k_adam<<<(unsigned)blocks, 256, 0, as_stream(stream)>>>(param, grad, exp_avg, exp_avg_sq, n, a);
    NEAT_LAUNCHED();
    return NEAT_OK;
}
