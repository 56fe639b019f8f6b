// Per-ray line integral (IntegrateSamplesXRay, geometry.cpp:393-424), local coordinates (ComputeLocalSamples,
// HyperTree.cu:177-192), node -> grid slot (GlobalNodeIdToLocalActiveId, HyperTree.h:183-187) and the FP64 MSE loss
// (reconstruct.cpp:378-381).
#include "common.hpp"

namespace neat
{

// Segmented, deterministic: one thread per (channel, ray) sums the ray's contiguous samples in order.
__global__ void __launch_bounds__(256) k_integrate_seg(const float* v, const float* w, int64_t S, int32_t C,
                                                       int64_t R, const int64_t* off, float* out)
{
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= R * C) return;
    int64_t r = idx % R;
    int32_t c = (int32_t)(idx / R);
    float acc = 0.f;
    for (int64_t i = off[r]; i < off[r + 1]; ++i) acc += v[i * C + c] * w[i];
    out[(int64_t)c * R + r] = acc;
}

__global__ void __launch_bounds__(256) k_zero_f32(float* p, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// Unordered input: float atomics, like the reference's index_add_ on the GPU. Out-of-range ray ids are skipped.
__global__ void __launch_bounds__(256) k_integrate_atomic(const float* v, const float* w, const int64_t* ray,
                                                          int64_t S, int32_t C, int64_t R, float* out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    int64_t r = ray[i];
    if (r < 0 || r >= R) return;
    for (int c = 0; c < C; ++c) atomicAdd(out + (int64_t)c * R + r, v[i * C + c] * w[i]);
}

__global__ void __launch_bounds__(256) k_integrate_bwd(const float* g, const float* w, const int64_t* ray, int64_t S,
                                                       int32_t C, int64_t R, float* gv)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    int64_t r = ray[i];
    bool ok   = r >= 0 && r < R;
    for (int c = 0; c < C; ++c) gv[i * C + c] = ok ? g[(int64_t)c * R + r] * w[i] : 0.f;
}

// u = (g - min) / (max - min) * 2 - 1, each step rounded separately like the chain of torch ops it replaces.
__global__ void __launch_bounds__(256) k_local(const float* g, const int64_t* node, int64_t S, const float* pmin,
                                               const float* pmax, int32_t num_nodes, float* out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    int64_t n = node[i];
    if (n < 0 || n >= num_nodes)
    {
        out[3 * i] = out[3 * i + 1] = out[3 * i + 2] = 0.f;
        return;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        float mn = pmin[3 * n + d], mx = pmax[3 * n + d];
        float size = mx - mn;
        out[3 * i + d] = (g[3 * i + d] - mn) / size * 2 - 1;
    }
}

// Autograd of the expression above with respect to g: (grad * 2) / size.
__global__ void __launch_bounds__(256) k_local_bwd(const float* gl, const int64_t* node, int64_t S, const float* pmin,
                                                   const float* pmax, int32_t num_nodes, float* gg)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    int64_t n = node[i];
    if (n < 0 || n >= num_nodes)
    {
        gg[3 * i] = gg[3 * i + 1] = gg[3 * i + 2] = 0.f;
        return;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        float size     = pmax[3 * n + d] - pmin[3 * n + d];
        gg[3 * i + d] = (gl[3 * i + d] * 2) / size;
    }
}

__global__ void __launch_bounds__(256) k_slot(const int64_t* node, int64_t S, const int64_t* prefix, int32_t num_nodes,
                                              int64_t* out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    int64_t n = node[i];
    out[i]    = (n >= 0 && n < num_nodes) ? prefix[n] : -1;
}

// ---- MSE in FP64 -----------------------------------------------------------------------------------------------------
constexpr int kMseThreads = 256;
constexpr int kMseBlocks  = 512;

__global__ void __launch_bounds__(kMseThreads) k_mse_partial(const float* p, const float* t, int64_t n,
                                                             float grad_scale, float* grad, double* partial)
{
    __shared__ double lds[kMseThreads];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    {
        double d = (double)p[i] - (double)t[i];
        acc += d * d;
        if (grad) grad[i] = (float)(2.0 * d / (double)n) * grad_scale;
    }
    lds[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kMseThreads / 2; s > 0; s >>= 1)
    {
        if (threadIdx.x < s) lds[threadIdx.x] += lds[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = lds[0];
}

__global__ void __launch_bounds__(kMseThreads) k_mse_final(const double* partial, int nb, int64_t n, double* out)
{
    __shared__ double lds[kMseThreads];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += partial[b];
    lds[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kMseThreads / 2; s > 0; s >>= 1)
    {
        if (threadIdx.x < s) lds[threadIdx.x] += lds[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = n > 0 ? lds[0] / (double)n : 0.0;
}

}  // namespace neat

using namespace neat;

extern "C"
{
    int neat_integrate_fwd(neat_stream_t stream, const float* v, const float* w, const int64_t* ray, int64_t S,
                           int32_t C, int64_t R, const int64_t* off, float* out)
    {
        NEAT_REQUIRE(S >= 0 && R >= 0 && C >= 1, "neat_integrate_fwd: bad sizes");
        NEAT_REQUIRE(out || R == 0, "neat_integrate_fwd: null output");
        hipStream_t s = as_stream(stream);
        if (R == 0) return NEAT_OK;
        if (off)
        {
            NEAT_REQUIRE(S == 0 || (v && w), "neat_integrate_fwd: null input");
            k_integrate_seg<<<grid_for(R * C, 256), 256, 0, s>>>(v, w, S, C, R, off, out);
            NEAT_LAUNCHED();
            return NEAT_OK;
        }
        k_zero_f32<<<grid_for(R * C, 256), 256, 0, s>>>(out, R * C);
        NEAT_LAUNCHED();
        if (S == 0) return NEAT_OK;
        NEAT_REQUIRE(v && w && ray, "neat_integrate_fwd: null input");
        k_integrate_atomic<<<grid_for(S, 256), 256, 0, s>>>(v, w, ray, S, C, R, out);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_integrate_bwd(neat_stream_t stream, const float* g, const float* w, const int64_t* ray, int64_t S,
                           int32_t C, int64_t R, float* gv)
    {
        NEAT_REQUIRE(S >= 0 && R >= 0 && C >= 1, "neat_integrate_bwd: bad sizes");
        if (S == 0) return NEAT_OK;
        NEAT_REQUIRE(g && w && ray && gv, "neat_integrate_bwd: null pointer");
        k_integrate_bwd<<<grid_for(S, 256), 256, 0, as_stream(stream)>>>(g, w, ray, S, C, R, gv);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_local_samples(neat_stream_t stream, const float* g, const int64_t* node, int64_t S, const float* pmin,
                           const float* pmax, int32_t num_nodes, float* out)
    {
        NEAT_REQUIRE(S >= 0, "num_samples < 0");
        if (S == 0) return NEAT_OK;
        NEAT_REQUIRE(g && node && pmin && pmax && out, "neat_local_samples: null pointer");
        k_local<<<grid_for(S, 256), 256, 0, as_stream(stream)>>>(g, node, S, pmin, pmax, num_nodes, out);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_local_samples_bwd(neat_stream_t stream, const float* gl, const int64_t* node, int64_t S,
                               const float* pmin, const float* pmax, int32_t num_nodes, float* gg)
    {
        NEAT_REQUIRE(S >= 0, "num_samples < 0");
        if (S == 0) return NEAT_OK;
        NEAT_REQUIRE(gl && node && pmin && pmax && gg, "neat_local_samples_bwd: null pointer");
        k_local_bwd<<<grid_for(S, 256), 256, 0, as_stream(stream)>>>(gl, node, S, pmin, pmax, num_nodes, gg);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_node_to_slot(neat_stream_t stream, const int64_t* node, int64_t S, const int64_t* prefix,
                          int32_t num_nodes, int64_t* out)
    {
        NEAT_REQUIRE(S >= 0, "num_samples < 0");
        if (S == 0) return NEAT_OK;
        NEAT_REQUIRE(node && prefix && out, "neat_node_to_slot: null pointer");
        k_slot<<<grid_for(S, 256), 256, 0, as_stream(stream)>>>(node, S, prefix, num_nodes, out);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    size_t neat_mse_workspace_bytes(int64_t n) { return kMseBlocks * sizeof(double); }

    int neat_mse_fp64(neat_stream_t stream, const float* p, const float* t, int64_t n, float grad_scale,
                      double* out_loss, float* grad, void* workspace, size_t workspace_bytes)
    {
        NEAT_REQUIRE(n >= 0 && out_loss, "neat_mse_fp64: bad arguments");
        NEAT_REQUIRE(n == 0 || (p && t), "neat_mse_fp64: null input");
        if (!workspace || workspace_bytes < neat_mse_workspace_bytes(n))
            return fail(NEAT_E_OVERFLOW, "mse workspace too small");
        hipStream_t s = as_stream(stream);
        double* partial = static_cast<double*>(workspace);
        int nb          = (int)std::max<int64_t>(1, std::min<int64_t>(kMseBlocks, (n + kMseThreads - 1) / kMseThreads));
        k_mse_partial<<<nb, kMseThreads, 0, s>>>(p, t, n, grad_scale, grad, partial);
        NEAT_LAUNCHED();
        k_mse_final<<<1, kMseThreads, 0, s>>>(partial, nb, n, out_loss);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }
}
