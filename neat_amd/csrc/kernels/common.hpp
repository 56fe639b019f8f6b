// Shared device helpers and C-ABI plumbing of libneat_amd (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "neat_amd.h"

namespace neat
{

// ---- error plumbing -------------------------------------------------------------------------------------------
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

#define NEAT_REQUIRE(cond, ...)                                           \
    do                                                                    \
    {                                                                     \
        if (!(cond)) return ::neat::fail(NEAT_E_ARG, __VA_ARGS__);        \
    } while (0)

#define NEAT_HIP(call)                                                                                    \
    do                                                                                                    \
    {                                                                                                     \
        hipError_t e_ = (call);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return ::neat::fail(NEAT_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                                __LINE__);                                                                \
    } while (0)

// Checks the launch that was just enqueued.
#define NEAT_LAUNCHED() NEAT_HIP(hipGetLastError())

inline hipStream_t as_stream(neat_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// ---- exact ray / box slab test -------------------------------------------------------------------------------------
// IntersectBoxRayPrecise (HyperTree.h:41-83): float slab distances, double running interval with the reference's
// sentinels, per-axis early exit, result narrowed to float. Compiled without contraction (-ffp-contract=off) so it
// is bit-identical to the oracle.
struct Hit
{
    bool hit;
    float tmin, tmax;
};

__device__ __forceinline__ Hit slab_test(const float* pos_min, const float* pos_max, const float o[3],
                                         const float d[3])
{
    double t_near = -236436436;
    double t_far  = 43637575;
#pragma unroll
    for (int i = 0; i < 3; i++)
    {
        if (d[i] == 0)
        {
            if ((o[i] < pos_min[i]) || (o[i] >= pos_max[i])) return {false, 0.f, 0.f};
        }
        else
        {
            float T_1 = (pos_min[i] - o[i]) / d[i];
            float T_2 = (pos_max[i] - o[i]) / d[i];
            if (T_1 > T_2)
            {
                float tmp = T_1;
                T_1       = T_2;
                T_2       = tmp;
            }
            if (T_1 > t_near) t_near = T_1;
            if (T_2 < t_far) t_far = T_2;
            if ((t_near > t_far) || (t_far < 0)) return {false, 0.f, 0.f};
        }
    }
    return {true, (float)t_near, (float)t_far};
}

// ---- jitter: splitmix64-finaliser hash of (seed, ray, segment rank, j) -----------------------------------------------
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// the part of the hash that depends on the seed and the ray only
__host__ __device__ __forceinline__ uint64_t jitter_ray_hash(uint64_t seed, int64_t ray)
{
    return mix64(seed ^ ((uint64_t)ray * 0x9E3779B97F4A7C15ull));
}

__host__ __device__ __forceinline__ float jitter_a_h(uint64_t hray, int32_t rank, int32_t j)
{
    const uint64_t h = mix64(hray ^ (((uint64_t)(uint32_t)rank << 32) | (uint64_t)(uint32_t)j));
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}

__host__ __device__ __forceinline__ float jitter_a(uint64_t seed, int64_t ray, int32_t rank, int32_t j)
{
    uint64_t h = jitter_ray_hash(seed, ray);
    h          = mix64(h ^ (((uint64_t)(uint32_t)rank << 32) | (uint64_t)(uint32_t)j));
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}

// Position of sample j in a segment (HyperTree.cu:339-358).
__device__ __forceinline__ float sample_t(float tmin, float step, int32_t j, float a)
{
    float t1 = tmin + (float)j * step;
    float t2 = tmin + (float)(j + 1) * step;
    return t1 * (1.0f - a) + t2 * a;
}

__device__ __forceinline__ float seg_a(int32_t jitter, uint64_t seed, int64_t ray, int32_t rank, int32_t j)
{
    return jitter ? jitter_a(seed, ray, rank, j) : 0.5f;
}

// Midpoint-cell weight (HyperTree.cu:372-389), t of the neighbours recomputed rather than re-read.
__device__ __forceinline__ float sample_w(float tmin, float tmax, float step, int32_t n, int32_t j, float t,
                                          int32_t jitter, uint64_t seed, int64_t ray, int32_t rank)
{
    float lo = (j == 0) ? tmin : (sample_t(tmin, step, j - 1, seg_a(jitter, seed, ray, rank, j - 1)) + t) * 0.5f;
    float hi = (j == n - 1) ? tmax : (sample_t(tmin, step, j + 1, seg_a(jitter, seed, ray, rank, j + 1)) + t) * 0.5f;
    return hi - lo;
}

// ---- activations ----------------------------------------------------------------------------------------------------
__device__ __forceinline__ float act_f(int a, float x, float beta)
{
    switch (a)
    {
        case NEAT_ACT_RELU:
            return x > 0.f ? x : 0.f;
        case NEAT_ACT_SILU:
            return x / (1.f + __expf(-x));
        case NEAT_ACT_SOFTPLUS:
            return (x * beta > 20.f) ? x : log1pf(__expf(x * beta)) / beta;
        case NEAT_ACT_ABS:
            return fabsf(x);
        default:
            return x;
    }
}

__device__ __forceinline__ float act_d(int a, float x, float beta)
{
    switch (a)
    {
        case NEAT_ACT_RELU:
            return x > 0.f ? 1.f : 0.f;
        case NEAT_ACT_SILU:
        {
            float s = 1.f / (1.f + __expf(-x));
            return s * (1.f + x * (1.f - s));
        }
        case NEAT_ACT_SOFTPLUS:
        {
            if (x * beta > 20.f) return 1.f;
            float z = __expf(x * beta);
            return z / (z + 1.f);
        }
        case NEAT_ACT_ABS:
            return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f);
        default:
            return 1.f;
    }
}

// ---- exclusive scan of int64 in place: a[0..n) counts -> a[0..n] offsets (a[n] = total) --------------------------------
size_t scan_workspace_bytes(int64_t n);
int exclusive_scan_i64(hipStream_t stream, int64_t* a, int64_t n, void* workspace, size_t workspace_bytes);

}  // namespace neat
