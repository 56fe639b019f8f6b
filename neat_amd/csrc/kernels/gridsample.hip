// Indirect trilinear grid sample, forward and backward (IndirectGridSample3D.cu:432-572 and :584-791).
//
// One thread per sample. The grid is addressed through element strides (N, C, Z, Y, X); when the channel stride is
// 1 (channels-last-3d, the layout the module layer stores its grid in) each corner's C channels are one contiguous
// run and are read with 16-byte vector loads. The arithmetic (corner weights from the unclamped neighbours, indices
// clamped to the border, per-channel sum in the corner order tnw tne tsw tse bnw bne bsw bse) follows the reference
// and is compiled without contraction, so the forward is bit-identical to the oracle.
#include "gridsample_dev.hpp"

namespace neat
{

// Channel-last fast path: C4 groups of 4 channels read with one 16-byte load per corner.
template <bool kChannelsLast>
__global__ void __launch_bounds__(256) k_gs3d_fwd(GridArgs g, const int64_t* index, const float* uv, int64_t S,
                                                  float* out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    const int64_t n = index[i];
    if (n < 0 || n >= g.N)
    {
        // a slot outside the grid (e.g. -1 for an inactive node) samples zeros instead of reading out of bounds
        for (int c = 0; c < g.C; ++c) out[i * g.C + c] = 0.f;
        return;
    }
    Tri t = trilinear(g, n, uv[3 * i], uv[3 * i + 1], uv[3 * i + 2]);
    if constexpr (kChannelsLast)
    {
        for (int c = 0; c < g.C; c += 4)
        {
            float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int k = 0; k < 8; ++k)
            {
                float4 v = *reinterpret_cast<const float4*>(g.grid + t.off[k] + c);
                acc.x += v.x * t.w[k];
                acc.y += v.y * t.w[k];
                acc.z += v.z * t.w[k];
                acc.w += v.w * t.w[k];
            }
            *reinterpret_cast<float4*>(out + i * g.C + c) = acc;
        }
    }
    else
    {
        for (int c = 0; c < g.C; ++c)
        {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < 8; ++k) acc += g.grid[t.off[k] + c * g.s1] * t.w[k];
            out[i * g.C + c] = acc;
        }
    }
}

// Backward: grid gradient with float atomics (like the reference, IndirectGridSample3D.cu:691-698) and the uv
// gradient with the reference's product order (:752-790).
__global__ void __launch_bounds__(256) k_gs3d_bwd(GridArgs g, const int64_t* index, const float* uv, int64_t S,
                                                  const float* grad_out, float* grad_grid, float* grad_uv)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    const int64_t n = index[i];
    if (n < 0 || n >= g.N)
    {
        if (grad_uv) grad_uv[3 * i] = grad_uv[3 * i + 1] = grad_uv[3 * i + 2] = 0.f;
        return;
    }
    Tri t    = trilinear(g, n, uv[3 * i], uv[3 * i + 1], uv[3 * i + 2]);
    float ix = t.ix, iy = t.iy, iz = t.iz;
    float x0 = (float)t.x0, y0 = (float)t.y0, z0 = (float)t.z0;
    float x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
    float gix = 0.f, giy = 0.f, giz = 0.f;
    for (int c = 0; c < g.C; ++c)
    {
        float go = grad_out[i * g.C + c];
        float val[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
        {
            int64_t o = t.off[k] + c * g.s1;
            if (grad_grid) atomicAdd(grad_grid + o, t.w[k] * go);
            val[k] = g.grid[o];
        }
        if (grad_uv)
        {
            float m1 = -1;
            gix += m1 * val[0] * (y1 - iy) * (z1 - iz) * go;
            gix += val[1] * (y1 - iy) * (z1 - iz) * go;
            gix += m1 * val[2] * (iy - y0) * (z1 - iz) * go;
            gix += val[3] * (iy - y0) * (z1 - iz) * go;
            gix += m1 * val[4] * (y1 - iy) * (iz - z0) * go;
            gix += val[5] * (y1 - iy) * (iz - z0) * go;
            gix += m1 * val[6] * (iy - y0) * (iz - z0) * go;
            gix += val[7] * (iy - y0) * (iz - z0) * go;

            giy += m1 * val[0] * (x1 - ix) * (z1 - iz) * go;
            giy += m1 * val[1] * (ix - x0) * (z1 - iz) * go;
            giy += val[2] * (x1 - ix) * (z1 - iz) * go;
            giy += val[3] * (ix - x0) * (z1 - iz) * go;
            giy += m1 * val[4] * (x1 - ix) * (iz - z0) * go;
            giy += m1 * val[5] * (ix - x0) * (iz - z0) * go;
            giy += val[6] * (x1 - ix) * (iz - z0) * go;
            giy += val[7] * (ix - x0) * (iz - z0) * go;

            giz += m1 * val[0] * (x1 - ix) * (y1 - iy) * go;
            giz += m1 * val[1] * (ix - x0) * (y1 - iy) * go;
            giz += m1 * val[2] * (x1 - ix) * (iy - y0) * go;
            giz += m1 * val[3] * (ix - x0) * (iy - y0) * go;
            giz += val[4] * (x1 - ix) * (y1 - iy) * go;
            giz += val[5] * (ix - x0) * (y1 - iy) * go;
            giz += val[6] * (x1 - ix) * (iy - y0) * go;
            giz += val[7] * (ix - x0) * (iy - y0) * go;
        }
    }
    if (grad_uv)
    {
        grad_uv[3 * i + 0] = gix * (float)(g.X - 1) / 2;
        grad_uv[3 * i + 1] = giy * (float)(g.Y - 1) / 2;
        grad_uv[3 * i + 2] = giz * (float)(g.Z - 1) / 2;
    }
}

}  // namespace neat

using namespace neat;

static int make_grid_args(GridArgs& g, const float* grid, int64_t N, int32_t C, int32_t Z, int32_t Y, int32_t X,
                          const int64_t* strides)
{
    NEAT_REQUIRE(grid && strides, "grid and strides must not be null");
    NEAT_REQUIRE(N > 0 && C > 0 && Z > 0 && Y > 0 && X > 0, "grid dims must be positive");
    g = GridArgs{grid, strides[0], strides[1], strides[2], strides[3], strides[4], N, C, Z, Y, X};
    return NEAT_OK;
}

static bool channels_last(const GridArgs& g)
{
    return g.s1 == 1 && g.C % 4 == 0 && g.s4 % 4 == 0 && g.s3 % 4 == 0 && g.s2 % 4 == 0 && g.s0 % 4 == 0 &&
           (reinterpret_cast<uintptr_t>(g.grid) % 16) == 0;
}

extern "C"
{
    int neat_gs3d_fwd(neat_stream_t stream, const float* grid, int64_t N, int32_t C, int32_t Z, int32_t Y, int32_t X,
                      const int64_t* strides, const int64_t* index, const float* uv, int64_t S, float* out)
    {
        GridArgs g;
        int st = make_grid_args(g, grid, N, C, Z, Y, X, strides);
        if (st) return st;
        NEAT_REQUIRE(S >= 0, "num_samples < 0");
        if (S == 0) return NEAT_OK;
        NEAT_REQUIRE(index && uv && out, "neat_gs3d_fwd: null pointer");
        hipStream_t s = as_stream(stream);
        if (channels_last(g) && (reinterpret_cast<uintptr_t>(out) % 16) == 0)
            k_gs3d_fwd<true><<<grid_for(S, 256), 256, 0, s>>>(g, index, uv, S, out);
        else
            k_gs3d_fwd<false><<<grid_for(S, 256), 256, 0, s>>>(g, index, uv, S, out);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_gs3d_bwd(neat_stream_t stream, const float* grid, int64_t N, int32_t C, int32_t Z, int32_t Y, int32_t X,
                      const int64_t* strides, const int64_t* index, const float* uv, int64_t S, const float* grad_out,
                      float* grad_grid, float* grad_uv)
    {
        GridArgs g;
        int st = make_grid_args(g, grid, N, C, Z, Y, X, strides);
        if (st) return st;
        NEAT_REQUIRE(S >= 0, "num_samples < 0");
        if (S == 0) return NEAT_OK;
        NEAT_REQUIRE(index && uv && grad_out, "neat_gs3d_bwd: null pointer");
        k_gs3d_bwd<<<grid_for(S, 256), 256, 0, as_stream(stream)>>>(g, index, uv, S, grad_out, grad_grid, grad_uv);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }
}
