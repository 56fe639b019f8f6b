// Decoder MLP (FCBlockImpl, ImplicitNet.h:18-71) fused with DecodeFeatures' last activation (geometry.cpp:123-143),
// forward and backward in FP32.
//
// The projector's decoder shape (8 inputs, one hidden layer of 32 or 64 relu or silu units, one output) runs on the
// matrix cores (v_mfma_f32_32x32x2f32, the tiles of mfma_mlp.hpp): persistent workgroups of 4 waves, one tile of 32
// samples per wave at a time, weights laid out once per workgroup in LDS. Other shapes run the vector-ALU tiles of
// mlp_tile.hpp, one wave per tile of 64 samples.
//
// Backward (both forms): each wave or workgroup recomputes its tiles' forward, back-propagates, and accumulates its
// share of the parameter gradient into its own slice of a partials buffer (every entry written by one thread); a
// second kernel sums the slices in a fixed order, so the result is deterministic.
#include "mfma_mlp.hpp"
#include "mlp_tile.hpp"

namespace neat
{

__device__ __forceinline__ void load_x_col(const float* x, int64_t i0, int ns, int in, float* X)
{
    const int s = threadIdx.x & (kTile - 1);
#pragma unroll
    for (int k = 0; k < kMaxIn; ++k) X[k * kS + s] = (s < ns && k < in) ? x[(i0 + s) * in + k] : 0.f;
}

template <int H>
__global__ void __launch_bounds__(kTile) k_dec_fwd(const float* x, int64_t S, MLP m, const float* __restrict__ wt,
                                                   int act, int last_act, float beta, float* y)
{
    __shared__ float lds[kTileFloats];
    TileBufs t       = tile_bufs(lds);
    const int64_t i0 = (int64_t)blockIdx.x * kTile;
    const int ns     = (int)min<int64_t>(kTile, S - i0);
    const int s      = threadIdx.x;
    load_x_col(x, i0, ns, m.in, t.X);
    __syncthreads();
    tile_forward<H>(m, wt, WLayout{false}, act, beta, t);
    if (s < ns)
    {
        for (int o = 0; o < m.out; ++o) y[(i0 + s) * m.out + o] = act_f(last_act, t.D[o * kS + s], beta);
    }
}

template <int H>
__global__ void __launch_bounds__(kTile) k_dec_bwd(const float* x, int64_t S, MLP m, const float* __restrict__ wt,
                                                   int act, int last_act, float beta, const float* gy, float* gx,
                                                   float* partials)
{
    __shared__ float lds[kTileFloats];
    TileBufs t          = tile_bufs(lds);
    float* gacc         = partials + (int64_t)blockIdx.x * m.nparams();
    const int64_t tiles = (S + kTile - 1) / kTile;
    const int s         = threadIdx.x;
    for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x)
    {
        const int64_t i0 = tile * kTile;
        const int ns     = (int)min<int64_t>(kTile, S - i0);
        __syncthreads();
        load_x_col(x, i0, ns, m.in, t.X);
        __syncthreads();
        tile_forward<H>(m, wt, WLayout{true}, act, beta, t);
        for (int o = 0; o < m.out; ++o)
        {
            float* p = t.D + o * kS + s;
            *p       = s < ns ? gy[(i0 + s) * m.out + o] * act_d(last_act, *p, beta) : 0.f;
        }
        tile_backward<H>(m, wt, WLayout{true}, act, beta, t, ns, gacc);
        if (gx && s < ns)
        {
            for (int k = 0; k < m.in; ++k) gx[(i0 + s) * m.in + k] = t.X[k * kS + s];
        }
    }
}

// ---- matrix-core form (8 -> H -> H -> 1) ----------------------------------------------------------------------
constexpr int kDecMfmaWaves = 4;

// the lane's 4 inputs (channels 4g..4g+3 of sample i, g = lane / 32), 0 past the end
__device__ __forceinline__ void load_x4(const float* __restrict__ x, int64_t i, int64_t S, float (&x4)[4])
{
    const int hg = (threadIdx.x & 63) >> 5;
#pragma unroll
    for (int j = 0; j < 4; ++j) x4[j] = i < S ? x[i * 8 + 4 * hg + j] : 0.f;
}

template <int H, int ACT>
__global__ void __launch_bounds__(256, 2) k_dec_mfma_fwd(const float* __restrict__ x, int64_t S,
                                                      const float* __restrict__ params, int last_act, float beta,
                                                      float* __restrict__ y)
{
    typedef PrecF32 P;
    __shared__ WeightLDS<P, H> w;
    load_weights<P, H, 8>(params, w);
    __syncthreads();
    const int l = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t tiles = (S + 31) / 32;
    for (int64_t t = (int64_t)blockIdx.x * kDecMfmaWaves + wave; t < tiles; t += (int64_t)gridDim.x * kDecMfmaWaves)
    {
        const int64_t i = t * 32 + (l & 31);
        float x4[4];
        load_x4(x, i, S, x4);
        Fwd<H / 32> f;
        tile_forward_mfma<P, H, ACT>(w, x4, beta, f);
        if (i < S && l < 32) y[i] = act_f(last_act, f.z, beta);
    }
}

// One partial slice per wave (kDecMfmaWaves per workgroup).
template <int H, int ACT>
__global__ void __launch_bounds__(256) k_dec_mfma_bwd(const float* __restrict__ x, int64_t S,
                                                      const float* __restrict__ params, int last_act, float beta,
                                                      const float* __restrict__ gy, float* __restrict__ gx,
                                                      float* __restrict__ partials, int32_t P_total)
{
    typedef PrecF32 P;
    typedef Stage<P> ST;
    __shared__ WeightLDS<P, H> w;
    __shared__ __align__(16) typename ST::E stage[kDecMfmaWaves][ST::elems(H)];
    load_weights<P, H, 8>(params, w);
    const int l = threadIdx.x & 63, wave = threadIdx.x >> 6, hg = l >> 5;
    BwdStage<P, H> sb(stage[wave]);
    BwdAcc<H / 32> acc;
    __syncthreads();
    const int64_t tiles = (S + 31) / 32;
    for (int64_t t = (int64_t)blockIdx.x * kDecMfmaWaves + wave; t < tiles; t += (int64_t)gridDim.x * kDecMfmaWaves)
    {
        const int64_t i  = t * 32 + (l & 31);
        const bool valid = i < S;
        float x4[4];
        load_x4(x, i, S, x4);
        const f32x16 dx = tile_backward_mfma<P, H, ACT>(w, x4, valid ? gy[i] : 0.f, valid, last_act, beta, sb, acc);
        if (gx && valid)
        {
#pragma unroll
            for (int j = 0; j < 4; ++j) gx[i * 8 + 4 * hg + j] = dx[j];
        }
    }
    write_param_partials<H, 8>(acc, partials + ((int64_t)blockIdx.x * kDecMfmaWaves + wave) * P_total);
}

// the shapes the matrix-core form covers
bool dec_mfma(int in, int hidden, int layers, int out, int act)
{
    return in == 8 && out == 1 && layers == 1 && (hidden == 32 || hidden == 64) &&
           (act == NEAT_ACT_RELU || act == NEAT_ACT_SILU);
}

int device_cus()
{
    int dev = 0, cu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                hipSuccess || cu <= 0)
        cu = 256;
    return cu;
}

constexpr int kMaxBwdWG = 1024;

int bwd_workgroups(int64_t S)
{
    int64_t tiles = (S + kTile - 1) / kTile;
    return (int)std::max<int64_t>(1, std::min<int64_t>(tiles, kMaxBwdWG));
}

int check_mlp(int in, int hidden, int layers, int out, int act, int last_act)
{
    NEAT_REQUIRE(in >= 1 && in <= kMaxIn, "decoder in_features %d outside [1, %d]", in, kMaxIn);
    NEAT_REQUIRE(out >= 1 && out <= kMaxOut, "decoder out_features %d outside [1, %d]", out, kMaxOut);
    NEAT_REQUIRE(act >= 0 && act <= 4 && last_act >= 0 && last_act <= 4, "unknown activation id");
    if (!((hidden == 16 || hidden == 32 || hidden == 64) && (layers == 0 || layers == 1)))
        return fail(NEAT_E_UNSUPPORTED,
                    "decoder hidden=%d hidden_layers=%d: this build supports hidden {16,32,64} and hidden_layers {0,1}",
                    hidden, layers);
    return NEAT_OK;
}

}  // namespace neat

using namespace neat;

extern "C"
{
    int64_t neat_decoder_num_params(int32_t in, int32_t hidden, int32_t layers, int32_t out)
    {
        MLP m{in, hidden, layers, out};
        return m.nparams();
    }

    size_t neat_decoder_workspace_bytes(int64_t S, int32_t in, int32_t hidden, int32_t layers, int32_t out)
    {
        MLP m{in, hidden, layers, out};
        size_t P = (size_t)m.nparams();
        return (P + (size_t)bwd_workgroups(S) * P) * sizeof(float);
    }

    int neat_decoder_fwd(neat_stream_t stream, const float* x, int64_t S, int32_t in, int32_t hidden, int32_t layers,
                         int32_t out, const float* params, int32_t act, int32_t last_act, float beta, float* y)
    {
        int st = check_mlp(in, hidden, layers, out, act, last_act);
        if (st) return st;
        NEAT_REQUIRE(S >= 0, "num_samples < 0");
        if (S == 0) return NEAT_OK;
        NEAT_REQUIRE(x && params && y, "neat_decoder_fwd: null pointer");
        MLP m{in, hidden, layers, out};
        hipStream_t s = as_stream(stream);
        if (dec_mfma(in, hidden, layers, out, act))
        {
            const int64_t tiles = (S + 31) / 32;
            const unsigned wg =
                (unsigned)std::max<int64_t>(1, std::min<int64_t>((tiles + kDecMfmaWaves - 1) / kDecMfmaWaves,
                                                                 (int64_t)device_cus() * 8));
            if (hidden == 32 && act == NEAT_ACT_RELU)
                k_dec_mfma_fwd<32, NEAT_ACT_RELU><<<wg, 256, 0, s>>>(x, S, params, last_act, beta, y);
            else if (hidden == 32)
                k_dec_mfma_fwd<32, NEAT_ACT_SILU><<<wg, 256, 0, s>>>(x, S, params, last_act, beta, y);
            else if (act == NEAT_ACT_RELU)
                k_dec_mfma_fwd<64, NEAT_ACT_RELU><<<wg, 256, 0, s>>>(x, S, params, last_act, beta, y);
            else
                k_dec_mfma_fwd<64, NEAT_ACT_SILU><<<wg, 256, 0, s>>>(x, S, params, last_act, beta, y);
            NEAT_LAUNCHED();
            return NEAT_OK;
        }
        const float* wt = params;  // packed layout read in place
        unsigned grid   = grid_for(S, kTile);
        if (hidden == 16)
            k_dec_fwd<16><<<grid, kTile, 0, s>>>(x, S, m, wt, act, last_act, beta, y);
        else if (hidden == 32)
            k_dec_fwd<32><<<grid, kTile, 0, s>>>(x, S, m, wt, act, last_act, beta, y);
        else
            k_dec_fwd<64><<<grid, kTile, 0, s>>>(x, S, m, wt, act, last_act, beta, y);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_decoder_bwd(neat_stream_t stream, const float* x, int64_t S, int32_t in, int32_t hidden, int32_t layers,
                         int32_t out, const float* params, int32_t act, int32_t last_act, float beta, const float* gy,
                         float* gparams, float* gx, void* workspace, size_t workspace_bytes)
    {
        int st = check_mlp(in, hidden, layers, out, act, last_act);
        if (st) return st;
        NEAT_REQUIRE(S >= 0, "num_samples < 0");
        NEAT_REQUIRE(params && gparams, "neat_decoder_bwd: null pointer");
        MLP m{in, hidden, layers, out};
        const int P   = m.nparams();
        hipStream_t s = as_stream(stream);
        if (S == 0)
        {
            NEAT_HIP(hipMemsetAsync(gparams, 0, (size_t)P * sizeof(float), s));
            return NEAT_OK;
        }
        NEAT_REQUIRE(x && gy, "neat_decoder_bwd: null pointer");
        size_t need = neat_decoder_workspace_bytes(S, in, hidden, layers, out);
        if (!workspace || workspace_bytes < need)
            return fail(NEAT_E_OVERFLOW, "decoder workspace too small: %zu < %zu", workspace_bytes, need);
        float* wt       = static_cast<float*>(workspace);
        float* partials = wt + P;
        if (dec_mfma(in, hidden, layers, out, act))
        {
            // kDecMfmaWaves slices per workgroup, within the bwd_workgroups(S) slices the workspace holds
            const int wg = std::max(1, std::min(bwd_workgroups(S) / kDecMfmaWaves, device_cus()));
            if (hidden == 32 && act == NEAT_ACT_RELU)
                k_dec_mfma_bwd<32, NEAT_ACT_RELU><<<wg, 256, 0, s>>>(x, S, params, last_act, beta, gy, gx, partials, P);
            else if (hidden == 32)
                k_dec_mfma_bwd<32, NEAT_ACT_SILU><<<wg, 256, 0, s>>>(x, S, params, last_act, beta, gy, gx, partials, P);
            else if (act == NEAT_ACT_RELU)
                k_dec_mfma_bwd<64, NEAT_ACT_RELU><<<wg, 256, 0, s>>>(x, S, params, last_act, beta, gy, gx, partials, P);
            else
                k_dec_mfma_bwd<64, NEAT_ACT_SILU><<<wg, 256, 0, s>>>(x, S, params, last_act, beta, gy, gx, partials, P);
            NEAT_LAUNCHED();
            k_dec_reduce_partials<<<reduce_partials_blocks(P), 256, 0, s>>>(partials, wg * kDecMfmaWaves, P, gparams);
            NEAT_LAUNCHED();
            return NEAT_OK;
        }
        const int nwg   = bwd_workgroups(S);
        k_transpose_params<<<grid_for(P, 256), 256, 0, s>>>(params, m, wt);
        NEAT_LAUNCHED();
        NEAT_HIP(hipMemsetAsync(partials, 0, (size_t)nwg * P * sizeof(float), s));
        if (hidden == 16)
            k_dec_bwd<16><<<nwg, kTile, 0, s>>>(x, S, m, wt, act, last_act, beta, gy, gx, partials);
        else if (hidden == 32)
            k_dec_bwd<32><<<nwg, kTile, 0, s>>>(x, S, m, wt, act, last_act, beta, gy, gx, partials);
        else
            k_dec_bwd<64><<<nwg, kTile, 0, s>>>(x, S, m, wt, act, last_act, beta, gy, gx, partials);
        NEAT_LAUNCHED();
        k_dec_reduce_partials<<<reduce_partials_blocks(P), 256, 0, s>>>(partials, nwg, P, gparams);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }
}
