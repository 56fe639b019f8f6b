// Learned proposal for in-node sampling (no counterpart in the reference): a per-segment MLP 6 -> 32 -> 32 -> 8 on the
// matrix cores in BF16 (FP32 accumulation) that turns a segment's entry and exit points into a piecewise-constant
// density over 8 equal bins of [tmin, tmax]; the projector then places the segment's samples in equal-probability
// cells of that density (neat_projector_args::seg_cdf). The network is trained online on the field's own values at
// uniform samples (neat_proposal_targets -> neat_proposal_bwd -> neat_adam_step).
//
//   x      = [o + d tmin, o + d tmax]                            (6 features, padded to 8)
//   logits = W2 relu(W1 relu(W0 x + b0) + b1) + b2               packed: W0 [32,6] b0 [32] W1 [32,32] b1 [32] W2 [8,32] b2 [8]
//   p      = (1 - 0.25) softmax(logits) + 0.25 / 8, cdf = cumsum(p) (cdf[7] = 1)
//   loss   = mean over segments of -sum_b q_b log softmax(logits)_b
//
// Tiles of 32 segments per wave, hidden-on-rows (the accumulator layout of a layer is the B operand of the next; see
// mfma_mlp.hpp). Weight gradients are products over the tile's segments staged [row][segment] in LDS.
#include "mfma_mlp.hpp"
#include "mlp_tile.hpp"

namespace neat
{
namespace
{

constexpr int kPIn = 6, kPH = 32, kPOut = 8;
constexpr int kPW0 = 0, kPB0 = kPW0 + kPH * kPIn, kPW1 = kPB0 + kPH, kPB1 = kPW1 + kPH * kPH, kPW2 = kPB1 + kPH,
              kPB2 = kPW2 + kPOut * kPH, kPParams = kPB2 + kPOut;  // 1544
constexpr float kPMix = 0.25f;
constexpr float kPFloor = 0.05f;

struct PropSegs
{
    const float* ro;
    const float* rd;
    const int32_t* seg_ray;
    const float* tmin;
    const float* tmax;
    int64_t K;
};

struct PropLDS
{
    s16x4 w1[64];        // layer 1: W0[l%32][4g + j] (0 for columns >= 6)
    bf16x8 w2[2][64];    // layer 2: W1[l%32][crow(8st + j, g)]
    bf16x8 w3[2][64];    // layer 3: W2[l%32][crow(8st + j, g)] (0 for rows >= 8)
    bf16x8 w2t[2][64];   // W1^T[l%32][crow(8st + j, g)] = W1[crow(..)][l%32]
    s16x4 w3t[64];       // W2^T[l%32][4g + j] = W2[4g + j][l%32]
    float b0[kPH], b1[kPH], b2[kPOut];
};

__device__ void prop_load_weights(const float* __restrict__ P, PropLDS& w)
{
    const float* W0 = P + kPW0;
    const float* W1 = P + kPW1;
    const float* W2 = P + kPW2;
    for (int l = threadIdx.x; l < 64; l += blockDim.x)
    {
        const int m = l & 31, g = l >> 5;
        s16x4 a, at;
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            const int c = 4 * g + j;
            a[j]        = bf16_bits(c < kPIn ? W0[m * kPIn + c] : 0.f);
            at[j]       = bf16_bits(W2[c * kPH + m]);
        }
        w.w1[l]  = a;
        w.w3t[l] = at;
#pragma unroll
        for (int st = 0; st < 2; ++st)
        {
            bf16x8 f2, f3, f2t;
#pragma unroll
            for (int j = 0; j < 8; ++j)
            {
                const int src = crow(8 * st + j, g);
                f2[j]         = (__bf16)W1[m * kPH + src];
                f3[j]         = (__bf16)(m < kPOut ? W2[m * kPH + src] : 0.f);
                f2t[j]        = (__bf16)W1[src * kPH + m];
            }
            w.w2[st][l]  = f2;
            w.w3[st][l]  = f3;
            w.w2t[st][l] = f2t;
        }
    }
    for (int i = threadIdx.x; i < kPH; i += blockDim.x)
    {
        w.b0[i] = P[kPB0 + i];
        w.b1[i] = P[kPB1 + i];
        if (i < kPOut) w.b2[i] = P[kPB2 + i];
    }
}

// The lane's 4 input features (group g: features 4g..4g+3, zero past the sixth) of segment k.
__device__ __forceinline__ void prop_features(const PropSegs& s, int64_t k, int g, float (&x4)[4])
{
    const int32_t r  = s.seg_ray[k];
    const float t0 = s.tmin[k], t1 = s.tmax[k];
    float f[6];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        const float o = s.ro[3 * (int64_t)r + d], dd = s.rd[3 * (int64_t)r + d];
        f[d]     = o + dd * t0;
        f[3 + d] = o + dd * t1;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) x4[j] = (4 * g + j < kPIn) ? f[4 * g + j] : 0.f;
}

struct PropFwd
{
    f32x16 z1, z2;
    float lg[4];  // logits of bins 4g..4g+3
};

__device__ __forceinline__ void prop_forward(const PropLDS& w, const float (&x4)[4], PropFwd& f)
{
    const int l = threadIdx.x & 63, g = l >> 5;
    s16x4 xb;
#pragma unroll
    for (int j = 0; j < 4; ++j) xb[j] = bf16_bits(x4[j]);
    f32x16 acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(w.w1[l], xb, f32x16{}, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] += w.b0[crow(i, g)];
    f.z1 = acc;
    acc  = f32x16{};
#pragma unroll
    for (int st = 0; st < 2; ++st)
        acc = PrecBF16::mma(w.w2[st][l], PrecBF16::from_acc(f.z1, st, [](float v, int) { return fmaxf(v, 0.f); }),
                            acc);
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] += w.b1[crow(i, g)];
    f.z2 = acc;
    acc  = f32x16{};
#pragma unroll
    for (int st = 0; st < 2; ++st)
        acc = PrecBF16::mma(w.w3[st][l], PrecBF16::from_acc(f.z2, st, [](float v, int) { return fmaxf(v, 0.f); }),
                            acc);
    // rows 0..7 are registers 0..3 of the two lane groups (crow(j, g) = 4g + j)
#pragma unroll
    for (int j = 0; j < 4; ++j) f.lg[j] = acc[j] + w.b2[4 * g + j];
}

// softmax over the 8 logits held 4 + 4 by lanes n and n + 32
__device__ __forceinline__ void prop_softmax(const float (&lg)[4], float (&p)[4])
{
    float m = fmaxf(fmaxf(lg[0], lg[1]), fmaxf(lg[2], lg[3]));
    m       = fmaxf(m, __shfl_xor(m, 32));
    float e[4], s = 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j)
    {
        e[j] = expf(lg[j] - m);
        s += e[j];
    }
    s += __shfl_xor(s, 32);
#pragma unroll
    for (int j = 0; j < 4; ++j) p[j] = e[j] / s;
}

__global__ void __launch_bounds__(256) k_prop_fwd(PropSegs sg, const int32_t* seg_count, const float* __restrict__ P,
                                                  int32_t nprop, float* cdf_out, int32_t* count_out, float* logits_out)
{
    __shared__ PropLDS w;
    prop_load_weights(P, w);
    __syncthreads();
    const int l = threadIdx.x & 63, wave = threadIdx.x >> 6, n = l & 31, g = l >> 5;
    const int64_t nt = (sg.K + 31) / 32;
    for (int64_t tile = (int64_t)blockIdx.x * 4 + wave; tile < nt; tile += (int64_t)gridDim.x * 4)
    {
        const int64_t k  = tile * 32 + n;
        const bool valid = k < sg.K;
        float x4[4]      = {0.f, 0.f, 0.f, 0.f};
        if (valid) prop_features(sg, k, g, x4);
        PropFwd f;
        prop_forward(w, x4, f);
        float p[4];
        prop_softmax(f.lg, p);
        float pm[4], run = 0.f;
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            pm[j] = (1.f - kPMix) * p[j] + kPMix / kPOut;
            run += pm[j];
        }
        // lanes of group 1 start after group 0's four bins
        const float lower = __shfl_xor(run, 32);
        float c[4], acc = g ? lower : 0.f;
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            acc += pm[j];
            c[j] = acc;
        }
        if (g == 1) c[3] = 1.f;
        if (!valid) continue;
        const int32_t cnt = seg_count[k];
        const bool use    = cnt > nprop;
        if (count_out && g == 0) count_out[k] = use ? nprop : cnt;
        *reinterpret_cast<float4*>(cdf_out + 8 * k + 4 * g) =
            use ? make_float4(c[0], c[1], c[2], c[3]) : make_float4(0.f, 0.f, 0.f, 0.f);
        if (logits_out)
            *reinterpret_cast<float4*>(logits_out + 8 * k + 4 * g) = make_float4(f.lg[0], f.lg[1], f.lg[2], f.lg[3]);
    }
}

// Gradient of scale * sum_k CE(q_k, softmax(logits_k)); one packed partial slice per wave.
__global__ void __launch_bounds__(256) k_prop_bwd(PropSegs sg, const float* __restrict__ P, const float* q,
                                                  float scale, float* partials)
{
    __shared__ PropLDS w;
    __shared__ __align__(16) float stage[4][(2 * kPH + kPOut) * kRS];
    prop_load_weights(P, w);
    __syncthreads();
    const int l = threadIdx.x & 63, wave = threadIdx.x >> 6, n = l & 31, g = l >> 5;
    float* sA = stage[wave];
    float* sB = sA + kPH * kRS;
    float* sX = sB + kPH * kRS;  // 8 rows: features 0..5, ones (6: b0), ones (7: b1)
    if (g == 0)
    {
        sX[6 * kRS + n] = 1.f;
        sX[7 * kRS + n] = 1.f;
    }
    f32x16 dW1 = {}, dW2 = {}, E = {};
    float db2[4] = {0.f, 0.f, 0.f, 0.f};
    const int64_t nt = (sg.K + 31) / 32;
    for (int64_t tile = (int64_t)blockIdx.x * 4 + wave; tile < nt; tile += (int64_t)gridDim.x * 4)
    {
        const int64_t k  = tile * 32 + n;
        const bool valid = k < sg.K;
        float x4[4]      = {0.f, 0.f, 0.f, 0.f};
        if (valid) prop_features(sg, k, g, x4);
        PropFwd f;
        prop_forward(w, x4, f);
        float p[4], dl[4];
        prop_softmax(f.lg, p);
        if (valid)
        {
            const float4 qq = *reinterpret_cast<const float4*>(q + 8 * k + 4 * g);
            dl[0] = (p[0] - qq.x) * scale;
            dl[1] = (p[1] - qq.y) * scale;
            dl[2] = (p[2] - qq.z) * scale;
            dl[3] = (p[3] - qq.w) * scale;
        }
        else
        {
            dl[0] = dl[1] = dl[2] = dl[3] = 0.f;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) db2[j] += dl[j];
        // dZ2 = (W2^T dl) * relu'(z2): K = the 8 bins, held 4 + 4 by the two lane groups
        s16x4 dlb;
#pragma unroll
        for (int j = 0; j < 4; ++j) dlb[j] = bf16_bits(dl[j]);
        f32x16 dZ2 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(w.w3t[l], dlb, f32x16{}, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 16; ++i) dZ2[i] = f.z2[i] > 0.f ? dZ2[i] : 0.f;
        f32x16 dZ1 = {};
#pragma unroll
        for (int st = 0; st < 2; ++st)
            dZ1 = PrecBF16::mma(w.w2t[st][l], PrecBF16::from_acc(dZ2, st, [](float v, int) { return v; }), dZ1);
#pragma unroll
        for (int i = 0; i < 16; ++i) dZ1[i] = f.z1[i] > 0.f ? dZ1[i] : 0.f;
        // phase 1: dW1 += dZ2 relu(z1)^T, E(col 7) += dZ2 1^T
#pragma unroll
        for (int i = 0; i < 16; ++i)
        {
            const int r     = crow(i, g);
            sA[r * kRS + n] = dZ2[i];
            sB[r * kRS + n] = fmaxf(f.z1[i], 0.f);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (4 * g + j < kPIn) sX[(4 * g + j) * kRS + n] = x4[j];
        wave_lds_fence();
#pragma unroll
        for (int st = 0; st < 2; ++st)
        {
            const auto a = stage_frag<PrecBF16>(sA, n, st, g, false);
            dW1          = PrecBF16::mma(a, stage_frag<PrecBF16>(sB, n, st, g, false), dW1);
            E            = PrecBF16::mma(a, stage_frag<PrecBF16>(sX, n, st, g, n != 7), E);
        }
        wave_lds_fence();
        // phase 2: E(cols 0..6) += dZ1 [x; 1]^T
#pragma unroll
        for (int i = 0; i < 16; ++i) sA[crow(i, g) * kRS + n] = dZ1[i];
        wave_lds_fence();
#pragma unroll
        for (int st = 0; st < 2; ++st)
            E = PrecBF16::mma(stage_frag<PrecBF16>(sA, n, st, g, false), stage_frag<PrecBF16>(sX, n, st, g, n > 6), E);
        wave_lds_fence();
        // phase 3: dW2 += dl relu(z2)^T (rows 0..7)
#pragma unroll
        for (int j = 0; j < 4; ++j) sA[(4 * g + j) * kRS + n] = dl[j];
#pragma unroll
        for (int i = 0; i < 16; ++i) sB[crow(i, g) * kRS + n] = fmaxf(f.z2[i], 0.f);
        wave_lds_fence();
#pragma unroll
        for (int st = 0; st < 2; ++st)
            dW2 = PrecBF16::mma(stage_frag<PrecBF16>(sA, n, st, g, n >= kPOut), stage_frag<PrecBF16>(sB, n, st, g, false),
                                dW2);
        wave_lds_fence();
    }
    float* part = partials + ((int64_t)blockIdx.x * 4 + wave) * kPParams;
#pragma unroll
    for (int i = 0; i < 16; ++i)
    {
        const int r = crow(i, g);
        part[kPW1 + r * kPH + n] = dW1[i];
        if (n < kPIn) part[kPW0 + r * kPIn + n] = E[i];
        if (n == 6) part[kPB0 + r] = E[i];
        if (n == 7) part[kPB1 + r] = E[i];
        if (r < kPOut) part[kPW2 + r * kPH + n] = dW2[i];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
    {
        float v = db2[j];
#pragma unroll
        for (int off = 16; off >= 1; off >>= 1) v += __shfl_xor(v, off);
        if (n == 0) part[kPB2 + 4 * g + j] = v;
    }
}

// q[k] from the field's values at segment k's uniform samples (values[off[k] .. off[k] + count[k])).
__global__ void __launch_bounds__(256) k_prop_targets(const float* values, const int64_t* off, const int32_t* count,
                                                      int64_t K, float* q)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int32_t n = count[k];
    const float* f  = values + off[k];
    float tv[kPOut] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float prev      = n > 0 ? f[0] : 0.f;
    for (int32_t j = 0; j + 1 < n; ++j)
    {
        const float cur = f[j + 1];
        const int b     = min(kPOut - 1, (int)(((int64_t)(j + 1) * kPOut) / n));
        const float dv  = fabsf(cur - prev);
#pragma unroll
        for (int i = 0; i < kPOut; ++i)
            if (i == b) tv[i] += dv;
        prev = cur;
    }
    float tot = 0.f;
#pragma unroll
    for (int i = 0; i < kPOut; ++i) tot += tv[i];
    const float add = kPFloor * tot / kPOut + 1e-12f;
    float s         = 0.f;
#pragma unroll
    for (int i = 0; i < kPOut; ++i)
    {
        tv[i] += add;
        s += tv[i];
    }
    float4* out = reinterpret_cast<float4*>(q + 8 * k);
    out[0]      = make_float4(tv[0] / s, tv[1] / s, tv[2] / s, tv[3] / s);
    out[1]      = make_float4(tv[4] / s, tv[5] / s, tv[6] / s, tv[7] / s);
}

// Per-slot roughness: decay, then the largest total variation of the field along any observed segment of the slot.
__global__ void __launch_bounds__(256) k_rough_decay(float* rough, int64_t n, float decay)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rough[i] *= decay;
}

__global__ void __launch_bounds__(256) k_rough_update(const float* values, const int64_t* off, const int32_t* count,
                                                      const int32_t* seg_node, const int64_t* prefix, int32_t num_nodes,
                                                      int64_t num_slots, int64_t K, float* rough)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int32_t node = seg_node[k];
    if (node < 0 || node >= num_nodes) return;
    const int64_t slot = prefix[node];
    if (slot < 0 || slot >= num_slots) return;
    const float* f = values + off[k];
    float tv       = 0.f;
    for (int32_t j = 0; j + 1 < count[k]; ++j) tv += fabsf(f[j + 1] - f[j]);
    atomicMax(reinterpret_cast<unsigned int*>(rough) + slot, __float_as_uint(tv));  // non-negative: bits order
}

// Fallback to the uniform samples for segments whose slot is rougher than tau.
// Three tiers by the slot's roughness r: r <= tau keeps what neat_proposal_fwd set; tau < r <= tau_uniform gives a
// segment that neat_proposal_fwd gave proposal samples min(uniform count, mid_samples) of them (so every long
// segment of such a slot samples its proposal, none switches back to the uniform rule); r > tau_uniform, or a segment
// outside every slot, gets its uniform samples back.
__global__ void __launch_bounds__(256) k_prop_gate(const int32_t* seg_node, const int64_t* prefix, int32_t num_nodes,
                                                   int64_t num_slots, const float* rough, float tau, float tau_uniform,
                                                   int32_t mid_samples, const int32_t* count_in, int64_t K, float* cdf,
                                                   int32_t* count_out)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int32_t node = seg_node[k];
    const int64_t slot = (node >= 0 && node < num_nodes) ? prefix[node] : -1;
    const float r      = (slot < 0 || slot >= num_slots) ? INFINITY : rough[slot];
    if (r <= tau) return;
    const int32_t n = count_in[k];
    if (r <= tau_uniform)
    {
        if (n > count_out[k]) count_out[k] = min(n, mid_samples);
        return;
    }
    float4* row = reinterpret_cast<float4*>(cdf + 8 * k);
    row[0] = make_float4(0.f, 0.f, 0.f, 0.f);
    row[1] = make_float4(0.f, 0.f, 0.f, 0.f);
    count_out[k] = n;
}

int prop_bwd_workgroups(int64_t K) { return (int)std::max<int64_t>(1, std::min<int64_t>((K + 127) / 128, 256)); }

int check_segs(const PropSegs& s, const float* params)
{
    NEAT_REQUIRE(s.K >= 0 && s.K < ((int64_t)1 << 40), "proposal: bad segment count");
    NEAT_REQUIRE(params, "proposal: null parameters");
    NEAT_REQUIRE(s.K == 0 || (s.ro && s.rd && s.seg_ray && s.tmin && s.tmax), "proposal: null segment arrays");
    return NEAT_OK;
}

}  // namespace
}  // namespace neat

using namespace neat;

extern "C"
{
    int64_t neat_proposal_num_params(void) { return kPParams; }

    int neat_proposal_fwd(neat_stream_t stream, const float* ray_origin, const float* ray_direction,
                          const int32_t* seg_ray, const float* seg_tmin, const float* seg_tmax,
                          const int32_t* seg_count, int64_t num_segments, const float* params, int32_t num_samples,
                          float* seg_cdf, int32_t* seg_count_out, float* logits)
    {
        PropSegs s{ray_origin, ray_direction, seg_ray, seg_tmin, seg_tmax, num_segments};
        int st = check_segs(s, params);
        if (st) return st;
        NEAT_REQUIRE(num_samples >= 1, "neat_proposal_fwd: num_samples must be >= 1");
        NEAT_REQUIRE(num_segments == 0 || (seg_count && seg_cdf), "neat_proposal_fwd: null seg_count / seg_cdf");
        NEAT_REQUIRE((reinterpret_cast<uintptr_t>(seg_cdf) | reinterpret_cast<uintptr_t>(logits)) % 16 == 0,
                     "neat_proposal_fwd: seg_cdf / logits must be 16-byte aligned");
        if (num_segments == 0) return NEAT_OK;
        const int64_t tiles = (num_segments + 31) / 32;
        const unsigned grid = (unsigned)std::min<int64_t>((tiles + 3) / 4, 2048);
        k_prop_fwd<<<grid, 256, 0, as_stream(stream)>>>(s, seg_count, params, num_samples, seg_cdf, seg_count_out,
                                                        logits);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    size_t neat_proposal_workspace_bytes(int64_t num_segments)
    {
        return (size_t)prop_bwd_workgroups(num_segments) * 4 * kPParams * sizeof(float);
    }

    int neat_proposal_bwd(neat_stream_t stream, const float* ray_origin, const float* ray_direction,
                          const int32_t* seg_ray, const float* seg_tmin, const float* seg_tmax, int64_t num_segments,
                          const float* params, const float* target, float scale, float* grad_params, void* workspace,
                          size_t workspace_bytes)
    {
        PropSegs s{ray_origin, ray_direction, seg_ray, seg_tmin, seg_tmax, num_segments};
        int st = check_segs(s, params);
        if (st) return st;
        NEAT_REQUIRE(grad_params && (num_segments == 0 || target), "neat_proposal_bwd: null target / gradient");
        NEAT_REQUIRE(reinterpret_cast<uintptr_t>(target) % 16 == 0, "neat_proposal_bwd: target must be 16-byte aligned");
        const size_t need = neat_proposal_workspace_bytes(num_segments);
        NEAT_REQUIRE(workspace && workspace_bytes >= need, "neat_proposal_bwd: workspace too small (%zu < %zu)",
                     workspace_bytes, need);
        hipStream_t hs = as_stream(stream);
        const int nwg  = prop_bwd_workgroups(num_segments);
        float* part    = static_cast<float*>(workspace);
        k_prop_bwd<<<nwg, 256, 0, hs>>>(s, params, target, scale, part);
        NEAT_LAUNCHED();
        k_dec_reduce_partials<<<reduce_partials_blocks(kPParams), 256, 0, hs>>>(part, nwg * 4, kPParams, grad_params);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_proposal_roughness_update(neat_stream_t stream, const float* values, const int64_t* seg_value_offset,
                                       const int32_t* seg_count, const int32_t* seg_node, int64_t num_segments,
                                       const int64_t* node_active_prefix_sum, int32_t num_nodes, int64_t num_slots,
                                       float decay, float* slot_roughness)
    {
        NEAT_REQUIRE(num_segments >= 0 && num_slots >= 0 && num_nodes >= 0 && decay >= 0.f && decay <= 1.f,
                     "neat_proposal_roughness_update: bad arguments");
        NEAT_REQUIRE(slot_roughness || num_slots == 0, "neat_proposal_roughness_update: null table");
        hipStream_t s = as_stream(stream);
        if (num_slots > 0)
        {
            k_rough_decay<<<grid_for(num_slots, 256), 256, 0, s>>>(slot_roughness, num_slots, decay);
            NEAT_LAUNCHED();
        }
        if (num_segments == 0) return NEAT_OK;
        NEAT_REQUIRE(values && seg_value_offset && seg_count && seg_node && node_active_prefix_sum,
                     "neat_proposal_roughness_update: null pointer");
        k_rough_update<<<grid_for(num_segments, 256), 256, 0, s>>>(values, seg_value_offset, seg_count, seg_node,
                                                                   node_active_prefix_sum, num_nodes, num_slots,
                                                                   num_segments, slot_roughness);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_proposal_gate(neat_stream_t stream, const int32_t* seg_node, const int32_t* seg_count,
                           int64_t num_segments, const int64_t* node_active_prefix_sum, int32_t num_nodes,
                           int64_t num_slots, const float* slot_roughness, float tau, float tau_uniform,
                           int32_t mid_samples, float* seg_cdf, int32_t* seg_count_out)
    {
        NEAT_REQUIRE(num_segments >= 0, "neat_proposal_gate: bad segment count");
        NEAT_REQUIRE(mid_samples >= 1, "neat_proposal_gate: mid_samples %d must be >= 1", mid_samples);
        if (num_segments == 0) return NEAT_OK;
        NEAT_REQUIRE(seg_node && seg_count && node_active_prefix_sum && slot_roughness && seg_cdf && seg_count_out,
                     "neat_proposal_gate: null pointer");
        NEAT_REQUIRE(reinterpret_cast<uintptr_t>(seg_cdf) % 16 == 0, "neat_proposal_gate: seg_cdf must be 16-byte aligned");
        k_prop_gate<<<grid_for(num_segments, 256), 256, 0, as_stream(stream)>>>(
            seg_node, node_active_prefix_sum, num_nodes, num_slots, slot_roughness, tau, tau_uniform, mid_samples,
            seg_count, num_segments, seg_cdf, seg_count_out);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_proposal_targets(neat_stream_t stream, const float* values, const int64_t* seg_value_offset,
                              const int32_t* seg_count, int64_t num_segments, float* target)
    {
        NEAT_REQUIRE(num_segments >= 0, "neat_proposal_targets: bad segment count");
        if (num_segments == 0) return NEAT_OK;
        NEAT_REQUIRE(values && seg_value_offset && seg_count && target, "neat_proposal_targets: null pointer");
        NEAT_REQUIRE(reinterpret_cast<uintptr_t>(target) % 16 == 0, "neat_proposal_targets: target must be 16-byte aligned");
        k_prop_targets<<<grid_for(num_segments, 256), 256, 0, as_stream(stream)>>>(values, seg_value_offset, seg_count,
                                                                                   num_segments, target);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }
}
