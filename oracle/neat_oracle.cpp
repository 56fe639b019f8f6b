/*
 * neat_oracle.cpp -- CPU restatement of NeAT's differentiable forward projector.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity oracle that the HIP path is checked against, and the
 * CPU baseline that bench.py times beside the GPU. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it, and only as the checker. The product package (neat_amd/) never links, loads or
 * calls anything in this directory.
 *
 * Every function restates a piece of the reference (darglein/NeAT) and cites the file:line it follows.
 * The real type T is float for the FP32 restatement and double for the FP64 oracle.
 *
 * Built with -O3 -ffp-contract=off: every a*b+c in this file rounds twice, exactly as the separately written
 * expressions of the reference round when they are not contracted. The reference's own nvcc build contracts
 * by default, so its last-ulp values are not knowable from source; this oracle fixes the uncontracted order.
 *
 * Canonical order (the reference's order is atomic-order dependent, HyperTree.cu:281,301):
 *   rays in input order; within a ray, segments sorted by (tmin, node_id); within a segment, samples by j.
 * Jitter: a counter-based hash of (seed, ray, segment rank, j), see neat_jitter().
 *
 * Pinning status of each function (the checks live in tests/test_oracle.py):
 *   tree_build            pinned: closed-form node counts, dyadic boxes, BFS child numbering.
 *   update_active         pinned: brute force on random masks.
 *   intersect             pinned: exact rational slab test (fractions) on random and special rays.
 *   raysample_*           pinned: exact-rational segment lengths, telescoping weight sums, closed-form counts
 *                         for axis-aligned rays, chord length of a full tiling.
 *   local_samples         pinned: closed form (faces map to +-1 exactly).
 *   grid_sample_fwd/bwd   pinned: torch.nn.functional.grid_sample (border, align_corners) in float64 + autograd.
 *   decoder_fwd/bwd       pinned: torch.nn.Linear / relu / silu / softplus in float64 + autograd.
 *   integrate_fwd         pinned: torch.index_add_.
 *   projector_fwd/bwd     pinned: constant-field and affine-field closed forms, central finite differences.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace
{

// ---------------------------------------------------------------------------------------------------------------
// Jitter: counter-based hash (splitmix64 finaliser) of (seed, ray, segment rank, sample index).
// The reference draws torch::rand_like per global sample index (HyperTree.cu:403,356), which depends on the
// atomic order of its output; this replaces it with an order-free definition shared with the HIP kernels.
// ---------------------------------------------------------------------------------------------------------------
inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline float neat_jitter(uint64_t seed, int64_t ray, int32_t seg_rank, int32_t j)
{
    uint64_t h = mix64(seed ^ ((uint64_t)ray * 0x9E3779B97F4A7C15ull));
    h          = mix64(h ^ (((uint64_t)(uint32_t)seg_rank << 32) | (uint64_t)(uint32_t)j));
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}

// ---------------------------------------------------------------------------------------------------------------
// IntersectBoxRayPrecise (HyperTree.h:41-83): float slab distances, double running t_near/t_far with the
// reference's sentinels, early exit per axis, result narrowed to float.
// ---------------------------------------------------------------------------------------------------------------
struct HitF
{
    bool hit;
    float tmin, tmax;
};

inline HitF intersect_f(const float* pos_min, const float* pos_max, const float* origin, const float* direction)
{
    double t_near = -236436436;
    double t_far  = 43637575;
    for (int i = 0; i < 3; i++)
    {
        if (direction[i] == 0)
        {
            if ((origin[i] < pos_min[i]) || (origin[i] >= pos_max[i])) return {false, 0, 0};
        }
        else
        {
            float T_1 = (pos_min[i] - origin[i]) / direction[i];
            float T_2 = (pos_max[i] - origin[i]) / direction[i];
            if (T_1 > T_2)
            {
                float tmp = T_1;
                T_1       = T_2;
                T_2       = tmp;
            }
            if (T_1 > t_near) t_near = T_1;
            if (T_2 < t_far) t_far = T_2;
            if ((t_near > t_far) || (t_far < 0)) return {false, 0, 0};
        }
    }
    return {true, (float)t_near, (float)t_far};
}

// The same slab test carried out entirely in double (FP64 oracle re-derivation of tmin/tmax).
struct HitD
{
    bool hit;
    double tmin, tmax;
};

inline HitD intersect_d(const float* pos_min, const float* pos_max, const double* origin, const double* direction)
{
    double t_near = -236436436;
    double t_far  = 43637575;
    for (int i = 0; i < 3; i++)
    {
        if (direction[i] == 0)
        {
            if ((origin[i] < pos_min[i]) || (origin[i] >= pos_max[i])) return {false, 0, 0};
        }
        else
        {
            double T_1 = ((double)pos_min[i] - origin[i]) / direction[i];
            double T_2 = ((double)pos_max[i] - origin[i]) / direction[i];
            if (T_1 > T_2) std::swap(T_1, T_2);
            if (T_1 > t_near) t_near = T_1;
            if (T_2 < t_far) t_far = T_2;
            if ((t_near > t_far) || (t_far < 0)) return {false, 0, 0};
        }
    }
    return {true, t_near, t_far};
}

struct Seg
{
    float tmin, tmax;
    int64_t node;
    int32_t count;
};

// Segment list of one ray: ComputeRaySamples' per-leaf loop (HyperTree.cu:260-295) over every active node,
// sliver filter tmax - tmin < 1e-5 compared in double (:274), sample count iCeil(|tmax-tmin| / diag * max_spn)
// with diag the Eigen norm of the box extent (:287-291; Eigen's fixed-size-3 reduction adds x + (y + z)),
// then the canonical order (tmin, node_id).
void ray_segments(const float* pos_min, const float* pos_max, const int64_t* active_ids, int64_t num_active,
                  const float* o, const float* d, int32_t max_spn, std::vector<Seg>& segs)
{
    segs.clear();
    for (int64_t i = 0; i < num_active; ++i)
    {
        int64_t node   = active_ids[i];
        const float* mn = pos_min + 3 * node;
        const float* mx = pos_max + 3 * node;
        HitF h         = intersect_f(mn, mx, o, d);
        if ((double)(h.tmax - h.tmin) < 1e-5) continue;
        if (!h.hit) continue;
        float ex   = mx[0] - mn[0];
        float ey   = mx[1] - mn[1];
        float ez   = mx[2] - mn[2];
        float diag = std::sqrt(ex * ex + (ey * ey + ez * ez));
        float dis  = h.tmax - h.tmin;
        float rel  = dis / diag;
        int32_t n  = (int32_t)std::ceil(rel * (float)max_spn);
        segs.push_back({h.tmin, h.tmax, node, n});
    }
    std::sort(segs.begin(), segs.end(), [](const Seg& a, const Seg& b)
              { return a.tmin < b.tmin || (a.tmin == b.tmin && a.node < b.node); });
}

// Sample position along the segment (HyperTree.cu:339-358): step = (tmax - tmin) / n,
// t1 = tmin + j*step, t2 = tmin + (j+1)*step, t = t1*(1-a) + t2*a.
template <typename T>
inline T sample_t(T tmin, T step, int32_t j, T a)
{
    T t1 = tmin + (T)j * step;
    T t2 = tmin + (T)(j + 1) * step;
    return t1 * ((T)1 - a) + t2 * a;
}

template <typename T>
inline T seg_a(int jitter, uint64_t seed, int64_t ray, int32_t rank, int32_t j)
{
    return jitter ? (T)neat_jitter(seed, ray, rank, j) : (T)0.5;
}

// Midpoint-cell weight (HyperTree.cu:372-389): w_j = t_half_max - t_half_min, clipped to [tmin, tmax].
template <typename T>
inline T sample_w(T tmin, T tmax, T step, int32_t n, int32_t j, int jitter, uint64_t seed, int64_t ray,
                  int32_t rank)
{
    T t  = sample_t<T>(tmin, step, j, seg_a<T>(jitter, seed, ray, rank, j));
    T lo = (j == 0) ? tmin : (sample_t<T>(tmin, step, j - 1, seg_a<T>(jitter, seed, ray, rank, j - 1)) + t) * (T)0.5;
    T hi = (j == n - 1) ? tmax
                        : (sample_t<T>(tmin, step, j + 1, seg_a<T>(jitter, seed, ray, rank, j + 1)) + t) * (T)0.5;
    return hi - lo;
}

// ---------------------------------------------------------------------------------------------------------------
// Trilinear sample, align_corners, border clamp (IndirectGridSample3D.cu:27-191 forward, :194-429 backward).
// Grid addressed through element strides (N, C, Z, Y, X) like the reference's read_grid lambdas.
// ---------------------------------------------------------------------------------------------------------------
struct GridDesc
{
    int64_t N;
    int32_t C, ID, IH, IW;
    int64_t s[5];
};

inline int clip(int v, int lim) { return std::min(lim - 1, std::max(v, 0)); }

template <typename T>
struct Corners
{
    int idx[8][3];  // clamped (x, y, z) per corner in the order tnw tne tsw tse bnw bne bsw bse
    T w[8];
    T ix, iy, iz;
    int x0, y0, z0;  // unclamped floor
};

template <typename T>
inline Corners<T> corners(const GridDesc& g, T u, T v, T wz)
{
    Corners<T> c;
    T ix = ((u + 1) / 2) * (T)(g.IW - 1);
    T iy = ((v + 1) / 2) * (T)(g.IH - 1);
    T iz = ((wz + 1) / 2) * (T)(g.ID - 1);
    int x0 = (int)std::floor(ix), y0 = (int)std::floor(iy), z0 = (int)std::floor(iz);
    int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
    // surfaces to each neighbour, products evaluated left to right as written in the reference
    c.w[0] = ((T)x1 - ix) * ((T)y1 - iy) * ((T)z1 - iz);  // tnw
    c.w[1] = (ix - (T)x0) * ((T)y1 - iy) * ((T)z1 - iz);  // tne
    c.w[2] = ((T)x1 - ix) * (iy - (T)y0) * ((T)z1 - iz);  // tsw
    c.w[3] = (ix - (T)x0) * (iy - (T)y0) * ((T)z1 - iz);  // tse
    c.w[4] = ((T)x1 - ix) * ((T)y1 - iy) * (iz - (T)z0);  // bnw
    c.w[5] = (ix - (T)x0) * ((T)y1 - iy) * (iz - (T)z0);  // bne
    c.w[6] = ((T)x1 - ix) * (iy - (T)y0) * (iz - (T)z0);  // bsw
    c.w[7] = (ix - (T)x0) * (iy - (T)y0) * (iz - (T)z0);  // bse
    const int xs[8] = {x0, x1, x0, x1, x0, x1, x0, x1};
    const int ys[8] = {y0, y0, y1, y1, y0, y0, y1, y1};
    const int zs[8] = {z0, z0, z0, z0, z1, z1, z1, z1};
    for (int k = 0; k < 8; ++k)
    {
        c.idx[k][0] = clip(xs[k], g.IW);
        c.idx[k][1] = clip(ys[k], g.IH);
        c.idx[k][2] = clip(zs[k], g.ID);
    }
    c.ix = ix;
    c.iy = iy;
    c.iz = iz;
    c.x0 = x0;
    c.y0 = y0;
    c.z0 = z0;
    return c;
}

template <typename T, typename G>
inline void grid_sample_one(const GridDesc& g, const G* grid, int64_t n, T u, T v, T wz, T* out)
{
    Corners<T> c = corners<T>(g, u, v, wz);
    for (int ch = 0; ch < g.C; ++ch)
    {
        T acc = 0;
        for (int k = 0; k < 8; ++k)
        {
            int64_t off = n * g.s[0] + ch * g.s[1] + c.idx[k][2] * g.s[2] + c.idx[k][1] * g.s[3] + c.idx[k][0] * g.s[4];
            acc += (T)grid[off] * c.w[k];
        }
        out[ch] = acc;
    }
}

// Backward of one sample: grid gradient += w_corner * g_c (IndirectGridSample3D.cu:322-329) and the uv gradient
// with the reference's product order (:383-416).
template <typename T, typename G>
inline void grid_sample_bwd_one(const GridDesc& g, const G* grid, int64_t n, T u, T v, T wz, const T* gout,
                                T* grad_grid, T* guv)
{
    Corners<T> c = corners<T>(g, u, v, wz);
    T ix = c.ix, iy = c.iy, iz = c.iz;
    T x0 = (T)c.x0, y0 = (T)c.y0, z0 = (T)c.z0, x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
    T gix = 0, giy = 0, giz = 0;
    for (int ch = 0; ch < g.C; ++ch)
    {
        T go = gout[ch];
        T val[8];
        for (int k = 0; k < 8; ++k)
        {
            int64_t off = n * g.s[0] + ch * g.s[1] + c.idx[k][2] * g.s[2] + c.idx[k][1] * g.s[3] + c.idx[k][0] * g.s[4];
            if (grad_grid) grad_grid[off] += c.w[k] * go;
            val[k] = (T)grid[off];
        }
        T m1 = -1;
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
    if (guv)
    {
        guv[0] = gix * (T)(g.IW - 1) / 2;
        guv[1] = giy * (T)(g.IH - 1) / 2;
        guv[2] = giz * (T)(g.ID - 1) / 2;
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Decoder (FCBlockImpl, ImplicitNet.h:18-71) + last activation (DecodeFeatures, geometry.cpp:102-163).
// Packed parameter layout: for each Linear layer l in order, W_l [out_l, in_l] row-major then b_l [out_l].
// Linear: y_o = sum_k x_k W_ok + b_o (sum over k first, then the bias, like addmm with beta = 1).
// ---------------------------------------------------------------------------------------------------------------
enum
{
    ACT_ID       = 0,
    ACT_RELU     = 1,
    ACT_SILU     = 2,
    ACT_SOFTPLUS = 3,
    ACT_ABS      = 4
};

template <typename T>
inline T act_f(int a, T x, T beta)
{
    switch (a)
    {
        case ACT_RELU:
            return x > 0 ? x : (T)0;
        case ACT_SILU:
            return x / ((T)1 + std::exp(-x));
        case ACT_SOFTPLUS:
            // torch::softplus(x, beta, threshold = 20): linear where beta*x > 20
            return (x * beta > (T)20) ? x : std::log1p(std::exp(x * beta)) / beta;
        case ACT_ABS:
            return std::fabs(x);
        default:
            return x;
    }
}

// d act / d x evaluated at the pre-activation x (torch's relu/silu/softplus/abs backward formulas).
template <typename T>
inline T act_d(int a, T x, T beta)
{
    switch (a)
    {
        case ACT_RELU:
            return x > 0 ? (T)1 : (T)0;
        case ACT_SILU:
        {
            T s = (T)1 / ((T)1 + std::exp(-x));
            return s * ((T)1 + x * ((T)1 - s));
        }
        case ACT_SOFTPLUS:
        {
            if (x * beta > (T)20) return (T)1;
            T z = std::exp(x * beta);
            return z / (z + (T)1);
        }
        case ACT_ABS:
            return x > 0 ? (T)1 : (x < 0 ? (T)-1 : (T)0);
        default:
            return (T)1;
    }
}

struct MLPDesc
{
    int32_t in, hidden, layers, out;  // layers = number of hidden->hidden Linear layers
    int32_t act, last_act;
    float beta;
    int nlin() const { return layers + 2; }
    int din(int l) const { return l == 0 ? in : hidden; }
    int dout(int l) const { return l == nlin() - 1 ? out : hidden; }
    int64_t wofs(int l) const
    {
        int64_t o = 0;
        for (int k = 0; k < l; ++k) o += (int64_t)dout(k) * din(k) + dout(k);
        return o;
    }
    int64_t bofs(int l) const { return wofs(l) + (int64_t)dout(l) * din(l); }
    int64_t nparams() const { return wofs(nlin()); }
};

// Forward of one sample; keeps every pre-activation (z) and post-activation (h) when asked.
template <typename T, typename P>
inline void mlp_fwd(const MLPDesc& m, const P* params, const T* x, T* y, std::vector<T>* zs = nullptr,
                    std::vector<T>* hs = nullptr)
{
    T bufA[256], bufB[256];
    const T* cur = x;
    T* nxt       = bufA;
    if (hs) hs->assign(x, x + m.in);
    if (zs) zs->clear();
    for (int l = 0; l < m.nlin(); ++l)
    {
        int di = m.din(l), dout = m.dout(l);
        const P* W = params + m.wofs(l);
        const P* b = params + m.bofs(l);
        bool last  = l == m.nlin() - 1;
        for (int o = 0; o < dout; ++o)
        {
            T acc = 0;
            for (int k = 0; k < di; ++k) acc += cur[k] * (T)W[(int64_t)o * di + k];
            T z = acc + (T)b[o];
            if (zs) zs->push_back(z);
            T h = act_f<T>(last ? m.last_act : m.act, z, (T)m.beta);
            if (last)
                y[o] = h;
            else
                nxt[o] = h;
        }
        if (!last)
        {
            if (hs) hs->insert(hs->end(), nxt, nxt + dout);
            cur = nxt;
            nxt = (nxt == bufA) ? bufB : bufA;
        }
    }
}

// Backward of one sample: accumulates the packed parameter gradient, returns dL/dx.
template <typename T, typename P>
inline void mlp_bwd(const MLPDesc& m, const P* params, const T* x, const T* gy, T* gparams, T* gx)
{
    std::vector<T> zs, hs;
    T ytmp[64];
    mlp_fwd<T, P>(m, params, x, ytmp, &zs, &hs);
    // offsets of each layer's z in zs and of each layer's input in hs
    int L = m.nlin();
    std::vector<int64_t> zoff(L + 1, 0), hoff(L + 1, 0);
    for (int l = 0; l < L; ++l)
    {
        zoff[l + 1] = zoff[l] + m.dout(l);
        hoff[l + 1] = hoff[l] + m.din(l);
    }
    T delta[256], nd[256];
    for (int o = 0; o < m.out; ++o) delta[o] = gy[o] * act_d<T>(m.last_act, zs[zoff[L - 1] + o], (T)m.beta);
    for (int l = L - 1; l >= 0; --l)
    {
        int di = m.din(l), dout = m.dout(l);
        const P* W = params + m.wofs(l);
        const T* hin = hs.data() + hoff[l];
        if (gparams)
        {
            T* gW = gparams + m.wofs(l);
            T* gb = gparams + m.bofs(l);
            for (int o = 0; o < dout; ++o)
            {
                for (int k = 0; k < di; ++k) gW[(int64_t)o * di + k] += delta[o] * hin[k];
                gb[o] += delta[o];
            }
        }
        for (int k = 0; k < di; ++k)
        {
            T acc = 0;
            for (int o = 0; o < dout; ++o) acc += (T)W[(int64_t)o * di + k] * delta[o];
            nd[k] = acc;
        }
        if (l > 0)
        {
            for (int k = 0; k < di; ++k) delta[k] = nd[k] * act_d<T>(m.act, zs[zoff[l - 1] + k], (T)m.beta);
        }
        else if (gx)
        {
            for (int k = 0; k < di; ++k) gx[k] = nd[k];
        }
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Fused projector composition over a canonical segment list (CreateSamplesForRays -> ComputeLocalSamples ->
// GlobalNodeIdToLocalActiveId -> IndirectGridSample3D -> DecodeFeatures -> IntegrateSamplesXRay;
// HyperTree.cu:393-438, :177-192; HyperTree.h:183-187; IndirectGridSample3D.cu:27-191; geometry.cpp:102-163,
// 393-424). Positions: o + d*t (the torch recompute of HyperTree.cu:426-432); local u = (g - min)/(max - min)*2 - 1.
// In the double instantiation with rederive_t, tmin/tmax come from the FP64 slab test of the same node
// (integer structure -- node and count -- frozen from the FP32 pass).
// ---------------------------------------------------------------------------------------------------------------
struct SegList
{
    int64_t num_segments;
    const int32_t* seg_node;
    const int32_t* seg_ray;
    const int32_t* seg_count;
    const float* seg_tmin;
    const float* seg_tmax;
    const int64_t* ray_seg_offset;  // to derive the segment rank within the ray
    const float* seg_cdf;           // optional [K][8] proposal bin CDF; a row with cdf[7] == 0 is uniform
};

// ---------------------------------------------------------------------------------------------------------------
// Proposal-guided sample placement (no reference counterpart; the build's own definition). A segment's proposal is
// a piecewise-constant density over 8 equal bins of [tmin, tmax], given as its CDF (cdf[b] = P(bin <= b)). Its n
// samples split [0, 1] into n equal-probability cells: cell j spans Q(j/n) .. Q((j+1)/n) of the segment (Q the
// inverse CDF, the ends pinned to tmin and tmax), the sample sits at t1 (1 - a) + t2 a inside its cell (a = 0.5 or
// the jitter, as for uniform samples) and its weight is the cell's length, so the weights sum to tmax - tmin and a
// uniform CDF gives the uniform cells.
// ---------------------------------------------------------------------------------------------------------------
template <typename T>
inline T inv_cdf8(const float* cdf, T u)
{
    int b = 0;
    T lo  = 0;
    while (b < 7 && (T)cdf[b] <= u)
    {
        lo = (T)cdf[b];
        ++b;
    }
    T pb = (T)cdf[b] - lo;
    T f  = pb > 0 ? (u - lo) / pb : (T)0;
    f    = std::min<T>(std::max<T>(f, (T)0), (T)1);
    return ((T)b + f) * (T)0.125;
}

template <typename T>
inline T proposal_cell(const float* cdf, T tmin, T tmax, int32_t n, int32_t j)
{
    if (j <= 0) return tmin;
    if (j >= n) return tmax;
    return tmin + (tmax - tmin) * inv_cdf8<T>(cdf, (T)j / (T)n);
}

template <typename T>
struct SampleCtx
{
    const float* pos_min;
    const float* pos_max;
    const int64_t* prefix;
    const float* ro;
    const float* rd;
    int jitter;
    uint64_t seed;
    bool rederive_t;
};

template <typename T>
inline void segment_params(const SampleCtx<T>& c, const SegList& s, int64_t k, T& tmin, T& tmax, T& step,
                           int32_t& n, int64_t& ray, int32_t& rank)
{
    ray  = s.seg_ray[k];
    n    = s.seg_count[k];
    rank = (int32_t)(k - s.ray_seg_offset[ray]);
    tmin = (T)s.seg_tmin[k];
    tmax = (T)s.seg_tmax[k];
    if (c.rederive_t)
    {
        double o[3] = {c.ro[3 * ray], c.ro[3 * ray + 1], c.ro[3 * ray + 2]};
        double d[3] = {c.rd[3 * ray], c.rd[3 * ray + 1], c.rd[3 * ray + 2]};
        int64_t node = s.seg_node[k];
        HitD h       = intersect_d(c.pos_min + 3 * node, c.pos_max + 3 * node, o, d);
        tmin         = (T)h.tmin;
        tmax         = (T)h.tmax;
    }
    step = (tmax - tmin) / (T)n;
}

template <typename T>
inline void sample_uv(const SampleCtx<T>& c, int64_t ray, int64_t node, T t, T* uv)
{
    for (int i = 0; i < 3; ++i)
    {
        T g   = (T)c.ro[3 * ray + i] + (T)c.rd[3 * ray + i] * t;
        T mn  = (T)c.pos_min[3 * node + i];
        T mx  = (T)c.pos_max[3 * node + i];
        uv[i] = (g - mn) / (mx - mn) * 2 - 1;
    }
}

// t and weight of sample j: uniform cells (HyperTree.cu:339-389) or, with a proposal CDF, its cells.
template <typename T>
inline void sample_tw(const SampleCtx<T>& c, const float* cdf, T tmin, T tmax, T step, int32_t n, int32_t j,
                      int64_t ray, int32_t rank, T& t, T& w)
{
    const T a = seg_a<T>(c.jitter, c.seed, ray, rank, j);
    if (!cdf)
    {
        t = sample_t<T>(tmin, step, j, a);
        w = sample_w<T>(tmin, tmax, step, n, j, c.jitter, c.seed, ray, rank);
        return;
    }
    const T t1 = proposal_cell<T>(cdf, tmin, tmax, n, j);
    const T t2 = proposal_cell<T>(cdf, tmin, tmax, n, j + 1);
    t          = t1 * ((T)1 - a) + t2 * a;
    w          = t2 - t1;
}

template <typename T, typename G>
void projector_fwd(const SampleCtx<T>& c, const SegList& s, const GridDesc& g, const G* grid, const MLPDesc& m,
                   const float* params, int64_t num_rays, T* proj, int nthreads)
{
    std::fill(proj, proj + num_rays, (T)0);
    // segments of one ray are contiguous and ordered; parallelise over rays.
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
    for (int64_t r = 0; r < num_rays; ++r)
    {
        T acc = 0;
        for (int64_t k = s.ray_seg_offset[r]; k < s.ray_seg_offset[r + 1]; ++k)
        {
            T tmin, tmax, step;
            int32_t n, rank;
            int64_t ray;
            segment_params<T>(c, s, k, tmin, tmax, step, n, ray, rank);
            int64_t node = s.seg_node[k];
            int64_t slot = c.prefix[node];
            const float* cdf = (s.seg_cdf && s.seg_cdf[8 * k + 7] > 0) ? s.seg_cdf + 8 * k : nullptr;
            for (int32_t j = 0; j < n; ++j)
            {
                T t, w;
                sample_tw<T>(c, cdf, tmin, tmax, step, n, j, ray, rank, t, w);
                T uv[3], feat[64], y[4];
                sample_uv<T>(c, ray, node, t, uv);
                grid_sample_one<T, G>(g, grid, slot, uv[0], uv[1], uv[2], feat);
                mlp_fwd<T, float>(m, params, feat, y);
                acc += y[0] * w;
            }
        }
        proj[r] = acc;
    }
}

template <typename T, typename G>
void projector_bwd(const SampleCtx<T>& c, const SegList& s, const GridDesc& g, const G* grid, const MLPDesc& m,
                   const float* params, int64_t num_rays, const T* grad_proj, T* grad_grid, int64_t grid_numel,
                   T* grad_params, int nthreads)
{
    int64_t P = m.nparams();
    std::fill(grad_grid, grad_grid + grid_numel, (T)0);
    std::fill(grad_params, grad_params + P, (T)0);
    int nt = std::max(1, nthreads);
    // node-privatised accumulation: one private grid gradient per thread, reduced in thread order.
    std::vector<std::vector<T>> priv_grid(nt > 1 ? nt : 0), priv_par(nt > 1 ? nt : 0);
#pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        T* gg = grad_grid;
        T* gp = grad_params;
        if (nt > 1)
        {
            priv_grid[tid].assign(grid_numel, (T)0);
            priv_par[tid].assign(P, (T)0);
            gg = priv_grid[tid].data();
            gp = priv_par[tid].data();
        }
#pragma omp for schedule(dynamic, 64)
        for (int64_t r = 0; r < num_rays; ++r)
        {
            T gr = grad_proj[r];
            for (int64_t k = s.ray_seg_offset[r]; k < s.ray_seg_offset[r + 1]; ++k)
            {
                T tmin, tmax, step;
                int32_t n, rank;
                int64_t ray;
                segment_params<T>(c, s, k, tmin, tmax, step, n, ray, rank);
                int64_t node = s.seg_node[k];
                int64_t slot = c.prefix[node];
                const float* cdf = (s.seg_cdf && s.seg_cdf[8 * k + 7] > 0) ? s.seg_cdf + 8 * k : nullptr;
                for (int32_t j = 0; j < n; ++j)
                {
                    T t, w;
                    sample_tw<T>(c, cdf, tmin, tmax, step, n, j, ray, rank, t, w);
                    T uv[3], feat[64], gfeat[64];
                    sample_uv<T>(c, ray, node, t, uv);
                    grid_sample_one<T, G>(g, grid, slot, uv[0], uv[1], uv[2], feat);
                    T gy[4] = {gr * w, 0, 0, 0};
                    mlp_bwd<T, float>(m, params, feat, gy, gp, gfeat);
                    grid_sample_bwd_one<T, G>(g, grid, slot, uv[0], uv[1], uv[2], gfeat, gg, (T*)nullptr);
                }
            }
        }
    }
    if (nt > 1)
    {
        // every entry sums the threads' partials in thread order (parallel over entries; the order per entry is fixed)
#pragma omp parallel for schedule(static) num_threads(nt)
        for (int64_t i = 0; i < grid_numel; ++i)
            for (int t = 0; t < nt; ++t) grad_grid[i] += priv_grid[t][i];
        for (int t = 0; t < nt; ++t)
            for (int64_t i = 0; i < P; ++i) grad_params[i] += priv_par[t][i];
    }
}

// Gradient of L = sum_r grad_proj[r] * proj[r] with respect to the rays' origins and directions, the samples' t and
// weights held fixed: the chain the reference runs when the rays require grad -- the global coordinate o + d t
// recomputed with torch ops (HyperTree.cu:426-432), ComputeLocalSamples' backward (grad * 2) / (max - min)
// (HyperTree.cu:116-199), the grid sample's uv gradient (IndirectGridSample3D.cu:752-790) of the decoder's
// dL/dfeature, and the index_select backward onto the rays (sums over each ray's samples).
template <typename T, typename G>
void projector_ray_grad(const SampleCtx<T>& c, const SegList& s, const GridDesc& g, const G* grid, const MLPDesc& m,
                        const float* params, int64_t num_rays, const T* grad_proj, T* grad_o, T* grad_d, int nthreads)
{
    const int64_t P = m.nparams();
#pragma omp parallel num_threads(std::max(1, nthreads))
    {
        std::vector<T> gpar(P);  // the decoder's parameter gradient, not used here
#pragma omp for schedule(dynamic, 64)
        for (int64_t r = 0; r < num_rays; ++r)
        {
            T go[3] = {0, 0, 0}, gd[3] = {0, 0, 0};
            for (int64_t k = s.ray_seg_offset[r]; k < s.ray_seg_offset[r + 1]; ++k)
            {
                T tmin, tmax, step;
                int32_t n, rank;
                int64_t ray;
                segment_params<T>(c, s, k, tmin, tmax, step, n, ray, rank);
                int64_t node = s.seg_node[k];
                int64_t slot = c.prefix[node];
                const float* cdf = (s.seg_cdf && s.seg_cdf[8 * k + 7] > 0) ? s.seg_cdf + 8 * k : nullptr;
                for (int32_t j = 0; j < n; ++j)
                {
                    T t, w;
                    sample_tw<T>(c, cdf, tmin, tmax, step, n, j, ray, rank, t, w);
                    T uv[3], feat[64], gfeat[64], guv[3];
                    sample_uv<T>(c, ray, node, t, uv);
                    grid_sample_one<T, G>(g, grid, slot, uv[0], uv[1], uv[2], feat);
                    T gy[4] = {grad_proj[r] * w, 0, 0, 0};
                    mlp_bwd<T, float>(m, params, feat, gy, gpar.data(), gfeat);
                    grid_sample_bwd_one<T, G>(g, grid, slot, uv[0], uv[1], uv[2], gfeat, (T*)nullptr, guv);
                    for (int i = 0; i < 3; ++i)
                    {
                        const T size = (T)c.pos_max[3 * node + i] - (T)c.pos_min[3 * node + i];
                        const T gpos = (guv[i] * 2) / size;
                        go[i] += gpos;
                        gd[i] += gpos * t;
                    }
                }
            }
            for (int i = 0; i < 3; ++i)
            {
                grad_o[3 * r + i] = go[i];
                grad_d[3 * r + i] = gd[i];
            }
        }
    }
}

GridDesc make_grid(int64_t N, int32_t C, int32_t ID, int32_t IH, int32_t IW, const int64_t* strides)
{
    GridDesc g;
    g.N  = N;
    g.C  = C;
    g.ID = ID;
    g.IH = IH;
    g.IW = IW;
    for (int i = 0; i < 5; ++i) g.s[i] = strides[i];
    return g;
}

}  // namespace

extern "C"
{
    // ---- tree (HyperTree.cpp:11-132) --------------------------------------------------------------------------
    int64_t oracle_tree_num_nodes(int32_t max_depth)
    {
        int64_t num_nodes = 0;
        for (int i = 0; i <= max_depth; ++i)
        {
            int64_t n = 1;
            for (int j = 0; j < i; ++j) n *= 8;
            num_nodes += n;
        }
        return num_nodes;
    }

    // Breadth-first node table: root = [-1,1]^3, child i of a node takes the upper half on axis k iff bit k of i
    // is set (HyperTree.cpp:94-129). scale = depth/max_depth*2 - 1 in float (:76-77).
    void oracle_tree_build(int32_t max_depth, int32_t* parent, int32_t* children, float* pos_min, float* pos_max,
                           float* scale, int32_t* depth_out)
    {
        int64_t num_nodes = oracle_tree_num_nodes(max_depth);
        for (int64_t i = 0; i < num_nodes * 8; ++i) children[i] = -1;
        auto set_node = [&](int64_t id, int32_t par, const float* mn, const float* mx, int32_t depth)
        {
            float s = ((float)depth / (max_depth)) * 2 - 1;
            if (max_depth == 0) s = 0;
            for (int k = 0; k < 3; ++k)
            {
                pos_min[3 * id + k] = mn[k];
                pos_max[3 * id + k] = mx[k];
            }
            depth_out[id] = depth;
            parent[id]    = par;
            scale[id]     = s;
        };
        int64_t current = 0, start_layer = 0;
        float rmin[3] = {-1, -1, -1}, rmax[3] = {1, 1, 1};
        set_node(current++, -1, rmin, rmax, 0);
        for (int32_t depth = 1; depth <= max_depth; ++depth)
        {
            int64_t n = current - start_layer;
            for (int64_t j = 0; j < n; ++j)
            {
                int64_t node = start_layer + j;
                float mn[3], mx[3], center[3];
                for (int k = 0; k < 3; ++k)
                {
                    mn[k]     = pos_min[3 * node + k];
                    mx[k]     = pos_max[3 * node + k];
                    center[k] = (mn[k] + mx[k]) * 0.5f;
                }
                for (int i = 0; i < 8; ++i)
                {
                    float cmn[3], cmx[3];
                    for (int k = 0; k < 3; ++k)
                    {
                        if (((i >> k) & 1) == 0)
                        {
                            cmn[k] = mn[k];
                            cmx[k] = center[k];
                        }
                        else
                        {
                            cmn[k] = center[k];
                            cmx[k] = mx[k];
                        }
                    }
                    children[node * 8 + i] = (int32_t)current;
                    set_node(current++, (int32_t)node, cmn, cmx, depth);
                }
            }
            start_layer = current - n * 8;
        }
    }

    // ---- active set (HyperTree.cpp:169-206) --------------------------------------------------------------------
    int64_t oracle_update_active(const int32_t* node_active, int64_t num_nodes, int64_t* active_ids, int64_t* prefix)
    {
        int64_t count = 0;
        for (int64_t i = 0; i < num_nodes; ++i)
        {
            if (node_active[i] == 1)
            {
                active_ids[count] = i;
                prefix[i]         = count;
                count++;
            }
            else
            {
                prefix[i] = -1;
            }
        }
        return count;
    }

    // ---- IntersectBoxRayPrecise (HyperTree.h:41-83) --------------------------------------------------------------
    int32_t oracle_intersect(const float* bmin, const float* bmax, const float* o, const float* d, float* out_tminmax)
    {
        HitF h          = intersect_f(bmin, bmax, o, d);
        out_tminmax[0] = h.tmin;
        out_tminmax[1] = h.tmax;
        return h.hit ? 1 : 0;
    }

    float oracle_jitter(uint64_t seed, int64_t ray, int32_t seg_rank, int32_t j)
    {
        return neat_jitter(seed, ray, seg_rank, j);
    }

    // ---- ray sampling (ComputeRaySamples HyperTree.cu:206-391 + CreateSamplesForRays :393-438) ----------------
    // Per-ray segment and sample counts; returns the total sample count.
    int64_t oracle_raysample_count(const float* pos_min, const float* pos_max, const int64_t* active_ids,
                                   int64_t num_active, const float* ray_o, const float* ray_d, int64_t num_rays,
                                   int32_t max_spn, int64_t* ray_nseg, int64_t* ray_nsamp, int32_t nthreads)
    {
        int64_t total = 0;
#pragma omp parallel num_threads(std::max(1, nthreads)) reduction(+ : total)
        {
            std::vector<Seg> segs;
#pragma omp for schedule(dynamic, 256)
            for (int64_t r = 0; r < num_rays; ++r)
            {
                ray_segments(pos_min, pos_max, active_ids, num_active, ray_o + 3 * r, ray_d + 3 * r, max_spn, segs);
                int64_t ns = 0;
                for (auto& s : segs) ns += s.count;
                ray_nseg[r]  = (int64_t)segs.size();
                ray_nsamp[r] = ns;
                total += ns;
            }
        }
        return total;
    }

    // Canonical segment list; ray_seg_offset / ray_sample_offset are the exclusive prefix sums of the counts.
    void oracle_raysample_segments(const float* pos_min, const float* pos_max, const int64_t* active_ids,
                                   int64_t num_active, const float* ray_o, const float* ray_d, int64_t num_rays,
                                   int32_t max_spn, const int64_t* ray_seg_offset, const int64_t* ray_sample_offset,
                                   int32_t* seg_node, int32_t* seg_ray, int32_t* seg_count, float* seg_tmin,
                                   float* seg_tmax, int64_t* seg_sample_offset, int32_t nthreads)
    {
#pragma omp parallel num_threads(std::max(1, nthreads))
        {
            std::vector<Seg> segs;
#pragma omp for schedule(dynamic, 256)
            for (int64_t r = 0; r < num_rays; ++r)
            {
                ray_segments(pos_min, pos_max, active_ids, num_active, ray_o + 3 * r, ray_d + 3 * r, max_spn, segs);
                int64_t k = ray_seg_offset[r], so = ray_sample_offset[r];
                for (auto& s : segs)
                {
                    seg_node[k]          = (int32_t)s.node;
                    seg_ray[k]           = (int32_t)r;
                    seg_count[k]         = s.count;
                    seg_tmin[k]          = s.tmin;
                    seg_tmax[k]          = s.tmax;
                    seg_sample_offset[k] = so;
                    so += s.count;
                    k++;
                }
            }
        }
    }

    // SampleList columns in canonical order (HyperTree.cu:347-389, positions as in :426-432).
    void oracle_raysample_emit(const float* ray_o, const float* ray_d, int64_t num_segments, const int32_t* seg_node,
                               const int32_t* seg_ray, const int32_t* seg_count, const float* seg_tmin,
                               const float* seg_tmax, const int64_t* seg_sample_offset, const int64_t* ray_seg_offset,
                               const int64_t* ray_sample_offset, int32_t jitter, uint64_t seed, float* out_t,
                               float* out_w, float* out_coord, int64_t* out_ray, int64_t* out_local, int64_t* out_node)
    {
        for (int64_t k = 0; k < num_segments; ++k)
        {
            int64_t ray  = seg_ray[k];
            int32_t n    = seg_count[k];
            int32_t rank = (int32_t)(k - ray_seg_offset[ray]);
            float tmin = seg_tmin[k], tmax = seg_tmax[k];
            float step = (tmax - tmin) / (float)n;
            for (int32_t j = 0; j < n; ++j)
            {
                int64_t i = seg_sample_offset[k] + j;
                float t   = sample_t<float>(tmin, step, j, seg_a<float>(jitter, seed, ray, rank, j));
                out_t[i]  = t;
                out_w[i]  = sample_w<float>(tmin, tmax, step, n, j, jitter, seed, ray, rank);
                for (int d = 0; d < 3; ++d) out_coord[3 * i + d] = ray_o[3 * ray + d] + ray_d[3 * ray + d] * t;
                out_ray[i]   = ray;
                out_local[i] = i - ray_sample_offset[ray];
                out_node[i]  = seg_node[k];
            }
        }
    }

    // ---- ComputeLocalSamples (HyperTree.cu:177-192) ------------------------------------------------------------
    void oracle_local_samples(const float* global, const int64_t* node_id, int64_t num_samples, const float* pos_min,
                              const float* pos_max, float* out)
    {
        for (int64_t i = 0; i < num_samples; ++i)
        {
            int64_t n = node_id[i];
            for (int d = 0; d < 3; ++d)
            {
                float mn = pos_min[3 * n + d], mx = pos_max[3 * n + d];
                float size = mx - mn;
                out[3 * i + d] = (global[3 * i + d] - mn) / size * 2 - 1;
            }
        }
    }

    // ---- IndirectGridSample3D (IndirectGridSample3D.cu:27-429) -------------------------------------------------
    void oracle_grid_sample_fwd_f32(const float* grid, int64_t N, int32_t C, int32_t ID, int32_t IH, int32_t IW,
                                    const int64_t* strides, const int64_t* index, const float* uv, int64_t S,
                                    float* out)
    {
        GridDesc g = make_grid(N, C, ID, IH, IW, strides);
        for (int64_t i = 0; i < S; ++i)
            grid_sample_one<float, float>(g, grid, index[i], uv[3 * i], uv[3 * i + 1], uv[3 * i + 2], out + i * C);
    }

    void oracle_grid_sample_fwd_f64(const double* grid, int64_t N, int32_t C, int32_t ID, int32_t IH, int32_t IW,
                                    const int64_t* strides, const int64_t* index, const double* uv, int64_t S,
                                    double* out)
    {
        GridDesc g = make_grid(N, C, ID, IH, IW, strides);
        for (int64_t i = 0; i < S; ++i)
            grid_sample_one<double, double>(g, grid, index[i], uv[3 * i], uv[3 * i + 1], uv[3 * i + 2], out + i * C);
    }

    // grad_grid must be zeroed by the caller (it is accumulated, like torch::zeros_like in :848).
    void oracle_grid_sample_bwd_f32(const float* grid, int64_t N, int32_t C, int32_t ID, int32_t IH, int32_t IW,
                                    const int64_t* strides, const int64_t* index, const float* uv, int64_t S,
                                    const float* grad_out, float* grad_grid, float* grad_uv)
    {
        GridDesc g = make_grid(N, C, ID, IH, IW, strides);
        for (int64_t i = 0; i < S; ++i)
            grid_sample_bwd_one<float, float>(g, grid, index[i], uv[3 * i], uv[3 * i + 1], uv[3 * i + 2],
                                              grad_out + i * C, grad_grid, grad_uv ? grad_uv + 3 * i : nullptr);
    }

    void oracle_grid_sample_bwd_f64(const double* grid, int64_t N, int32_t C, int32_t ID, int32_t IH, int32_t IW,
                                    const int64_t* strides, const int64_t* index, const double* uv, int64_t S,
                                    const double* grad_out, double* grad_grid, double* grad_uv)
    {
        GridDesc g = make_grid(N, C, ID, IH, IW, strides);
        for (int64_t i = 0; i < S; ++i)
            grid_sample_bwd_one<double, double>(g, grid, index[i], uv[3 * i], uv[3 * i + 1], uv[3 * i + 2],
                                                grad_out + i * C, grad_grid, grad_uv ? grad_uv + 3 * i : nullptr);
    }

    // ---- decoder (ImplicitNet.h:18-71 + geometry.cpp:123-143) ----------------------------------------------------
    int64_t oracle_decoder_num_params(int32_t in, int32_t hidden, int32_t layers, int32_t out)
    {
        MLPDesc m{in, hidden, layers, out, 0, 0, 1.f};
        return m.nparams();
    }

    void oracle_decoder_fwd_f64(const double* x, int64_t S, int32_t in, int32_t hidden, int32_t layers, int32_t out,
                                const double* params, int32_t act, int32_t last_act, float beta, double* y)
    {
        MLPDesc m{in, hidden, layers, out, act, last_act, beta};
        for (int64_t i = 0; i < S; ++i) mlp_fwd<double, double>(m, params, x + i * in, y + i * out);
    }

    void oracle_decoder_fwd_f32(const float* x, int64_t S, int32_t in, int32_t hidden, int32_t layers, int32_t out,
                                const float* params, int32_t act, int32_t last_act, float beta, float* y)
    {
        MLPDesc m{in, hidden, layers, out, act, last_act, beta};
        for (int64_t i = 0; i < S; ++i) mlp_fwd<float, float>(m, params, x + i * in, y + i * out);
    }

    // grad_params is accumulated (caller zeroes it).
    void oracle_decoder_bwd_f64(const double* x, int64_t S, int32_t in, int32_t hidden, int32_t layers, int32_t out,
                                const double* params, int32_t act, int32_t last_act, float beta, const double* gy,
                                double* grad_params, double* grad_x)
    {
        MLPDesc m{in, hidden, layers, out, act, last_act, beta};
        for (int64_t i = 0; i < S; ++i)
            mlp_bwd<double, double>(m, params, x + i * in, gy + i * out, grad_params, grad_x ? grad_x + i * in : nullptr);
    }

    void oracle_decoder_bwd_f32(const float* x, int64_t S, int32_t in, int32_t hidden, int32_t layers, int32_t out,
                                const float* params, int32_t act, int32_t last_act, float beta, const float* gy,
                                float* grad_params, float* grad_x)
    {
        MLPDesc m{in, hidden, layers, out, act, last_act, beta};
        for (int64_t i = 0; i < S; ++i)
            mlp_bwd<float, float>(m, params, x + i * in, gy + i * out, grad_params, grad_x ? grad_x + i * in : nullptr);
    }

    // ---- IntegrateSamplesXRay (geometry.cpp:393-424): out[c, ray[i]] += v[i, c] * w[i] in sample order --------
    void oracle_integrate_fwd(const float* values, const float* weight, const int64_t* ray_index, int64_t S,
                              int32_t C, int64_t R, float* out)
    {
        std::fill(out, out + (int64_t)C * R, 0.f);
        for (int64_t i = 0; i < S; ++i)
            for (int c = 0; c < C; ++c) out[(int64_t)c * R + ray_index[i]] += values[i * C + c] * weight[i];
    }

    // ---- fused projector over a canonical segment list -----------------------------------------------------------
    // prec: 0 = float restatement (FP32), 1 = double with tmin/tmax rederived in FP64 (integer structure frozen).
    void oracle_projector_fwd(int32_t prec, const float* pos_min, const float* pos_max, const int64_t* prefix,
                              const float* ray_o, const float* ray_d, int64_t num_rays, const int64_t* ray_seg_offset,
                              int64_t num_segments, const int32_t* seg_node, const int32_t* seg_ray,
                              const int32_t* seg_count, const float* seg_tmin, const float* seg_tmax,
                              const float* seg_cdf, int32_t jitter,
                              uint64_t seed, const float* grid, int64_t N, int32_t C, int32_t ID, int32_t IH,
                              int32_t IW, const int64_t* strides, int32_t in, int32_t hidden, int32_t layers,
                              int32_t out, const float* params, int32_t act, int32_t last_act, float beta,
                              double* proj, int32_t nthreads)
    {
        SegList s{num_segments, seg_node, seg_ray, seg_count, seg_tmin, seg_tmax, ray_seg_offset, seg_cdf};
        GridDesc g = make_grid(N, C, ID, IH, IW, strides);
        MLPDesc m{in, hidden, layers, out, act, last_act, beta};
        if (prec == 0)
        {
            SampleCtx<float> c{pos_min, pos_max, prefix, ray_o, ray_d, jitter, seed, false};
            std::vector<float> p(num_rays);
            projector_fwd<float, float>(c, s, g, grid, m, params, num_rays, p.data(), nthreads);
            for (int64_t r = 0; r < num_rays; ++r) proj[r] = p[r];
        }
        else
        {
            // prec 1: tmin / tmax re-derived by the FP64 slab test; prec 2: the segments' FP32 t values as given
            SampleCtx<double> c{pos_min, pos_max, prefix, ray_o, ray_d, jitter, seed, prec == 1};
            projector_fwd<double, float>(c, s, g, grid, m, params, num_rays, proj, nthreads);
        }
    }

    // Gradients of L = sum_r grad_proj[r] * proj[r] with respect to the grid (same strides as the grid) and the
    // packed decoder parameters.
    void oracle_projector_bwd(int32_t prec, const float* pos_min, const float* pos_max, const int64_t* prefix,
                              const float* ray_o, const float* ray_d, int64_t num_rays, const int64_t* ray_seg_offset,
                              int64_t num_segments, const int32_t* seg_node, const int32_t* seg_ray,
                              const int32_t* seg_count, const float* seg_tmin, const float* seg_tmax,
                              const float* seg_cdf, int32_t jitter,
                              uint64_t seed, const float* grid, int64_t N, int32_t C, int32_t ID, int32_t IH,
                              int32_t IW, const int64_t* strides, int64_t grid_numel, int32_t in, int32_t hidden,
                              int32_t layers, int32_t out, const float* params, int32_t act, int32_t last_act,
                              float beta, const double* grad_proj, double* grad_grid, double* grad_params,
                              int32_t nthreads)
    {
        SegList s{num_segments, seg_node, seg_ray, seg_count, seg_tmin, seg_tmax, ray_seg_offset, seg_cdf};
        GridDesc g = make_grid(N, C, ID, IH, IW, strides);
        MLPDesc m{in, hidden, layers, out, act, last_act, beta};
        int64_t P  = m.nparams();
        if (prec == 0)
        {
            SampleCtx<float> c{pos_min, pos_max, prefix, ray_o, ray_d, jitter, seed, false};
            std::vector<float> gp(num_rays), gg(grid_numel), gpar(P);
            for (int64_t r = 0; r < num_rays; ++r) gp[r] = (float)grad_proj[r];
            projector_bwd<float, float>(c, s, g, grid, m, params, num_rays, gp.data(), gg.data(), grid_numel,
                                        gpar.data(), nthreads);
            for (int64_t i = 0; i < grid_numel; ++i) grad_grid[i] = gg[i];
            for (int64_t i = 0; i < P; ++i) grad_params[i] = gpar[i];
        }
        else
        {
            SampleCtx<double> c{pos_min, pos_max, prefix, ray_o, ray_d, jitter, seed, true};
            projector_bwd<double, float>(c, s, g, grid, m, params, num_rays, grad_proj, grad_grid, grid_numel,
                                         grad_params, nthreads);
        }
    }

    // prec 0: the FP32 restatement; prec 1: FP64 with the samples' t taken from the given segments (the quantities
    // the reference's positional gradient holds fixed)
    void oracle_projector_ray_grad(int32_t prec, const float* pos_min, const float* pos_max, const int64_t* prefix,
                                   const float* ray_o, const float* ray_d, int64_t num_rays,
                                   const int64_t* ray_seg_offset, int64_t num_segments, const int32_t* seg_node,
                                   const int32_t* seg_ray, const int32_t* seg_count, const float* seg_tmin,
                                   const float* seg_tmax, const float* seg_cdf, int32_t jitter, uint64_t seed,
                                   const float* grid, int64_t N, int32_t C, int32_t ID, int32_t IH, int32_t IW,
                                   const int64_t* strides, int32_t in, int32_t hidden, int32_t layers, int32_t out,
                                   const float* params, int32_t act, int32_t last_act, float beta,
                                   const double* grad_proj, double* grad_o, double* grad_d, int32_t nthreads)
    {
        SegList s{num_segments, seg_node, seg_ray, seg_count, seg_tmin, seg_tmax, ray_seg_offset, seg_cdf};
        GridDesc g = make_grid(N, C, ID, IH, IW, strides);
        MLPDesc m{in, hidden, layers, out, act, last_act, beta};
        if (prec == 0)
        {
            SampleCtx<float> c{pos_min, pos_max, prefix, ray_o, ray_d, jitter, seed, false};
            std::vector<float> gp(num_rays), go(3 * num_rays), gd(3 * num_rays);
            for (int64_t r = 0; r < num_rays; ++r) gp[r] = (float)grad_proj[r];
            projector_ray_grad<float, float>(c, s, g, grid, m, params, num_rays, gp.data(), go.data(), gd.data(),
                                             nthreads);
            for (int64_t i = 0; i < 3 * num_rays; ++i)
            {
                grad_o[i] = go[i];
                grad_d[i] = gd[i];
            }
        }
        else
        {
            SampleCtx<double> c{pos_min, pos_max, prefix, ray_o, ray_d, jitter, seed, false};
            projector_ray_grad<double, float>(c, s, g, grid, m, params, num_rays, grad_proj, grad_o, grad_d,
                                              nthreads);
        }
    }

    int32_t oracle_max_threads()
    {
#ifdef _OPENMP
        return omp_get_max_threads();
#else
        return 1;
#endif
    }
}
