// Ray / active-leaf intersection and per-ray sample generation (replaces ComputeRaySamples, HyperTree.cu:206-391,
// and CreateSamplesForRays, HyperTree.cu:393-438).
//
// Instead of testing every active leaf per ray (R x A slab tests) each thread walks the octree top-down and only
// visits children of nodes the ray hits. The pruning is exact: a child box lies inside its parent box, and the
// correctly rounded subtraction and division in the slab test are monotonic, so the child's per-axis interval is a
// sub-interval of the parent's; a parent miss therefore implies a miss for every descendant, and every active node
// the brute-force loop would report is visited with the very same slab test. The walk is stackless: in the
// breadth-first numbering node n's children are 8n+1..8n+8.
//
// Output order is canonical (segments of a ray sorted by (tmin, node id)); the whole file is compiled without
// floating-point contraction so counts and t values are bit-identical to the oracle.
#include <algorithm>
#include <mutex>

#include "common.hpp"

namespace neat
{

struct TreeArgs
{
    const float* pmin;
    const float* pmax;
    const uint8_t* flags;  // bit 0: active, bit 1: has an active strict descendant
    int32_t first_leaf;    // first node of the deepest level: nodes below it have children 8n+1..8n+8
};

// Sample count of a segment (HyperTree.cu:287-291): iCeil(|tmax - tmin| / |max - min|_2 * max_spn), the norm added
// as x + (y + z) like Eigen's fixed-size reduction.
__device__ __forceinline__ int32_t segment_count(const float mn[3], const float mx[3], float tmin, float tmax,
                                                 int32_t max_spn)
{
    float ex   = mx[0] - mn[0];
    float ey   = mx[1] - mn[1];
    float ez   = mx[2] - mn[2];
    float diag = sqrtf(ex * ex + (ey * ey + ez * ez));
    float dis  = tmax - tmin;
    float rel  = dis / diag;
    return (int32_t)ceilf(rel * (float)max_spn);
}

// Children are visited in the order i ^ mask (bit k of mask set when the ray runs towards -axis k), a front-to-back
// order for most rays, so segments mostly arrive sorted by entry t; the order does not change what is visited.
//
// Boxes are not read from the node table: the walk halves them as HyperTreeBaseImpl's constructor does
// (HyperTree.cpp:94-129: center = (min + max) * 0.5, child i takes the upper half on axis k iff bit k of i), starting
// from the root's box, and recovers a parent's box from a child's by doubling its extent (exact: every box of the
// construction from [-1, 1] is dyadic). The only memory the walk reads per step is the 8 flag bytes of a sibling
// group, fetched together when the walk enters the group.
__device__ __forceinline__ uint64_t sibling_flags(const TreeArgs& T, int32_t parent)
{
    const uint8_t* f = T.flags + 8 * (int64_t)parent + 1;
    uint64_t v       = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) v |= (uint64_t)f[i] << (8 * i);
    return v;
}

template <typename Visit>
__device__ __forceinline__ void traverse(const TreeArgs& T, const float o[3], const float d[3], Visit&& visit)
{
    const int32_t mask = (d[0] < 0.f ? 1 : 0) | (d[1] < 0.f ? 2 : 0) | (d[2] < 0.f ? 4 : 0);
    float mn[3] = {T.pmin[0], T.pmin[1], T.pmin[2]};
    float mx[3] = {T.pmax[0], T.pmax[1], T.pmax[2]};
    int32_t n    = 0;
    uint32_t f   = T.flags[0];
    uint64_t sib = 0;  // flags of n's sibling group (byte c = child c of n's parent)
    while (true)
    {
        bool descend = false;
        if (f)
        {
            Hit h = slab_test(mn, mx, o, d);
            if (h.hit)
            {
                // sliver filter (HyperTree.cu:274): float difference compared in double
                if ((f & 1) && !((double)(h.tmax - h.tmin) < 1e-5)) visit(n, h.tmin, h.tmax, mn, mx);
                descend = (f & 2) != 0 && n < T.first_leaf;  // never leave the node table
            }
        }
        if (descend)
        {
            sib = sibling_flags(T, n);
            n   = 8 * n + 1 + mask;
#pragma unroll
            for (int k = 0; k < 3; ++k)
            {
                const float c = (mn[k] + mx[k]) * 0.5f;
                if ((mask >> k) & 1)
                    mn[k] = c;
                else
                    mx[k] = c;
            }
            f = (uint32_t)(sib >> (8 * mask)) & 0xFFu;
            continue;
        }
        // to the next sibling in visit order, going up while n is the last of its group
        bool moved = false;
        while (n != 0)
        {
            const int32_t c = (n - 1) & 7, i = c ^ mask;
            // the parent's box, from n's
            float pmn[3], pmx[3];
#pragma unroll
            for (int k = 0; k < 3; ++k)
            {
                const float e = mx[k] - mn[k];
                pmn[k]        = ((c >> k) & 1) ? mn[k] - e : mn[k];
                pmx[k]        = ((c >> k) & 1) ? mx[k] : mx[k] + e;
            }
            if (i < 7)
            {
                const int32_t c2 = (i + 1) ^ mask;
                n                = ((n - 1) & ~7) + 1 + c2;
#pragma unroll
                for (int k = 0; k < 3; ++k)
                {
                    const float cc = (pmn[k] + pmx[k]) * 0.5f;
                    mn[k]          = ((c2 >> k) & 1) ? cc : pmn[k];
                    mx[k]          = ((c2 >> k) & 1) ? pmx[k] : cc;
                }
                f     = (uint32_t)(sib >> (8 * c2)) & 0xFFu;
                moved = true;
                break;
            }
            n = (n - 1) >> 3;
#pragma unroll
            for (int k = 0; k < 3; ++k)
            {
                mn[k] = pmn[k];
                mx[k] = pmx[k];
            }
            if (n != 0) sib = sibling_flags(T, (n - 1) >> 3);
        }
        if (!moved) break;
    }
}

__device__ __forceinline__ void load_ray(const float* ro, const float* rd, int64_t r, float o[3], float d[3])
{
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        o[k] = ro[3 * r + k];
        d[k] = rd[3 * r + k];
    }
}

__global__ void k_zero_i32(int32_t* p)
{
    if (p) *p = 0;
}

__global__ void __launch_bounds__(128) k_ray_count(TreeArgs T, const float* ro, const float* rd, int64_t R,
                                                   int32_t max_spn, int64_t* nseg, int64_t* nsamp, int32_t* max_per_ray)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    float o[3], d[3];
    load_ray(ro, rd, r, o, d);
    int64_t ns = 0, nk = 0;
    traverse(T, o, d,
             [&](int32_t, float tmin, float tmax, const float* mn, const float* mx)
             {
                 nk += 1;
                 ns += segment_count(mn, mx, tmin, tmax, max_spn);
             });
    nseg[r]  = nk;
    nsamp[r] = ns;
    if (max_per_ray) atomicMax(max_per_ray, (int32_t)ns);
}

__device__ __forceinline__ bool seg_less(float ta, int32_t na, float tb, int32_t nb)
{
    return ta < tb || (ta == tb && na < nb);
}

// Segment list of one ray in canonical order into columns at stride `st` from position `base` (at most `cap`
// entries; `count` and `samples` receive the ray's full segment and sample counts): insertion into the sorted
// prefix while walking (segments per ray are few, and mostly arrive in order, which the register copy of the last
// entry answers without a memory read).
__device__ __forceinline__ void walk_sorted(const TreeArgs& T, const float o[3], const float d[3], int32_t max_spn,
                                            int32_t* node_col, int32_t* count_col, float* tmin_col, float* tmax_col,
                                            int64_t base, int64_t st, int64_t cap, int64_t& count, int64_t& samples)
{
    int64_t k      = 0;  // entries written
    int64_t nk     = 0;  // segments seen
    int64_t ns     = 0;
    float last_t   = 0.f;  // the largest (tmin, node) so far, i.e. the entry at k - 1 (valid when k > 0)
    int32_t last_n = 0;
    auto at = [&](int64_t i) { return base + i * st; };
    traverse(T, o, d,
             [&](int32_t node, float tmin, float tmax, const float* mn, const float* mx)
             {
                 const int32_t cnt = segment_count(mn, mx, tmin, tmax, max_spn);
                 nk += 1;
                 ns += cnt;
                 if (k >= cap) return;
                 int64_t p = k;
                 if (k == 0 || !seg_less(tmin, node, last_t, last_n))
                 {
                     last_t = tmin;
                     last_n = node;
                 }
                 else
                 {
                     tmin_col[at(p)]  = last_t;
                     tmax_col[at(p)]  = tmax_col[at(p - 1)];
                     node_col[at(p)]  = last_n;
                     count_col[at(p)] = count_col[at(p - 1)];
                     --p;
                 }
                 while (p > 0 && p < k && seg_less(tmin, node, tmin_col[at(p - 1)], node_col[at(p - 1)]))
                 {
                     tmin_col[at(p)]  = tmin_col[at(p - 1)];
                     tmax_col[at(p)]  = tmax_col[at(p - 1)];
                     node_col[at(p)]  = node_col[at(p - 1)];
                     count_col[at(p)] = count_col[at(p - 1)];
                     --p;
                 }
                 tmin_col[at(p)]  = tmin;
                 tmax_col[at(p)]  = tmax;
                 node_col[at(p)]  = node;
                 count_col[at(p)] = cnt;
                 ++k;
             });
    count   = nk;
    samples = ns;
}

__global__ void __launch_bounds__(128) k_ray_segments(TreeArgs T, const float* ro, const float* rd, int64_t R,
                                                      int32_t max_spn, const int64_t* ray_seg_offset,
                                                      const int64_t* ray_sample_offset, int32_t* seg_node,
                                                      int32_t* seg_ray, int32_t* seg_count, float* seg_tmin,
                                                      float* seg_tmax, int64_t* seg_sample_offset)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    float o[3], d[3];
    load_ray(ro, rd, r, o, d);
    const int64_t base = ray_seg_offset[r];
    const int64_t end  = ray_seg_offset[r + 1];
    int64_t nk, ns;
    // the capacity end - base holds every segment when the offsets come from the same walk
    walk_sorted(T, o, d, max_spn, seg_node, seg_count, seg_tmin, seg_tmax, base, 1, end - base, nk, ns);
    int64_t so = ray_sample_offset[r];
    for (int64_t q = base; q < end; ++q)
    {
        seg_ray[q]           = (int32_t)r;
        seg_sample_offset[q] = so;
        so += seg_count[q];
    }
}

// Single-walk sampler: segment i of ray r goes to column entry i * R + r of a scratch of `cap` entries per ray
// (four columns: node, count, tmin, tmax), with the ray's segment and sample counts; once the offsets are known,
// k_ray_rewalk walks the rays with more than `cap` segments again straight into the canonical list and k_ray_compact
// copies the scratch of the others.
struct RayScratch
{
    int32_t* node;
    int32_t* count;
    float* tmin;
    float* tmax;
    int64_t cap;
};

__global__ void __launch_bounds__(128) k_ray_walk(TreeArgs T, const float* ro, const float* rd, int64_t R,
                                                  int32_t max_spn, RayScratch sc, int64_t* nseg, int64_t* nsamp,
                                                  int32_t* max_per_ray)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    float o[3], d[3];
    load_ray(ro, rd, r, o, d);
    int64_t nk, ns;
    walk_sorted(T, o, d, max_spn, sc.node, sc.count, sc.tmin, sc.tmax, r, R, sc.cap, nk, ns);
    nseg[r]  = nk;
    nsamp[r] = ns;
    if (max_per_ray) atomicMax(max_per_ray, (int32_t)ns);
}

// Rays with more segments than the scratch holds (cap): walked again, straight into the canonical list.
__global__ void __launch_bounds__(128) k_ray_rewalk(TreeArgs T, const float* ro, const float* rd, int64_t R,
                                                    int32_t max_spn, int64_t cap, const int64_t* ray_seg_offset,
                                                    int32_t* seg_node, int32_t* seg_count, float* seg_tmin,
                                                    float* seg_tmax)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const int64_t base = ray_seg_offset[r];
    const int64_t end  = ray_seg_offset[r + 1];
    if (end - base <= cap) return;
    float o[3], d[3];
    load_ray(ro, rd, r, o, d);
    int64_t nk, ns;
    walk_sorted(T, o, d, max_spn, seg_node, seg_count, seg_tmin, seg_tmax, base, 1, end - base, nk, ns);
}

constexpr int kCompactWaves = 4;

// The scratch copied into the canonical list, one wave per 64 consecutive rays: their segments are the contiguous
// output range [ray_seg_offset[r0], ray_seg_offset[r0 + 64]), which the wave writes 64 entries at a time. Each lane
// finds the ray of its entry by a binary search over the 64 ray offsets the lanes hold (largest ray whose first
// segment is at or before the entry; rays without segments share their successor's offset and are passed over).
// seg_sample_offset is the exclusive prefix sum of seg_count in canonical order: the ray's sample offset plus a wave
// scan. Segments of rays over the scratch capacity were written by k_ray_rewalk; only their counts are read back.
__global__ void __launch_bounds__(64 * kCompactWaves) k_ray_compact(int64_t R, RayScratch sc,
                                                                   const int64_t* ray_seg_offset,
                                                                   const int64_t* ray_sample_offset, int32_t* seg_node,
                                                                   int32_t* seg_ray, int32_t* seg_count,
                                                                   float* seg_tmin, float* seg_tmax,
                                                                   int64_t* seg_sample_offset)
{
    const int l      = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * kCompactWaves + (threadIdx.x >> 6)) * 64;
    if (r0 >= R) return;  // uniform over the wave
    const int64_t nr    = min<int64_t>(64, R - r0);
    const int64_t off   = ray_seg_offset[r0 + min<int64_t>(l, nr)];  // lanes >= nr hold the range end
    const int64_t q_beg = __shfl(off, 0);
    const int64_t q_end = ray_seg_offset[r0 + nr];
    int64_t run         = ray_sample_offset[r0];
    for (int64_t qb = q_beg; qb < q_end; qb += 64)
    {
        const int64_t q = qb + l;
        int j           = 0;
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1)
        {
            const int cand   = j + step;  // <= 63
            const int64_t oc = __shfl(off, cand);
            if (oc <= q) j = cand;
        }
        const int64_t oj   = __shfl(off, j);
        const int64_t onx  = __shfl(off, min(j + 1, 63));
        const int64_t next = j == 63 ? q_end : onx;
        int32_t cnt        = 0;
        if (q < q_end)
        {
            const int64_t r = r0 + j;
            if (next - oj > sc.cap)
            {
                cnt = seg_count[q];
            }
            else
            {
                const int64_t e = (q - oj) * R + r;
                cnt             = sc.count[e];
                seg_node[q]     = sc.node[e];
                seg_count[q]    = cnt;
                seg_tmin[q]     = sc.tmin[e];
                seg_tmax[q]     = sc.tmax[e];
            }
            seg_ray[q] = (int32_t)r;
        }
        // inclusive wave scan of the counts
        int32_t inc = cnt;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1)
        {
            const int32_t v = __shfl_up(inc, d);
            if (l >= d) inc += v;
        }
        if (q < q_end) seg_sample_offset[q] = run + (inc - cnt);
        run += __shfl(inc, 63);
    }
}

// One thread per sample; the owning segment is the last one whose first sample is <= i.
__global__ void __launch_bounds__(256) k_ray_emit(const float* ro, const float* rd, const int64_t* ray_seg_offset,
                                                  const int64_t* ray_sample_offset, int64_t K, const int32_t* seg_node,
                                                  const int32_t* seg_ray, const int32_t* seg_count,
                                                  const float* seg_tmin, const float* seg_tmax,
                                                  const int64_t* seg_sample_offset, int64_t S, int32_t jitter,
                                                  uint64_t seed, float* out_t, float* out_w, float* out_coord,
                                                  int64_t* out_ray, int64_t* out_local, int64_t* out_node)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    int64_t lo = 0, hi = K - 1;
    while (lo < hi)
    {
        int64_t mid = (lo + hi + 1) >> 1;
        if (seg_sample_offset[mid] <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int64_t k    = lo;
    const int64_t ray  = seg_ray[k];
    const int32_t n    = seg_count[k];
    const int32_t j    = (int32_t)(i - seg_sample_offset[k]);
    const int32_t rank = (int32_t)(k - ray_seg_offset[ray]);
    const float tmin = seg_tmin[k], tmax = seg_tmax[k];
    const float step = (tmax - tmin) / (float)n;
    const float t    = sample_t(tmin, step, j, seg_a(jitter, seed, ray, rank, j));
    out_t[i]         = t;
    out_w[i]         = sample_w(tmin, tmax, step, n, j, t, jitter, seed, ray, rank);
    if (out_coord)
    {
#pragma unroll
        for (int c = 0; c < 3; ++c) out_coord[3 * i + c] = ro[3 * ray + c] + rd[3 * ray + c] * t;
    }
    out_ray[i]   = ray;
    out_local[i] = i - ray_sample_offset[ray];
    out_node[i]  = seg_node[k];
}

}  // namespace neat

using namespace neat;

static int32_t first_leaf(int32_t num_nodes) { return (num_nodes - 1) / 8; }

static int check_tree(const float* pmin, const float* pmax, const uint8_t* flags, int32_t num_nodes, int32_t max_depth)
{
    NEAT_REQUIRE(pmin && pmax && flags, "octree pointers must not be null");
    NEAT_REQUIRE(max_depth >= 0 && max_depth <= 9, "max_depth %d outside [0, 9]", max_depth);
    int64_t expect = 0, p = 1;
    for (int d = 0; d <= max_depth; ++d, p *= 8) expect += p;
    NEAT_REQUIRE(num_nodes == expect, "num_nodes %d is not a complete octree of depth %d (%lld)", num_nodes, max_depth,
                 (long long)expect);
    return NEAT_OK;
}

// Scratch entries per ray of the single-walk sampler: the most leaves a ray can cross in a depth-D octree with
// positive length (3 * 2^D - 2, one per crossed cell plane, plus margin), at most 128.
static int64_t scratch_cap(int32_t max_depth)
{
    const int64_t c = 3 * ((int64_t)1 << std::min(max_depth, 9)) + 8;
    return std::min<int64_t>(c, 128);
}

static RayScratch scratch_view(void* p, int64_t R, int32_t max_depth)
{
    RayScratch sc;
    sc.cap        = scratch_cap(max_depth);
    const size_t n = (size_t)(sc.cap * R);
    char* b       = static_cast<char*>(p);
    sc.node       = reinterpret_cast<int32_t*>(b);
    sc.count      = reinterpret_cast<int32_t*>(b + 4 * n);
    sc.tmin       = reinterpret_cast<float*>(b + 8 * n);
    sc.tmax       = reinterpret_cast<float*>(b + 12 * n);
    return sc;
}

// ---- zero-allocation count pass ----------------------------------------------------------------------------------
// The scan of the count pass needs (rays / 1024 + 1) int64 tile sums. neat_raysample_count_noalloc keeps them in
// `scratch`, a file-scope device array sized for kNoAllocMaxRays rays (131 KiB; every GPU the library runs on holds
// one copy until the process ends). Host code reaches it only through the address hipGetSymbolAddress returns for
// the current device, fetched once per device and kept in scratch_addr.
constexpr int64_t kNoAllocMaxRays = (int64_t)1 << 24;
static __device__ __align__(256) int64_t scratch[kNoAllocMaxRays / 1024 + 2];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static void* g_scratch_addr[kMaxDevices] = {};

// this device's copy of `scratch` (NEAT_OK), or an error status
static int scratch_addr(void** out)
{
    int dev = 0;
    NEAT_HIP(hipGetDevice(&dev));
    NEAT_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d outside the %d this library tracks", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_addr[dev])
    {
        void* p = nullptr;
        NEAT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        NEAT_REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address for the count-pass scratch");
        g_scratch_addr[dev] = p;
    }
    *out = g_scratch_addr[dev];
    return NEAT_OK;
}

extern "C"
{
    size_t neat_raysample_workspace_bytes(int64_t num_rays) { return scan_workspace_bytes(num_rays); }

    size_t neat_raysample_scratch_bytes(int64_t num_rays, int32_t max_depth)
    {
        if (num_rays < 0 || max_depth < 0 || max_depth > 9) return 0;
        return (size_t)(16 * scratch_cap(max_depth) * std::max<int64_t>(num_rays, 1));
    }

    int neat_raysample_walk(neat_stream_t stream, const float* pmin, const float* pmax, const uint8_t* flags,
                            int32_t num_nodes, int32_t max_depth, const float* ro, const float* rd, int64_t R,
                            int32_t max_spn, int64_t* ray_seg_offset, int64_t* ray_sample_offset,
                            int32_t* max_per_ray, void* scratch, size_t scratch_bytes, void* workspace,
                            size_t workspace_bytes)
    {
        int st = check_tree(pmin, pmax, flags, num_nodes, max_depth);
        if (st) return st;
        NEAT_REQUIRE(R >= 0, "num_rays %lld < 0", (long long)R);
        NEAT_REQUIRE(max_spn > 0, "max_samples_per_node %d must be > 0", max_spn);
        NEAT_REQUIRE(ray_seg_offset && ray_sample_offset, "offset arrays must not be null");
        NEAT_REQUIRE(R == 0 || (ro && rd), "ray arrays must not be null");
        NEAT_REQUIRE(R == 0 || (scratch && scratch_bytes >= neat_raysample_scratch_bytes(R, max_depth)),
                     "neat_raysample_walk: scratch too small");
        NEAT_REQUIRE((reinterpret_cast<uintptr_t>(scratch) % 16) == 0, "neat_raysample_walk: scratch alignment");
        hipStream_t s = as_stream(stream);
        k_zero_i32<<<1, 1, 0, s>>>(max_per_ray);
        NEAT_LAUNCHED();
        if (R > 0)
        {
            TreeArgs T{pmin, pmax, flags, first_leaf(num_nodes)};
            k_ray_walk<<<grid_for(R, 128), 128, 0, s>>>(T, ro, rd, R, max_spn, scratch_view(scratch, R, max_depth),
                                                         ray_seg_offset, ray_sample_offset, max_per_ray);
            NEAT_LAUNCHED();
        }
        st = exclusive_scan_i64(s, ray_seg_offset, R, workspace, workspace_bytes);
        if (st) return st;
        return exclusive_scan_i64(s, ray_sample_offset, R, workspace, workspace_bytes);
    }

    int neat_raysample_compact(neat_stream_t stream, const float* pmin, const float* pmax, const uint8_t* flags,
                               int32_t num_nodes, int32_t max_depth, const float* ro, const float* rd, int64_t R,
                               int32_t max_spn, const int64_t* ray_seg_offset, const int64_t* ray_sample_offset,
                               const void* scratch, size_t scratch_bytes, int64_t K, int32_t* seg_node,
                               int32_t* seg_ray, int32_t* seg_count, float* seg_tmin, float* seg_tmax,
                               int64_t* seg_sample_offset)
    {
        int st = check_tree(pmin, pmax, flags, num_nodes, max_depth);
        if (st) return st;
        NEAT_REQUIRE(R >= 0 && K >= 0, "negative size");
        NEAT_REQUIRE(max_spn > 0, "max_samples_per_node %d must be > 0", max_spn);
        if (R == 0 || K == 0) return NEAT_OK;
        NEAT_REQUIRE(ro && rd && ray_seg_offset && ray_sample_offset && seg_node && seg_ray && seg_count && seg_tmin &&
                         seg_tmax && seg_sample_offset,
                     "neat_raysample_compact: null pointer");
        NEAT_REQUIRE(scratch && scratch_bytes >= neat_raysample_scratch_bytes(R, max_depth),
                     "neat_raysample_compact: scratch too small");
        TreeArgs T{pmin, pmax, flags, first_leaf(num_nodes)};
        hipStream_t s        = as_stream(stream);
        const RayScratch sc  = scratch_view(const_cast<void*>(scratch), R, max_depth);
        k_ray_rewalk<<<grid_for(R, 128), 128, 0, s>>>(T, ro, rd, R, max_spn, sc.cap, ray_seg_offset, seg_node,
                                                       seg_count, seg_tmin, seg_tmax);
        NEAT_LAUNCHED();
        k_ray_compact<<<grid_for(R, 64 * kCompactWaves), 64 * kCompactWaves, 0, s>>>(
            R, sc, ray_seg_offset, ray_sample_offset, seg_node, seg_ray, seg_count, seg_tmin, seg_tmax,
            seg_sample_offset);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_raysample_count(neat_stream_t stream, const float* pmin, const float* pmax, const uint8_t* flags,
                             int32_t num_nodes, int32_t max_depth, const float* ro, const float* rd, int64_t R,
                             int32_t max_spn, int64_t* ray_seg_offset, int64_t* ray_sample_offset,
                             int32_t* max_per_ray, void* workspace, size_t workspace_bytes)
    {
        int st = check_tree(pmin, pmax, flags, num_nodes, max_depth);
        if (st) return st;
        NEAT_REQUIRE(R >= 0, "num_rays %lld < 0", (long long)R);
        NEAT_REQUIRE(max_spn > 0, "max_samples_per_node %d must be > 0", max_spn);
        NEAT_REQUIRE(ray_seg_offset && ray_sample_offset, "offset arrays must not be null");
        NEAT_REQUIRE(R == 0 || (ro && rd), "ray arrays must not be null");
        hipStream_t s = as_stream(stream);
        k_zero_i32<<<1, 1, 0, s>>>(max_per_ray);
        NEAT_LAUNCHED();
        if (R > 0)
        {
            TreeArgs T{pmin, pmax, flags, first_leaf(num_nodes)};
            k_ray_count<<<grid_for(R, 128), 128, 0, s>>>(T, ro, rd, R, max_spn, ray_seg_offset, ray_sample_offset,
                                                          max_per_ray);
            NEAT_LAUNCHED();
        }
        st = exclusive_scan_i64(s, ray_seg_offset, R, workspace, workspace_bytes);
        if (st) return st;
        return exclusive_scan_i64(s, ray_sample_offset, R, workspace, workspace_bytes);
    }

    int neat_raysample_count_noalloc(neat_stream_t stream, const float* pmin, const float* pmax,
                                     const uint8_t* flags, int32_t num_nodes, int32_t max_depth, const float* ro,
                                     const float* rd, int64_t R, int32_t max_spn, int64_t* ray_seg_offset,
                                     int64_t* ray_sample_offset, int32_t* max_per_ray)
    {
        if (R > kNoAllocMaxRays)
            return fail(NEAT_E_OVERFLOW,
                        "neat_raysample_count_noalloc: %lld rays exceed the %lld its scan scratch holds; call "
                        "neat_raysample_count with a workspace",
                        (long long)R, (long long)kNoAllocMaxRays);
        void* ptr = nullptr;
        int st    = scratch_addr(&ptr);
        if (st) return st;
        // the scan writes every tile sum it reads: nothing in the scratch needs clearing between calls
        return neat_raysample_count(stream, pmin, pmax, flags, num_nodes, max_depth, ro, rd, R, max_spn,
                                    ray_seg_offset, ray_sample_offset, max_per_ray, ptr, sizeof(scratch));
    }

    int neat_raysample_segments(neat_stream_t stream, const float* pmin, const float* pmax, const uint8_t* flags,
                                int32_t num_nodes, int32_t max_depth, const float* ro, const float* rd, int64_t R,
                                int32_t max_spn, const int64_t* ray_seg_offset, const int64_t* ray_sample_offset,
                                int64_t K, int32_t* seg_node, int32_t* seg_ray, int32_t* seg_count, float* seg_tmin,
                                float* seg_tmax, int64_t* seg_sample_offset)
    {
        int st = check_tree(pmin, pmax, flags, num_nodes, max_depth);
        if (st) return st;
        NEAT_REQUIRE(R >= 0 && K >= 0, "negative size");
        NEAT_REQUIRE(max_spn > 0, "max_samples_per_node %d must be > 0", max_spn);
        if (R == 0 || K == 0) return NEAT_OK;
        NEAT_REQUIRE(ro && rd && ray_seg_offset && ray_sample_offset && seg_node && seg_ray && seg_count && seg_tmin &&
                         seg_tmax && seg_sample_offset,
                     "neat_raysample_segments: null pointer");
        TreeArgs T{pmin, pmax, flags, first_leaf(num_nodes)};
        k_ray_segments<<<grid_for(R, 128), 128, 0, as_stream(stream)>>>(
            T, ro, rd, R, max_spn, ray_seg_offset, ray_sample_offset, seg_node, seg_ray, seg_count, seg_tmin, seg_tmax,
            seg_sample_offset);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_raysample_emit(neat_stream_t stream, const float* ro, const float* rd, int64_t R,
                            const int64_t* ray_seg_offset, const int64_t* ray_sample_offset, int64_t K,
                            const int32_t* seg_node, const int32_t* seg_ray, const int32_t* seg_count,
                            const float* seg_tmin, const float* seg_tmax, const int64_t* seg_sample_offset, int64_t S,
                            int32_t jitter, uint64_t seed, float* out_t, float* out_w, float* out_coord,
                            int64_t* out_ray, int64_t* out_local, int64_t* out_node)
    {
        NEAT_REQUIRE(R >= 0 && K >= 0 && S >= 0, "negative size");
        if (S == 0) return NEAT_OK;
        NEAT_REQUIRE(K > 0, "num_samples %lld > 0 with no segments", (long long)S);
        NEAT_REQUIRE(ro && rd && ray_seg_offset && ray_sample_offset && seg_node && seg_ray && seg_count && seg_tmin &&
                         seg_tmax && seg_sample_offset && out_t && out_w && out_ray && out_local && out_node,
                     "neat_raysample_emit: null pointer");
        k_ray_emit<<<grid_for(S, 256), 256, 0, as_stream(stream)>>>(ro, rd, ray_seg_offset, ray_sample_offset, K,
                                                                     seg_node, seg_ray, seg_count, seg_tmin, seg_tmax,
                                                                     seg_sample_offset, S, jitter, seed, out_t, out_w,
                                                                     out_coord, out_ray, out_local, out_node);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }
}
