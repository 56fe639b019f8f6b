/*
 * neat_amd.h -- C ABI of the MI355X (gfx950) differentiable forward projector.
 *
 * Plain pointers, sizes and a HIP stream; no torch types. Every call is asynchronous on `stream` (a
 * hipStream_t passed as void*, NULL = the default stream) and returns a status code; on failure
 * neat_last_error() returns a thread-local message. Device arrays are caller-allocated; functions that
 * need scratch memory take (workspace, workspace_bytes) and publish the size through a *_workspace_bytes
 * query. No entry point keeps state between calls, so concurrent calls on different streams are safe.
 *
 * Conventions shared by every entry point (they restate the reference's, see the citations):
 *  - Octree: a complete octree of depth max_depth, numbered breadth first; node n's children are
 *    8n+1 .. 8n+8 and child i takes the upper half on axis k iff bit k of i is set
 *    (HyperTree.cpp:94-129). node_pos_min / node_pos_max are float [num_nodes, 3].
 *  - node_flags: uint8 [num_nodes], bit 0 = node active, bit 1 = some strict descendant is active.
 *    neat_tree_flags_host() derives it from the reference's node_active buffer (HyperTree.h:229-230).
 *  - Rays: float [R, 3] origin and direction (RayList, HelperStructs.h:92-98).
 *  - Canonical sample order: rays in input order; within a ray, segments by (tmin, node id); within a
 *    segment, samples by index j. (The reference's order depends on atomics, HyperTree.cu:281,301.)
 *  - Grids: float [N, C, Z, Y, X] addressed through element strides, so both the reference's contiguous
 *    layout and the channels-last-3d layout (C stride 1) are accepted; channels-last is the fast path.
 *  - Decoder parameters are one packed float array: for each Linear layer in order, W [out, in] row-major,
 *    then b [out] (FCBlockImpl, ImplicitNet.h:18-71).
 */
#ifndef NEAT_AMD_H
#define NEAT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ----------------------------------------------------------------------------------------------- */
#define NEAT_OK 0
#define NEAT_E_ARG 1         /* invalid argument: null pointer, negative size, inconsistent shape        */
#define NEAT_E_OVERFLOW 2    /* a caller-provided buffer (workspace or output) is too small               */
#define NEAT_E_HIP 3         /* the HIP runtime reported an error                                         */
#define NEAT_E_UNSUPPORTED 4 /* a configuration this build has no kernel for (message names it)          */

/* activation ids (Saiga::ActivationFromString names and DecodeFeatures' last activation, geometry.cpp:123-143) */
#define NEAT_ACT_ID 0
#define NEAT_ACT_RELU 1
#define NEAT_ACT_SILU 2
#define NEAT_ACT_SOFTPLUS 3
#define NEAT_ACT_ABS 4

typedef void* neat_stream_t;

/* No reference counterpart (the reference aborts through glog CHECK). Thread-local message of the last
 * failing call on this thread. */
const char* neat_last_error(void);
int neat_abi_version(void);
/* No reference counterpart. Hash of the sources this library was built from (16 hex digits; neat_amd/buildinfo.py
 * computes it the same way), so a loader can refuse a library built from other sources. */
const char* neat_build_id(void);

/* ---- octree (host helpers, no device work) ------------------------------------------------------------------ */

/* Number of nodes of the complete octree of depth max_depth (HyperTree.cpp:13-25), -1 outside [0, 9]. */
int64_t neat_tree_num_nodes(int32_t max_depth);

/* Node table of HyperTreeBaseImpl's constructor (HyperTree.cpp:11-132) into host arrays: parent [n], children
 * [n, 8] (-1 on the deepest level), pos_min / pos_max [n, 3], scale [n], depth [n]. */
int neat_tree_build_host(int32_t max_depth, int32_t* parent, int32_t* children, float* pos_min, float* pos_max,
                         float* scale, int32_t* depth);

/* active_node_ids (ascending ids of nodes with node_active == 1) and node_active_prefix_sum (the node's grid slot,
 * -1 when inactive), the host loop of HyperTreeBaseImpl::UpdateActive (HyperTree.cpp:169-206). Returns the active
 * count, -1 on bad arguments. */
int64_t neat_tree_update_active_host(const int32_t* node_active, int32_t num_nodes, int64_t* active_ids,
                                     int64_t* node_active_prefix_sum);

/* Derives node_flags from node_active (int32 0/1 [num_nodes], HyperTree.h:229-230) for the traversal used by
 * the ray sampler. Host memory in and out. Replaces the host loop of HyperTreeBaseImpl::UpdateActive
 * (HyperTree.cpp:169-206) as far as the sampler needs it. */
int neat_tree_flags_host(const int32_t* node_active, int32_t num_nodes, int32_t max_depth, uint8_t* node_flags);

/* Checks on the host that a node table is the canonical breadth-first dyadic octree of depth max_depth
 * (the table HyperTreeBaseImpl's constructor builds, HyperTree.cpp:11-132) that the sampler's traversal
 * relies on. Returns NEAT_E_ARG, with the first mismatch in the message, otherwise NEAT_OK. */
int neat_tree_validate_host(const float* node_pos_min, const float* node_pos_max, const int32_t* node_children,
                            int32_t num_nodes, int32_t max_depth);

/* ---- ray / leaf intersection and sample generation --------------------------------------------------------------
 * Together these replace HyperTreeBaseImpl::CreateSamplesForRays (HyperTree.h:157, HyperTree.cu:393-438) and its
 * kernel ComputeRaySamples (HyperTree.cu:206-391): exact allocation (count -> scan -> emit) instead of a
 * predicted buffer, no 256-segment cap, canonical order. */

size_t neat_raysample_workspace_bytes(int64_t num_rays);

/* Pass 1: per-ray segment and sample counts, written as exclusive offsets [num_rays + 1] (totals at index
 * num_rays; the caller reads them back to size pass 2). max_samples_per_ray_out (device int32, may be NULL)
 * receives the largest per-ray sample count (SampleList::max_samples_per_ray, HyperTree.cu:302,422). */
int neat_raysample_count(neat_stream_t stream, const float* node_pos_min, const float* node_pos_max,
                         const uint8_t* node_flags, int32_t num_nodes, int32_t max_depth, const float* ray_origin,
                         const float* ray_direction, int64_t num_rays, int32_t max_samples_per_node,
                         int64_t* ray_seg_offset, int64_t* ray_sample_offset, int32_t* max_samples_per_ray_out,
                         void* workspace, size_t workspace_bytes);

/* No counterpart in the reference: neat_raysample_count without a workspace, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The scan's tile sums live in
 * a static device array of 131 KiB per GPU, which limits num_rays to 2^24 (NEAT_E_OVERFLOW beyond); it allocates
 * nothing. Calls that overlap on one GPU share that array and race: serialize them (one stream, or events). */
int neat_raysample_count_noalloc(neat_stream_t stream, const float* node_pos_min, const float* node_pos_max,
                                 const uint8_t* node_flags, int32_t num_nodes, int32_t max_depth,
                                 const float* ray_origin, const float* ray_direction, int64_t num_rays,
                                 int32_t max_samples_per_node, int64_t* ray_seg_offset, int64_t* ray_sample_offset,
                                 int32_t* max_samples_per_ray_out);

/* Pass 2: the canonical segment list [num_segments]: node id, ray, sample count, tmin, tmax and the global
 * offset of the segment's first sample. */
int neat_raysample_segments(neat_stream_t stream, const float* node_pos_min, const float* node_pos_max,
                            const uint8_t* node_flags, int32_t num_nodes, int32_t max_depth, const float* ray_origin,
                            const float* ray_direction, int64_t num_rays, int32_t max_samples_per_node,
                            const int64_t* ray_seg_offset, const int64_t* ray_sample_offset, int64_t num_segments,
                            int32_t* seg_node, int32_t* seg_ray, int32_t* seg_count, float* seg_tmin, float* seg_tmax,
                            int64_t* seg_sample_offset);

/* Single-walk form of passes 1 and 2 (the path the module layer takes): neat_raysample_walk walks every ray once,
 * writes the counts and offsets as neat_raysample_count does, and keeps each ray's canonical segments in a caller
 * scratch of neat_raysample_scratch_bytes(num_rays, max_depth) bytes (16-byte aligned); neat_raysample_compact then
 * copies them into the canonical segment list as neat_raysample_segments writes it (a ray with more segments than
 * the scratch holds per ray is walked again there). Same replacement as above (HyperTree.cu:206-438). */
size_t neat_raysample_scratch_bytes(int64_t num_rays, int32_t max_depth);
int neat_raysample_walk(neat_stream_t stream, const float* node_pos_min, const float* node_pos_max,
                        const uint8_t* node_flags, int32_t num_nodes, int32_t max_depth, const float* ray_origin,
                        const float* ray_direction, int64_t num_rays, int32_t max_samples_per_node,
                        int64_t* ray_seg_offset, int64_t* ray_sample_offset, int32_t* max_samples_per_ray_out,
                        void* scratch, size_t scratch_bytes, void* workspace, size_t workspace_bytes);
int neat_raysample_compact(neat_stream_t stream, const float* node_pos_min, const float* node_pos_max,
                           const uint8_t* node_flags, int32_t num_nodes, int32_t max_depth, const float* ray_origin,
                           const float* ray_direction, int64_t num_rays, int32_t max_samples_per_node,
                           const int64_t* ray_seg_offset, const int64_t* ray_sample_offset, const void* scratch,
                           size_t scratch_bytes, int64_t num_segments, int32_t* seg_node, int32_t* seg_ray,
                           int32_t* seg_count, float* seg_tmin, float* seg_tmax, int64_t* seg_sample_offset);

/* Pass 3: SampleList columns (HelperStructs.h:168-189) for every sample: t, weight (midpoint cell width),
 * global coordinate o + d*t [S,3] (may be NULL), ray index, index within the ray, node id. jitter != 0 draws the
 * in-cell position from the counter hash of (seed, ray, segment rank, j); jitter == 0 uses the cell midpoint. */
int neat_raysample_emit(neat_stream_t stream, const float* ray_origin, const float* ray_direction, int64_t num_rays,
                        const int64_t* ray_seg_offset, const int64_t* ray_sample_offset, int64_t num_segments,
                        const int32_t* seg_node, const int32_t* seg_ray, const int32_t* seg_count,
                        const float* seg_tmin, const float* seg_tmax, const int64_t* seg_sample_offset,
                        int64_t num_samples, int32_t jitter, uint64_t seed, float* out_t, float* out_weight,
                        float* out_global_coordinate, int64_t* out_ray_index, int64_t* out_local_index,
                        int64_t* out_node_id);

/* ---- local coordinates and slots ---------------------------------------------------------------------------- */

/* u = (g - min[node]) / (max[node] - min[node]) * 2 - 1, float [S,3] (HyperTreeBaseImpl::ComputeLocalSamples,
 * HyperTree.h:134, HyperTree.cu:177-192). */
int neat_local_samples(neat_stream_t stream, const float* global_coordinate, const int64_t* node_id,
                       int64_t num_samples, const float* node_pos_min, const float* node_pos_max, int32_t num_nodes,
                       float* out_local);

/* grad_global = grad_local * 2 / (max - min): the autograd backward of the expression above. */
int neat_local_samples_bwd(neat_stream_t stream, const float* grad_local, const int64_t* node_id,
                           int64_t num_samples, const float* node_pos_min, const float* node_pos_max,
                           int32_t num_nodes, float* grad_global);

/* slot = node_active_prefix_sum[node_id] (GlobalNodeIdToLocalActiveId, HyperTree.h:183-187).
 * Node ids outside [0, num_nodes) give local coordinates / gradients of 0 and slot -1 instead of reading out of
 * bounds. */
int neat_node_to_slot(neat_stream_t stream, const int64_t* node_id, int64_t num_samples,
                      const int64_t* node_active_prefix_sum, int32_t num_nodes, int64_t* out_slot);

/* ---- indirect trilinear grid sample ------------------------------------------------------------------------------
 * out[i, c] = trilinear(grid[index[i], c], uv[i]) with align_corners and border clamp
 * (IndirectGridSample3D, IndirectGridSample3D.h:17-30; kernels IndirectGridSample3D.cu:432-572 / 584-791).
 * strides: element strides of grid dims (N, C, Z, Y, X). uv: float [S,3] contiguous; out: float [S,C]. */
int neat_gs3d_fwd(neat_stream_t stream, const float* grid, int64_t N, int32_t C, int32_t Z, int32_t Y, int32_t X,
                  const int64_t* strides, const int64_t* index, const float* uv, int64_t num_samples, float* out);

/* grad_grid (same strides as grid) is accumulated and must be zeroed by the caller; grad_uv [S,3] is written
 * (may be NULL). */
int neat_gs3d_bwd(neat_stream_t stream, const float* grid, int64_t N, int32_t C, int32_t Z, int32_t Y, int32_t X,
                  const int64_t* strides, const int64_t* index, const float* uv, int64_t num_samples,
                  const float* grad_out, float* grad_grid, float* grad_uv);

/* ---- decoder MLP ---------------------------------------------------------------------------------------------------
 * FCBlockImpl::forward (ImplicitNet.h:61-67) followed by the last activation of DecodeFeatures
 * (geometry.cpp:123-143): Linear(in,hidden) -> act -> [Linear(hidden,hidden) -> act] x hidden_layers ->
 * Linear(hidden,out) -> last_act. Supported: hidden in {16,32,64}, hidden_layers in {0,1,2}, in <= 16, out <= 4. */
int64_t neat_decoder_num_params(int32_t in, int32_t hidden, int32_t hidden_layers, int32_t out);

int neat_decoder_fwd(neat_stream_t stream, const float* x, int64_t num_samples, int32_t in, int32_t hidden,
                     int32_t hidden_layers, int32_t out, const float* params, int32_t act, int32_t last_act,
                     float softplus_beta, float* y);

size_t neat_decoder_workspace_bytes(int64_t num_samples, int32_t in, int32_t hidden, int32_t hidden_layers,
                                    int32_t out);

/* grad_params (packed, written -- not accumulated) and grad_x [S,in] (may be NULL) for upstream grad_y [S,out]. */
int neat_decoder_bwd(neat_stream_t stream, const float* x, int64_t num_samples, int32_t in, int32_t hidden,
                     int32_t hidden_layers, int32_t out, const float* params, int32_t act, int32_t last_act,
                     float softplus_beta, const float* grad_y, float* grad_params, float* grad_x, void* workspace,
                     size_t workspace_bytes);

/* ---- per-ray line integral --------------------------------------------------------------------------------------
 * out[c, ray_index[i]] += values[i, c] * weight[i]  (NeuralGeometry::IntegrateSamplesXRay, geometry.h:112-113,
 * geometry.cpp:393-424). With ray_sample_offset [R+1] (samples of ray r are [off[r], off[r+1])) the sum is a
 * deterministic segmented reduction; with NULL it falls back to atomics like the reference's index_add_.
 * out [C, R] is written (not accumulated). */
int neat_integrate_fwd(neat_stream_t stream, const float* values, const float* weight, const int64_t* ray_index,
                       int64_t num_samples, int32_t C, int64_t num_rays, const int64_t* ray_sample_offset, float* out);

/* grad_values[i, c] = grad_out[c, ray_index[i]] * weight[i]. */
int neat_integrate_bwd(neat_stream_t stream, const float* grad_out, const float* weight, const int64_t* ray_index,
                       int64_t num_samples, int32_t C, int64_t num_rays, float* grad_values);

/* ---- loss -----------------------------------------------------------------------------------------------------------
 * loss = mean((pred - target)^2) reduced in FP64 (reconstruct.cpp:378-381 computes it in FP32); grad_pred (may be
 * NULL) = 2 (pred - target) / n * grad_scale. out_loss: one device double. */
size_t neat_mse_workspace_bytes(int64_t n);
int neat_mse_fp64(neat_stream_t stream, const float* pred, const float* target, int64_t n, float grad_scale,
                  double* out_loss, float* grad_pred, void* workspace, size_t workspace_bytes);

/* ---- fused projector -------------------------------------------------------------------------------------------------
 * One call per direction for the whole chain of HierarchicalNeuralGeometry::ComputeImage (geometry.h:170,
 * geometry.cpp:347-389) over GeometryExEx::SampleVolumeIndirect (geometry_ex_ex.cpp:161-180): sample positions are
 * recomputed from the canonical segment list (neat_raysample_segments), gathered from the node grids, decoded,
 * integrated per ray; the backward returns grid and decoder gradients. Covered configuration: grid cubic and
 * channels-last contiguous, in = channels in {4, 8}, hidden in {32, 64}, one hidden layer, out = 1, relu or silu
 * hidden activation, softplus output; anything else returns NEAT_E_UNSUPPORTED (use the unfused entry points). */
typedef struct neat_projector_args
{
    /* octree */
    const float* node_pos_min;
    const float* node_pos_max;
    const int64_t* node_active_prefix_sum; /* node -> grid slot, -1 when inactive */
    int32_t num_nodes;
    /* rays and their canonical segments */
    const float* ray_origin;
    const float* ray_direction;
    int64_t num_rays;
    const int64_t* ray_seg_offset; /* [num_rays + 1] */
    int64_t num_segments;
    const int32_t* seg_node;
    const int32_t* seg_ray;
    const int32_t* seg_count;
    const float* seg_tmin;
    const float* seg_tmax;
    int64_t num_samples; /* total of seg_count */
    int32_t jitter;
    uint64_t seed;
    /* feature grid [num_slots, channels, G, G, G] with element strides */
    const float* grid;
    int64_t num_slots;
    int32_t channels;
    int32_t grid_size;
    int64_t grid_strides[5];
    /* decoder (packed parameters) */
    int32_t in_features;
    int32_t hidden;
    int32_t hidden_layers;
    int32_t out_features;
    const float* params;
    int32_t act;
    int32_t last_act;
    float softplus_beta;
    /* arithmetic: NEAT_PREC_FP32 (default), NEAT_PREC_BF16 (grid read from grid_bf16, decoder weights rounded to BF16,
     * FP32 accumulation), NEAT_PREC_FP32_VALU (the first FP32 kernels on the vector ALUs, kept for comparison) */
    int32_t precision;
    /* BF16 copy of the grid, same shape and channels-last layout (NEAT_PREC_BF16 only; see neat_grid_to_bf16) */
    const uint16_t* grid_bf16;
    /* optional proposal CDF [num_segments][8] (see neat_proposal_fwd): a segment whose row has cdf[7] > 0 places its
     * seg_count samples in equal-probability cells of that piecewise-constant density, each weighted by its cell's
     * length; rows with cdf[7] == 0, or a NULL pointer, keep the uniform samples of the reference */
    const float* seg_cdf;
} neat_projector_args;

#define NEAT_PREC_FP32 0
#define NEAT_PREC_BF16 1
#define NEAT_PREC_FP32_VALU 2

/* dst[i] = round-to-nearest-even BF16 of src[i] (the BF16 grid copy the NEAT_PREC_BF16 projector reads). */
int neat_grid_to_bf16(neat_stream_t stream, const float* src, int64_t n, uint16_t* dst);

/* Workspace for plan + forward + backward of these arguments (depends on the counts, not on the data). */
size_t neat_projector_workspace_bytes(const neat_projector_args* args);

/* Sorts the segments by grid slot and cuts the samples into node-local chunks. Must run (on the same workspace)
 * after the segments change and before neat_projector_fwd / _bwd. */
int neat_projector_plan(neat_stream_t stream, const neat_projector_args* args, void* workspace, size_t workspace_bytes);

/* proj [num_rays] = per-ray line integral of softplus(decoder(trilinear(grid))) (compensated != 0: Neumaier sum of
 * the per-segment partials). */
int neat_projector_fwd(neat_stream_t stream, const neat_projector_args* args, int32_t compensated, float* proj,
                       void* workspace, size_t workspace_bytes);

/* Gradients of sum_r grad_proj[r] * proj[r]: grad_grid (same layout as the grid, fully written) and grad_params
 * (packed, written). No activations are kept from the forward; the BF16 form with relu hidden layers reads each
 * sample's BF16 features, weight and grid cell, which neat_projector_fwd leaves in the workspace (24 B per sample,
 * the bits it computed), instead of recomputing them: like the reference's autograd backward, it must follow
 * neat_projector_fwd on the same workspace and arguments. The other forms recompute the samples. */
int neat_projector_bwd(neat_stream_t stream, const neat_projector_args* args, const float* grad_proj,
                       float* grad_grid, float* grad_params, void* workspace, size_t workspace_bytes);

/* The backward in its three stages, for timing them and for overlapping them with other work: 1 the decoder backward
 * (dL/dfeature per sample and per-workgroup parameter partials, into the workspace), 2 the parameter gradient (the
 * partials summed in a fixed order, into grad_params), 3 the grid gradient (the scatter, into grad_grid). Stages 1, 2,
 * 3 in order on one stream are neat_projector_bwd; only the buffer of a stage's output needs to be non-NULL. */
int neat_projector_bwd_stage(neat_stream_t stream, const neat_projector_args* args, int32_t stage,
                             const float* grad_proj, float* grad_grid, float* grad_params, void* workspace,
                             size_t workspace_bytes);

/* Gradients of sum_r grad_proj[r] * proj[r] with respect to the rays' origins and directions, grad_origin and
 * grad_direction [num_rays, 3] (either may be NULL), the samples' t and weights held fixed: the positional gradient
 * the reference's chain yields for rays that require grad (the o + d t recompute, HyperTree.cu:426-432; the local
 * coordinates' backward, HyperTree.cu:116-199; the grid sample's uv gradient, IndirectGridSample3D.cu:752-790; the
 * index_select backward onto the rays). Reads the dL/dfeature of the decoder backward: call it after
 * neat_projector_bwd (or its stage 1) on the same workspace and arguments. Deterministic: per-segment sums in sample
 * order, then per-ray sums in canonical segment order. */
int neat_projector_ray_grad(neat_stream_t stream, const neat_projector_args* args, float* grad_origin,
                            float* grad_direction, void* workspace, size_t workspace_bytes);

/* Per-sample decoded values softplus(decoder(trilinear(grid))) (unweighted) of the planned segment list, in the
 * plan's node-major sample order (values [num_samples]), and seg_value_offset [num_segments] = the first sample of each
 * canonical segment in that order. Matrix-core kernels only (8 channels, NEAT_PREC_FP32 or NEAT_PREC_BF16). */
int neat_projector_sample_values(neat_stream_t stream, const neat_projector_args* args, float* values,
                                 int64_t* seg_value_offset, void* workspace, size_t workspace_bytes);

/* ---- learned proposal (no reference counterpart) -----------------------------------------------------------------
 * A per-segment MLP 6 -> 32 -> 32 -> 8 (BF16 on the matrix cores, FP32 accumulation) from the segment's entry and exit
 * points o + d tmin, o + d tmax to a piecewise-constant density over 8 equal bins of [tmin, tmax]:
 * p = 0.75 softmax(logits) + 0.25 / 8. Packed parameters W0 [32,6], b0 [32], W1 [32,32], b1 [32], W2 [8,32], b2 [8]. */
int64_t neat_proposal_num_params(void);

/* seg_cdf [num_segments][8] = cumulative p (cdf[7] = 1) for segments with seg_count > num_samples, whose count becomes
 * num_samples in seg_count_out (may be NULL); other segments keep their count and get a zero row (uniform sampling,
 * the fallback). logits [num_segments][8] (may be NULL) receives the raw logits. */
int neat_proposal_fwd(neat_stream_t stream, const float* ray_origin, const float* ray_direction,
                      const int32_t* seg_ray, const float* seg_tmin, const float* seg_tmax, const int32_t* seg_count,
                      int64_t num_segments, const float* params, int32_t num_samples, float* seg_cdf,
                      int32_t* seg_count_out, float* logits);

/* Training target [num_segments][8]: bin b gets the total variation of the field's values between consecutive uniform
 * samples whose midpoint lies in bin b, plus 5 % of the mean, normalised (values[seg_value_offset[k] + j], j <
 * seg_count[k]; see neat_projector_sample_values). */
int neat_proposal_targets(neat_stream_t stream, const float* values, const int64_t* seg_value_offset,
                          const int32_t* seg_count, int64_t num_segments, float* target);

/* The fallback's table, one value per grid slot: every slot's value is multiplied by decay, then raised to the total
 * variation sum_j |f_{j+1} - f_j| of the field's values along any of the given uniformly sampled segments that lies in
 * the slot (values as for neat_proposal_targets). */
int neat_proposal_roughness_update(neat_stream_t stream, const float* values, const int64_t* seg_value_offset,
                                   const int32_t* seg_count, const int32_t* seg_node, int64_t num_segments,
                                   const int64_t* node_active_prefix_sum, int32_t num_nodes, int64_t num_slots,
                                   float decay, float* slot_roughness);

/* Uniform in-node sampling as the fallback, by the slot's roughness r: segments with r <= tau keep what
 * neat_proposal_fwd set (seg_count_out, seg_cdf); with tau < r <= tau_uniform a segment that neat_proposal_fwd gave
 * proposal samples gets min(its uniform count, mid_samples) of them (the same CDF row, more cells); with
 * r > tau_uniform (or outside every slot) a segment gets a zero CDF row and its uniform count back (seg_count, the
 * sampler's counts). tau_uniform = tau gives the two-tier gate. */
int neat_proposal_gate(neat_stream_t stream, const int32_t* seg_node, const int32_t* seg_count, int64_t num_segments,
                       const int64_t* node_active_prefix_sum, int32_t num_nodes, int64_t num_slots,
                       const float* slot_roughness, float tau, float tau_uniform, int32_t mid_samples, float* seg_cdf,
                       int32_t* seg_count_out);

/* grad_params = gradient of scale * sum_k CE(target_k, softmax(logits_k)) (scale = 1 / num_segments for the mean). */
size_t neat_proposal_workspace_bytes(int64_t num_segments);
int neat_proposal_bwd(neat_stream_t stream, const float* ray_origin, const float* ray_direction, const int32_t* seg_ray,
                      const float* seg_tmin, const float* seg_tmax, int64_t num_segments, const float* params,
                      const float* target, float scale, float* grad_params, void* workspace, size_t workspace_bytes);

/* ---- volume regulariser (GeometryExEx::VolumeRegularizer, geometry_ex_ex.cpp:86-160) --------------------------------
 * Neighbour pairs of HyperTreeBaseImpl::NodeNeighborSamples (HyperTree.cu:830-940): for every active node a (in
 * active_node_ids order), side s in 0..2, front f in 0..1 and face point sid in 0..G*G-1, the point
 * p = local * (max - min) + min with local[s] = f, local[(s+1)%3] = (sid % G) / (G - 1.f),
 * local[(s+2)%3] = (sid / G) / (G - 1.f) pairs a with the active node holding p + epsilon * (f ? 1 : -1) along axis s
 * (ActiveNodeIdForGlobalPosition, HyperTree.cu:600-640), when that node exists and is not a. Each pair gives two
 * SampleList rows (the reference's layout): rows 2k, 2k + 1 with coordinate p, node ids a and the neighbour,
 * ray_index = row, local_index_in_ray = 0 / 1, weight = 1. Canonical order: (a, s, f, sid) ascending (the reference's
 * follows an atomic counter). max_pairs must be at least num_active * 6 * G * G (the reference's prediction); the
 * row arrays hold 2 * max_pairs rows; out_ray, out_local and out_weight may be NULL. *num_pairs (device) receives the
 * number of pairs. */
size_t neat_neighbor_pairs_workspace_bytes(int64_t num_active, int32_t grid_size);
int neat_neighbor_pairs(neat_stream_t stream, const float* node_pos_min, const float* node_pos_max,
                        const uint8_t* node_flags, int32_t num_nodes, const int64_t* active_node_ids,
                        int64_t num_active, int32_t grid_size, float epsilon, int64_t max_pairs, float* out_coord,
                        int64_t* out_node, int64_t* out_ray, int64_t* out_local, float* out_weight,
                        int64_t* num_pairs, void* workspace, size_t workspace_bytes);

/* Scratch of neat_edge_loss_fwd (num_items = pairs) and of neat_tv_loss_fwd. */
size_t neat_reg_workspace_bytes(int64_t num_items);
size_t neat_tv_workspace_bytes(int64_t num_active, int32_t C, int32_t G);

/* Edge loss (geometry_ex_ex.cpp:121-134) over the features of the neighbour rows, float [2 * num_pairs, C]:
 * loss = mean_k(mean_c |f[2k, c] - f[2k + 1, c]| * r_k) * weight * C with r_k in [0, 1) a splitmix64 hash of
 * (seed, k) in place of torch::rand_like. *loss is FP64 on the device. The backward writes
 * grad_features = d loss / d features * (*grad_loss) (a device scalar). */
int neat_edge_loss_fwd(neat_stream_t stream, const float* features, int64_t num_pairs, int32_t C, uint64_t seed,
                       double weight, double* loss, void* workspace, size_t workspace_bytes);
int neat_edge_loss_bwd(neat_stream_t stream, const float* features, int64_t num_pairs, int32_t C, uint64_t seed,
                       double weight, const float* grad_loss, float* grad_features);

/* Total variation (TVLoss::forward without a per-batch weight, geometry_ex_ex.h:18-61, times loss_tv,
 * geometry_ex_ex.cpp:95-103) of the grids of the given slots of a cubic grid [N, C, G, G, G] (element strides on the
 * host, as for neat_gs3d_fwd): loss = sum over the three axes of mean |forward difference| * C * weight. The backward
 * adds d loss / d grid * (*grad_loss) to grad_grid (same strides). Workspace: neat_tv_workspace_bytes(num_active,
 * C, G); C * G^3 must be below 2^24. */
int neat_tv_loss_fwd(neat_stream_t stream, const float* grid, int64_t N, int32_t C, int32_t G, const int64_t* strides,
                     const int64_t* slots, int64_t num_active, double weight, double* loss, void* workspace,
                     size_t workspace_bytes);
int neat_tv_loss_bwd(neat_stream_t stream, const float* grid, int64_t N, int32_t C, int32_t G, const int64_t* strides,
                     const int64_t* slots, int64_t num_active, double weight, const float* grad_loss,
                     float* grad_grid);
/* Both in one pass over the grid: loss as neat_tv_loss_fwd, and d loss / d grid * (*grad_loss) written into the
 * given slots of grad_grid (overwrite != 0: entries of other slots are left untouched, the caller clears them) or
 * added to it (overwrite == 0). Same workspace as neat_tv_loss_fwd. No counterpart in the reference (its autograd runs
 * the two halves separately); for callers that want the value and the gradient together. */
int neat_tv_loss_fwd_bwd(neat_stream_t stream, const float* grid, int64_t N, int32_t C, int32_t G,
                         const int64_t* strides, const int64_t* slots, int64_t num_active, double weight, double* loss,
                         const float* grad_loss, float* grad_grid, int32_t overwrite, void* workspace,
                         size_t workspace_bytes);

/* The edge loss of GeometryExEx::VolumeRegularizer (geometry_ex_ex.cpp:114-134) in one pass over the candidate face
 * points, replacing the chain NodeNeighborSamples -> ComputeLocalSamples -> IndirectGridSample3D -> edge loss (and its
 * backward through the grid sample) without materialising the neighbour rows or their features. Same pairs, pair
 * order, random factor and per-feature gradient as that chain. node_slot[node] = the node's grid slot (N for none);
 * grid [N, C, G, G, G] with element strides. loss (nullable) = the edge loss (0 without pairs); grad_grid (nullable,
 * needs grad_loss) gets d loss / d grid * (*grad_loss) added with float atomics. */
size_t neat_edge_loss_direct_workspace_bytes(int64_t num_active, int32_t grid_size);
int neat_edge_loss_direct(neat_stream_t stream, const float* node_pos_min, const float* node_pos_max,
                          const uint8_t* node_flags, int32_t num_nodes, const int64_t* active_node_ids,
                          int64_t num_active, const int64_t* node_slot, const float* grid, int64_t N, int32_t C,
                          int32_t G, const int64_t* strides, float epsilon, uint64_t seed, double weight, double* loss,
                          const float* grad_loss, float* grad_grid, void* workspace, size_t workspace_bytes);

/* ---- camera poses of the ray generation (FP64) ----------------------------------------------------------------------------
 * Quaternions (x, y, z, w) per pose, [num_poses, 4]; points [num_points, 3]; index[i] = pose of point i (int32, as
 * the reference's image ids).
 * neat_rotate_points_fwd: out = q[index] * p (RotatePointGPUk, CameraPose.cu:20-34).
 * neat_rotate_points_bwd: grad_points = R^T grad_out (nullable); grad_tangent[pose] += (R p) x grad_out, the gradient
 * with respect to a left tangent perturbation exp(t) * q at t = 0, accumulated with FP64 atomics (nullable;
 * RotatePointBackwardGPUk, CameraPose.cu:37-63).
 * neat_so3_apply_tangent: q <- exp(t) * q (renormalised), t <- 0 (ApplyTangent, CameraPose.cu:160-180). */
int neat_rotate_points_fwd(neat_stream_t stream, const double* points, const int32_t* index, const double* rotation,
                           int64_t num_points, int64_t num_poses, double* out);
int neat_rotate_points_bwd(neat_stream_t stream, const double* points, const int32_t* index, const double* rotation,
                           int64_t num_points, int64_t num_poses, const double* grad_out, double* grad_tangent,
                           double* grad_points);
int neat_so3_apply_tangent(neat_stream_t stream, double* tangent, double* rotation, int64_t num_poses);

/* ---- uniform volume samples of the volume export --------------------------------------------------------------------------
 * HyperTreeBaseImpl::UniformPhantomSamplesGPU (HyperTree.cu:547-597) with NodeIdForPositionGPU (:643-686): voxel i of
 * a size[0] x size[1] x size[2] lattice (dimension 0 fastest; the last with swap_xy) at c / (size - 1) * 2 - 1;
 * out_coord float [n, 3], out_index int64 [n] (= i), out_node int64 [n] (the active node containing it, else
 * first_active_node), out_weight float [n] (1, else 0); size is a host array of 3. */
int neat_uniform_volume_samples(neat_stream_t stream, const float* node_pos_min, const float* node_pos_max,
                                const uint8_t* node_flags, int32_t num_nodes, int64_t first_active_node,
                                const int32_t* size, int32_t swap_xy, float* out_coord, int64_t* out_index,
                                int64_t* out_node, float* out_weight);

/* ---- per-node loss of the evaluation pass --------------------------------------------------------------------------------
 * HierarchicalNeuralGeometry::AccumulateSampleLossPerNode (geometry.cpp:38-58): loss_sum[node] = sum over the node's
 * samples of per_ray_loss[ray] * weight, weight_sum[node] = sum of the weights, float [num_nodes], over the canonical
 * segment list of a uniformly sampled batch (neat_raysample_segments with the same jitter / seed as the emitted
 * samples; the weights are recomputed bit-identical to neat_raysample_emit's). Sums: FP32 per segment in j order,
 * then FP64 per node over its segments in canonical order (the reference's scatter_add follows atomic order). */
size_t neat_node_loss_workspace_bytes(int64_t num_segments, int32_t num_nodes);
int neat_node_loss_accumulate(neat_stream_t stream, const int64_t* ray_seg_offset, int64_t num_rays,
                              const int32_t* seg_node, const int32_t* seg_ray, const int32_t* seg_count,
                              const float* seg_tmin, const float* seg_tmax, int64_t num_segments, int32_t jitter,
                              uint64_t seed, const float* per_ray_loss, int32_t num_nodes, float* loss_sum,
                              float* weight_sum, void* workspace, size_t workspace_bytes);

/* ---- optimizer --------------------------------------------------------------------------------------------------------
 * One Adam step over n FP32 parameters in place (replaces torch::optim::Adam::step for the grid and decoder parameter
 * groups of HierarchicalNeuralGeometry::OptimizerStep, geometry.h:66-88; groups geometry_ex_ex.cpp:40-49,
 * geometry.cpp:90-97). step counts from 1; weight_decay adds weight_decay * param to the gradient as torch does.
 * All four buffers 16-byte aligned. */
int neat_adam_step(neat_stream_t stream, float* param, const float* grad, float* exp_avg, float* exp_avg_sq,
                   int64_t n, double lr, double beta1, double beta2, double eps, double weight_decay, int64_t step);

#ifdef __cplusplus
}
#endif

#endif /* NEAT_AMD_H */
