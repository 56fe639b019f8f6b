// Tensor bundles passed across the projector path (HelperStructs.h:92-234 in the reference): RayList and
// SampleList with the reference's field names, dtypes and shapes. SampleList additionally carries the canonical
// segment list the samples were emitted from, which the fused projector consumes instead of the per-sample columns.
#pragma once

#include <torch/torch.h>

struct RayList
{
    // float [num_rays, D]
    torch::Tensor origin;
    // float [num_rays, D]
    torch::Tensor direction;

    RayList() {}
    RayList(torch::Tensor o, torch::Tensor d) : origin(o), direction(d) {}

    RayList SubSample(torch::Tensor index)
    {
        return RayList(torch::index_select(origin, 0, index), torch::index_select(direction, 0, index));
    }
    void to(torch::Device device)
    {
        origin    = origin.to(device);
        direction = direction.to(device);
    }
    int size() const { return origin.size(0); }
    int Dim() const { return origin.size(1); }
};

// Canonical segments of a ray batch: rays in input order, segments of a ray sorted by (tmin, node id).
struct SegmentList
{
    torch::Tensor ray_seg_offset;     // long [num_rays + 1]
    torch::Tensor ray_sample_offset;  // long [num_rays + 1]
    torch::Tensor node_id;            // int [num_segments]
    torch::Tensor ray;                // int [num_segments]
    torch::Tensor count;              // int [num_segments]
    torch::Tensor tmin;               // float [num_segments]
    torch::Tensor tmax;               // float [num_segments]
    torch::Tensor sample_offset;      // long [num_segments]
    torch::Tensor ray_origin;         // the rays the segments belong to (FP32, contiguous, detached)
    torch::Tensor ray_direction;
    torch::Tensor ray_origin_in;      // the caller's ray tensors themselves: gradients flow to these
    torch::Tensor ray_direction_in;
    int64_t num_samples = 0;
    bool jitter         = false;
    uint64_t seed       = 0;

    bool defined() const { return ray_seg_offset.defined(); }
    int64_t size() const { return node_id.defined() ? node_id.size(0) : 0; }
};

struct SampleList
{
    int max_samples_per_ray = 0;
    // float [num_samples, D]
    torch::Tensor global_coordinate;
    // float [num_samples]
    torch::Tensor ray_t;
    // float [num_samples]
    torch::Tensor weight;
    // long [num_samples]
    torch::Tensor ray_index;
    // long [num_samples]
    torch::Tensor local_index_in_ray;
    // long [num_samples]
    torch::Tensor node_id;

    SegmentList segments;

    int size() const
    {
        if (global_coordinate.defined()) return global_coordinate.size(0);
        return (int)segments.num_samples;
    }

    SampleList Slice(int start, int end)
    {
        SampleList r;
        r.global_coordinate  = global_coordinate.slice(0, start, end);
        r.weight             = weight.slice(0, start, end);
        r.ray_index          = ray_index.slice(0, start, end);
        r.ray_t              = ray_t.slice(0, start, end);
        r.local_index_in_ray = local_index_in_ray.slice(0, start, end);
        r.node_id            = node_id.slice(0, start, end);
        return r;
    }
};
