// Neural geometry of the projector path (geometry.h:21-249 and geometry_ex_ex.h:64-84 in the reference):
// NeuralGeometry::IntegrateSamplesXRay, HierarchicalNeuralGeometry::ComputeImage / DecodeFeatures and
// GeometryExEx::SampleVolumeIndirect, on the C ABI. ComputeImage runs the fused projector when the SampleList carries
// its canonical segments and the decoder / grid fall in the fused kernels' coverage, the unfused kernel chain
// otherwise.
#pragma once

#include <memory>

#include "HyperTree.h"
#include "ImplicitNet.h"
#include "IndirectGridSample3D.h"
#include "Settings.h"
#include "optim.h"

// out[c, ray_index[i]] += sample_values[i, c] * integration_weight[i]; deterministic when ray_sample_offset
// (CSR offsets of ray-sorted samples) is given.
torch::Tensor IntegrateSamples(torch::Tensor sample_values, torch::Tensor integration_weight, torch::Tensor ray_index,
                               int num_channels, int num_rays, torch::Tensor ray_sample_offset = {});

class NeuralGeometry : public torch::nn::Module
{
   public:
    NeuralGeometry(int num_channels, int D, std::shared_ptr<CombinedParams> params)
        : num_channels(num_channels), D(D), params(params)
    {
    }

    // sample_values [num_samples, num_channels] or [groups, group_size, num_channels], integration_weight of the same
    // rank, ray_index [num_samples] -> [num_channels, num_rays] (geometry.h:112-113, geometry.cpp:393-424).
    torch::Tensor IntegrateSamplesXRay(torch::Tensor sample_values, torch::Tensor integration_weight,
                                       torch::Tensor ray_index, int num_channels, int num_rays);

    int num_channels;
    int D;
    std::shared_ptr<CombinedParams> params;
};

class HierarchicalNeuralGeometry : public NeuralGeometry
{
   public:
    HierarchicalNeuralGeometry(int num_channels, int D, std::shared_ptr<CombinedParams> params, HyperTreeBase tree);

    // [num_channels, num_pixels] line integrals of the decoded field over the samples (geometry.cpp:347-389).
    torch::Tensor ComputeImage(SampleList all_samples, int num_channels, int num_pixels);

    // Input: global_coordinate [num_samples, D], node_id [num_samples] -> density [num_samples, num_channels].
    virtual torch::Tensor SampleVolumeIndirect(torch::Tensor global_coordinate, torch::Tensor node_id) { return {}; }

    // decoder + last activation (geometry.cpp:102-163), one fused kernel.
    torch::Tensor DecodeFeatures(torch::Tensor neural_features);

    // The fused projector for a SampleList with segments, or an undefined tensor when not covered.
    virtual torch::Tensor ComputeImageFused(const SampleList& samples, int num_pixels) { return {}; }

    // Per-node sums of the samples' per-ray loss times weight, and of the weights (geometry.cpp:38-58), float
    // [num_nodes] each; per_ray_loss is [channels, num_rays] (averaged over channels). Needs the samples' segment
    // list (CreateSamplesForRays keeps it in SampleList::segments).
    // the field on a uniform lattice over [-1, 1]^3 (geometry.cpp:166-249): volume [num_channels, shape...] (the
    // decoded value where the voxel lies in an active node, 0 elsewhere), node id and validity [shape...] (int64),
    // on the CPU as there; evaluated in batches of 2^20 voxels on the GPU
    std::tuple<torch::Tensor, torch::Tensor, torch::Tensor> UniformSampledVolume(std::vector<long> shape,
                                                                                 int num_channels);
    std::pair<torch::Tensor, torch::Tensor> AccumulateSampleLossPerNode(const SampleList& combined_samples,
                                                                        torch::Tensor per_ray_loss);

    // Adam groups as the reference creates them (geometry.h:46-56, geometry.cpp:87-97): the decoder in its own
    // optimizer at net_params.decoder_lr; subclasses add their parameters to optimizer_adam.
    void CreateGeometryOptimizer();
    void ResetGeometryOptimizer() { CreateGeometryOptimizer(); }
    virtual void AddParametersToOptimizer();
    // step and zero the gradients of every optimizer (geometry.h:66-88)
    void OptimizerStep(int epoch_id);
    void UpdateLearningRate(double factor);

    HyperTreeBase tree = nullptr;
    FCBlock decoder    = nullptr;
    std::shared_ptr<NeatAdam> optimizer_adam, optimizer_decoder;
};

class GeometryExEx : public HierarchicalNeuralGeometry
{
   public:
    GeometryExEx(int num_channels, int D, HyperTreeBase tree, std::shared_ptr<CombinedParams> params);

    torch::Tensor SampleVolumeIndirect(torch::Tensor global_coordinate, torch::Tensor node_id) override;
    torch::Tensor ComputeImageFused(const SampleList& samples, int num_pixels) override;

    // why the fused projector does not cover this configuration ("" when it does)
    std::string FusedUnsupportedReason();

    // the grid group at train_params.lr_exex_grid_adam (geometry_ex_ex.cpp:40-49)
    void AddParametersToOptimizer() override;

    // total variation of the active grids times loss_tv, plus the edge loss along the faces of neighbouring active
    // nodes times loss_edge (geometry_ex_ex.cpp:86-160); differentiable with respect to the grid. seed < 0 draws the
    // per-pair factors' seed from torch's default generator (the reference draws torch::rand_like).
    torch::Tensor VolumeRegularizer(int64_t seed = -1);
    // true: the edge loss in one fused pass over the candidate face points; false: the reference's chain
    // (NodeNeighborSamples, ComputeLocalSamples, IndirectGridSample3D, edge loss). Same values either way.
    bool fused_edge_loss = true;

    ExplicitFeatureGrid explicit_grid_generator = nullptr;
};

// torch::save / torch::load of the geometry and the tree as prefix + "geometry.pth" / "tree.pth"
// (reconstruct.cpp:36-62)
void SaveCheckpoint(std::shared_ptr<GeometryExEx> geometry, std::shared_ptr<HyperTreeBaseImpl> tree,
                    const std::string& prefix);
void LoadCheckpoint(std::shared_ptr<GeometryExEx> geometry, std::shared_ptr<HyperTreeBaseImpl> tree,
                    const std::string& prefix);
