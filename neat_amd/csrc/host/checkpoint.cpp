// Checkpoints as the reference trainer writes them (reconstruct.cpp:36-62): torch::save of the geometry module and of
// the tree module, one file each, under the reference's parameter and buffer names (grid_data, decoder.seq.*,
// node_* / active_node_ids). Loading goes through torch::load onto the modules' device and then rebuilds the tree's
// derived state (active ids, prefix sums, traversal flags) from node_active, so a file whose node_active_id buffer
// is int32, as the reference registers it, loads too.
#include <torch/serialize.h>

#include "geometry.h"

void SaveCheckpoint(std::shared_ptr<GeometryExEx> geometry, std::shared_ptr<HyperTreeBaseImpl> tree,
                    const std::string& prefix)
{
    torch::nn::ModuleHolder<GeometryExEx> holder(geometry);
    torch::save(holder, prefix + "geometry.pth");
    HyperTreeBase t(tree);
    torch::save(t, prefix + "tree.pth");
}

void LoadCheckpoint(std::shared_ptr<GeometryExEx> geometry, std::shared_ptr<HyperTreeBaseImpl> tree,
                    const std::string& prefix)
{
    const auto dev = tree->node_position_min.device();
    HyperTreeBase t(tree);
    torch::load(t, prefix + "tree.pth", dev);
    tree->to(dev);
    tree->UpdateActive();
    torch::nn::ModuleHolder<GeometryExEx> holder(geometry);
    torch::load(holder, prefix + "geometry.pth", dev);
}
