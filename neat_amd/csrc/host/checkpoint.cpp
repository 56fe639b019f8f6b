// Checkpoints as the reference trainer writes them (reconstruct.cpp:36-62): torch::save of the geometry module and of
// the tree module, one file each, under the reference's parameter and buffer names (grid_data, decoder.seq.*,
// node_* / active_node_ids). The tree loads buffer by buffer, each converted to the dtype this tree keeps, then
// rebuilds its derived state (active ids, prefix sums, traversal flags) from node_active, so a file whose
// node_active_id buffer is int32, as the reference registers it, loads too.
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
    {
        // buffer by buffer, each converted to the dtype this tree keeps (torch::load would refuse a buffer stored in
        // another dtype, such as the reference's int32 node_active_id)
        torch::serialize::InputArchive ar;
        ar.load_from(prefix + "tree.pth", dev);
        torch::NoGradGuard ng;
        for (auto& kv : tree->named_buffers(/*recurse=*/false))
        {
            torch::Tensor t;
            if (ar.try_read(kv.key(), t, /*is_buffer=*/true))
                kv.value().set_data(t.to(kv.value().dtype()).to(kv.value().device()));
        }
    }
    tree->UpdateActive();
    torch::nn::ModuleHolder<GeometryExEx> holder(geometry);
    torch::load(holder, prefix + "geometry.pth", dev);
}
