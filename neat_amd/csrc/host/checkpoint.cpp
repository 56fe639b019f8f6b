// Checkpoints as the reference trainer writes them (reconstruct.cpp:36-62): torch::save of the geometry module and of
// the tree module, one file each, under the reference's parameter and buffer names (grid_data, decoder.seq.*,
// node_* / active_node_ids).
//
// Loading reads every stored tensor into a temporary first, checks that each expected key is there and that its
// shape matches the live tensor's, and only then copies the values into the live tensors. Nothing changes when a
// check fails. Copying (rather than swapping the stored tensor in, as torch::load does) keeps each live tensor's
// layout: the grid stays channels-last-3d whatever layout the file holds (the reference saves it as plain NCDHW,
// ImplicitNet.h:79-119), and optimizer moments created before the load keep matching their parameters. Values are
// converted to the live tensor's dtype, so a tree whose node_active_id buffer is int32, as the reference registers
// it, loads too. The tree's derived state (active ids, prefix sums, traversal flags) is rebuilt from node_active.
#include <torch/serialize.h>

#include <algorithm>

#include "geometry.h"

namespace
{

struct PendingCopy
{
    torch::Tensor live;
    torch::Tensor stored;
};

// Collects (live, stored) pairs of module m and its children from archive ar; every parameter and buffer of the live
// module must be in the archive with the live shape, except the names in `any_shape` (buffers whose length is state,
// such as active_node_ids, rebuilt after the load).
void collect(torch::nn::Module& m, torch::serialize::InputArchive& ar, const std::string& path,
             const std::vector<std::string>& any_shape, std::vector<PendingCopy>& out)
{
    for (auto& kv : m.named_parameters(/*recurse=*/false))
    {
        torch::Tensor t;
        TORCH_CHECK(ar.try_read(kv.key(), t), "checkpoint: parameter '", path + kv.key(), "' is missing");
        TORCH_CHECK(t.sizes() == kv.value().sizes(), "checkpoint: parameter '", path + kv.key(), "' has shape ",
                    t.sizes(), ", the module holds ", kv.value().sizes());
        out.push_back({kv.value(), t});
    }
    for (auto& kv : m.named_buffers(/*recurse=*/false))
    {
        torch::Tensor t;
        TORCH_CHECK(ar.try_read(kv.key(), t, /*is_buffer=*/true), "checkpoint: buffer '", path + kv.key(),
                    "' is missing");
        const bool free_shape = std::find(any_shape.begin(), any_shape.end(), kv.key()) != any_shape.end();
        TORCH_CHECK(free_shape || t.sizes() == kv.value().sizes(), "checkpoint: buffer '", path + kv.key(),
                    "' has shape ", t.sizes(), ", the module holds ", kv.value().sizes());
        out.push_back({kv.value(), t});
    }
    for (auto& kv : m.named_children())
    {
        torch::serialize::InputArchive sub;
        TORCH_CHECK(ar.try_read(kv.key(), sub), "checkpoint: module '", path + kv.key(), "' is missing");
        collect(*kv.value(), sub, path + kv.key() + ".", any_shape, out);
    }
}

void apply(std::vector<PendingCopy>& copies)
{
    torch::NoGradGuard ng;
    for (auto& c : copies)
    {
        auto v = c.stored.to(c.live.device()).to(c.live.dtype());
        if (v.sizes() == c.live.sizes())
            c.live.copy_(v);
        else
            c.live.set_data(v.contiguous());  // a buffer whose length is state (see collect)
    }
}

}  // namespace

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
    std::vector<PendingCopy> tree_copies, geo_copies;
    {
        torch::serialize::InputArchive ar;
        ar.load_from(prefix + "tree.pth", dev);
        // every [num_nodes] buffer is checked against this tree's node count (its depth), so a tree of another
        // depth is refused before anything changes
        collect(*tree, ar, "", {"active_node_ids"}, tree_copies);
    }
    {
        torch::serialize::InputArchive ar;
        ar.load_from(prefix + "geometry.pth", dev);
        collect(*geometry, ar, "", {}, geo_copies);
    }
    // node_active and node_culled must describe a valid active set of this tree before anything is written
    {
        torch::Tensor active, culled;
        for (auto& c : tree_copies)
        {
            if (c.live.is_same(tree->node_active)) active = c.stored;
            if (c.live.is_same(tree->node_culled)) culled = c.stored;
        }
        TORCH_CHECK(active.defined() && culled.defined(), "checkpoint: tree buffers node_active / node_culled missing");
        auto a = active.to(torch::kCPU).to(torch::kInt32).contiguous();
        auto cu = culled.to(torch::kCPU).to(torch::kInt32);
        TORCH_CHECK((a * cu).sum().item<int64_t>() == 0, "checkpoint: a culled node is active in the stored tree");
        auto flags = torch::empty({a.size(0)}, torch::kUInt8);
        TORCH_CHECK(neat_tree_flags_host(a.data_ptr<int32_t>(), (int32_t)a.size(0), tree->MaxDepth(),
                                         flags.data_ptr<uint8_t>()) == NEAT_OK,
                    "checkpoint: node_active is not a valid active set of this tree: ", neat_last_error());
    }
    apply(tree_copies);
    apply(geo_copies);
    tree->UpdateActive();
}
