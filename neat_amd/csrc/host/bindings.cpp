// Python bindings of the libtorch host layer (used by the tests and the benchmark; C++ callers link the layer
// directly).
#include <torch/extension.h>
#include <torch/python.h>

#include "CameraPose.h"
#include "geometry.h"

namespace py = pybind11;

namespace
{
// nn::Module plumbing for plain pybind11 classes: parameters, buffers, state (names as registered in C++).
template <typename M>
void module_methods(py::class_<M, std::shared_ptr<M>>& c)
{
    c.def("parameters", [](M& m) { return m.parameters(); })
        .def("named_parameters",
             [](M& m)
             {
                 std::vector<std::pair<std::string, torch::Tensor>> out;
                 for (auto& kv : m.named_parameters()) out.emplace_back(kv.key(), kv.value());
                 return out;
             })
        .def("named_buffers",
             [](M& m)
             {
                 std::vector<std::pair<std::string, torch::Tensor>> out;
                 for (auto& kv : m.named_buffers()) out.emplace_back(kv.key(), kv.value());
                 return out;
             })
        .def("to", [](M& m, const std::string& dev) { m.to(torch::Device(dev)); })
        .def("train", [](M& m, bool on) { m.train(on); })
        .def("zero_grad", [](M& m) { m.zero_grad(); });
}
}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m)
{
    m.doc() = "NeAT projector path (HyperTreeBase, IndirectGridSample3D, FCBlock, GeometryExEx) on MI355X kernels";
// the build passes the source hash as a bare token (id + 16 hex digits), stringified here
#ifndef NEAT_HOST_BUILD_ID
#define NEAT_HOST_BUILD_ID idunknown
#endif
#define NEAT_STR2_(x) #x
#define NEAT_STR_(x) NEAT_STR2_(x)
    // hash of the sources this module was built from (neat_amd/buildinfo.py)
    m.def("build_id", []() { return std::string(NEAT_STR_(NEAT_HOST_BUILD_ID) + 2); });

    py::class_<RayList>(m, "RayList")
        .def(py::init<>())
        .def(py::init<torch::Tensor, torch::Tensor>())
        .def_readwrite("origin", &RayList::origin)
        .def_readwrite("direction", &RayList::direction)
        .def("size", &RayList::size);

    py::class_<SegmentList>(m, "SegmentList")
        .def_readonly("ray_seg_offset", &SegmentList::ray_seg_offset)
        .def_readonly("ray_sample_offset", &SegmentList::ray_sample_offset)
        .def_readonly("node_id", &SegmentList::node_id)
        .def_readonly("ray", &SegmentList::ray)
        .def_readonly("count", &SegmentList::count)
        .def_readonly("tmin", &SegmentList::tmin)
        .def_readonly("tmax", &SegmentList::tmax)
        .def_readonly("sample_offset", &SegmentList::sample_offset)
        .def_readonly("num_samples", &SegmentList::num_samples)
        .def_readonly("jitter", &SegmentList::jitter)
        .def_readonly("seed", &SegmentList::seed)
        .def("size", &SegmentList::size);

    py::class_<SampleList>(m, "SampleList")
        .def(py::init<>())
        .def_readwrite("max_samples_per_ray", &SampleList::max_samples_per_ray)
        .def_readwrite("global_coordinate", &SampleList::global_coordinate)
        .def_readwrite("ray_t", &SampleList::ray_t)
        .def_readwrite("weight", &SampleList::weight)
        .def_readwrite("ray_index", &SampleList::ray_index)
        .def_readwrite("local_index_in_ray", &SampleList::local_index_in_ray)
        .def_readwrite("node_id", &SampleList::node_id)
        .def_readwrite("segments", &SampleList::segments)
        .def("size", &SampleList::size);

    py::class_<HyperTreeBaseImpl, std::shared_ptr<HyperTreeBaseImpl>> tree_cls(m, "HyperTreeBase");
    module_methods(tree_cls);
    tree_cls.def(py::init<int, int>())
        .def("SetActive", &HyperTreeBaseImpl::SetActive)
        .def("UpdateActive", &HyperTreeBaseImpl::UpdateActive)
        .def("CreateSamplesForRays", &HyperTreeBaseImpl::CreateSamplesForRays)
        .def("CreateSegmentsForRays", &HyperTreeBaseImpl::CreateSegmentsForRays)
        .def("ComputeLocalSamples", &HyperTreeBaseImpl::ComputeLocalSamples)
        .def("GlobalNodeIdToLocalActiveId", &HyperTreeBaseImpl::GlobalNodeIdToLocalActiveId)
        .def("UniformPhantomSamplesGPU", &HyperTreeBaseImpl::UniformPhantomSamplesGPU, py::arg("size"),
             py::arg("swap_xy") = false)
        .def("NodeNeighborSamples", &HyperTreeBaseImpl::NodeNeighborSamples)
        .def("SetErrorForActiveNodes", &HyperTreeBaseImpl::SetErrorForActiveNodes, py::arg("error"),
             py::arg("strategy") = "override")
        .def_readonly("node_error", &HyperTreeBaseImpl::node_error)
        .def("ActiveNodeTensor", &HyperTreeBaseImpl::ActiveNodeTensor)
        .def("InactiveNodeIds", &HyperTreeBaseImpl::InactiveNodeIds)
        .def("NumNodes", &HyperTreeBaseImpl::NumNodes)
        .def("NumActiveNodes", &HyperTreeBaseImpl::NumActiveNodes)
        .def("MaxDepth", &HyperTreeBaseImpl::MaxDepth)
        .def_readwrite("node_active", &HyperTreeBaseImpl::node_active)
        .def_readonly("node_parent", &HyperTreeBaseImpl::node_parent)
        .def_readonly("node_children", &HyperTreeBaseImpl::node_children)
        .def_readonly("node_position_min", &HyperTreeBaseImpl::node_position_min)
        .def_readonly("node_position_max", &HyperTreeBaseImpl::node_position_max)
        .def_readonly("node_scale", &HyperTreeBaseImpl::node_scale)
        .def_readonly("node_depth", &HyperTreeBaseImpl::node_depth)
        .def_readonly("node_active_prefix_sum", &HyperTreeBaseImpl::node_active_prefix_sum)
        .def_readonly("active_node_ids", &HyperTreeBaseImpl::active_node_ids)
        .def_readonly("node_flags", &HyperTreeBaseImpl::node_flags);

    m.def("IndirectGridSample3D", &IndirectGridSample3D, py::arg("multi_grid"), py::arg("index"), py::arg("uv"));
    m.def("IntegrateSamples", &IntegrateSamples, py::arg("sample_values"), py::arg("integration_weight"),
          py::arg("ray_index"), py::arg("num_channels"), py::arg("num_rays"),
          py::arg("ray_sample_offset") = torch::Tensor());

    py::class_<FCBlockImpl, std::shared_ptr<FCBlockImpl>> fc_cls(m, "FCBlock");
    module_methods(fc_cls);
    fc_cls.def(py::init<int, int, int, float, bool, std::string>(), py::arg("in_features"), py::arg("out_features"),
             py::arg("num_hidden_layers"), py::arg("hidden_features"), py::arg("outermost_linear") = true,
             py::arg("non_linearity") = "relu")
        .def("forward", &FCBlockImpl::forward)
        .def("forward_with_last", &FCBlockImpl::forward_with_last)
        .def("packed_parameters", &FCBlockImpl::packed_parameters)
        .def("load_packed", &FCBlockImpl::load_packed);

    py::class_<ExplicitFeatureGridImpl, std::shared_ptr<ExplicitFeatureGridImpl>> grid_cls(m, "ExplicitFeatureGrid");
    module_methods(grid_cls);
    grid_cls.def(py::init<int, int, std::vector<long>, std::string>())
        .def_readonly("grid_data", &ExplicitFeatureGridImpl::grid_data);

    m.def("SaveCheckpoint", &SaveCheckpoint, py::arg("geometry"), py::arg("tree"), py::arg("prefix"));
    m.def("LoadCheckpoint", &LoadCheckpoint, py::arg("geometry"), py::arg("tree"), py::arg("prefix"));
    m.def("UVToPixel", &UVToPixel, py::arg("uv"), py::arg("w"), py::arg("h"), py::arg("align_corners"));
    m.def("PixelToUV", &PixelToUV, py::arg("px"), py::arg("w"), py::arg("h"), py::arg("align_corners"));
    py::class_<CameraPoseModuleImpl, std::shared_ptr<CameraPoseModuleImpl>> pose_cls(m, "CameraPoseModule");
    module_methods(pose_cls);
    pose_cls.def(py::init<torch::Tensor, torch::Tensor>(), py::arg("rotation"), py::arg("translation"))
        .def("RotatePoint", &CameraPoseModuleImpl::RotatePoint)
        .def("ApplyTangent", &CameraPoseModuleImpl::ApplyTangent)
        .def_readonly("rotation", &CameraPoseModuleImpl::rotation)
        .def_readonly("rotation_tangent", &CameraPoseModuleImpl::rotation_tangent)
        .def_readonly("translation", &CameraPoseModuleImpl::translation);
    py::class_<CameraModelModuleImpl, std::shared_ptr<CameraModelModuleImpl>> cam_cls(m, "CameraModelModule");
    module_methods(cam_cls);
    cam_cls.def(py::init<int, int, torch::Tensor>(), py::arg("h"), py::arg("w"), py::arg("intrinsics"))
        .def("Unproject", &CameraModelModuleImpl::Unproject)
        .def_readonly("intrinsics", &CameraModelModuleImpl::intrinsics)
        .def_readonly("h", &CameraModelModuleImpl::h)
        .def_readonly("w", &CameraModelModuleImpl::w);
    m.def(
        "GetRays",
        [](std::shared_ptr<CameraPoseModuleImpl> pose, std::shared_ptr<CameraModelModuleImpl> cam, torch::Tensor uv,
           torch::Tensor image_id, torch::Tensor camera_id, bool align)
        { return GetRays(CameraPoseModule(pose), CameraModelModule(cam), uv, image_id, camera_id, align); },
        py::arg("pose"), py::arg("camera"), py::arg("uv"), py::arg("image_id"), py::arg("camera_id"),
        py::arg("uv_align_corners") = false);

    py::class_<Netparams>(m, "Netparams")
        .def(py::init<>())
        .def_readwrite("grid_size", &Netparams::grid_size)
        .def_readwrite("grid_features", &Netparams::grid_features)
        .def_readwrite("last_activation_function", &Netparams::last_activation_function)
        .def_readwrite("softplus_beta", &Netparams::softplus_beta)
        .def_readwrite("decoder_activation", &Netparams::decoder_activation)
        .def_readwrite("decoder_hidden_layers", &Netparams::decoder_hidden_layers)
        .def_readwrite("decoder_hidden_features", &Netparams::decoder_hidden_features)
        .def_readwrite("decoder_lr", &Netparams::decoder_lr);
    py::class_<TreeOptimizerParams>(m, "TreeOptimizerParams")
        .def(py::init<>())
        .def_readwrite("max_active_nodes", &TreeOptimizerParams::max_active_nodes);
    py::class_<OctreeParams>(m, "OctreeParams")
        .def(py::init<>())
        .def_readwrite("max_samples_per_node", &OctreeParams::max_samples_per_node)
        .def_readwrite("tree_depth", &OctreeParams::tree_depth)
        .def_readwrite("tree_optimizer_params", &OctreeParams::tree_optimizer_params);
    py::class_<MyTrainParams>(m, "MyTrainParams")
        .def(py::init<>())
        .def_readwrite("grid_init", &MyTrainParams::grid_init)
        .def_readwrite("lr_exex_grid_adam", &MyTrainParams::lr_exex_grid_adam)
        .def_readwrite("use_fused_projector", &MyTrainParams::use_fused_projector)
        .def_readwrite("compensated_sum", &MyTrainParams::compensated_sum)
        .def_readwrite("projector_precision", &MyTrainParams::projector_precision)
        .def_readwrite("loss_tv", &MyTrainParams::loss_tv)
        .def_readwrite("loss_edge", &MyTrainParams::loss_edge);
    py::class_<CombinedParams, std::shared_ptr<CombinedParams>>(m, "CombinedParams")
        .def(py::init<>())
        .def_readwrite("net_params", &CombinedParams::net_params)
        .def_readwrite("octree_params", &CombinedParams::octree_params)
        .def_readwrite("train_params", &CombinedParams::train_params);

    py::class_<GeometryExEx, std::shared_ptr<GeometryExEx>> geo_cls(m, "GeometryExEx");
    module_methods(geo_cls);
    geo_cls.def(py::init([](int num_channels, int D, std::shared_ptr<HyperTreeBaseImpl> tree,
                         std::shared_ptr<CombinedParams> params)
                      { return std::make_shared<GeometryExEx>(num_channels, D, HyperTreeBase(tree), params); }))
        .def("ComputeImage", &GeometryExEx::ComputeImage)
        .def("SampleVolumeIndirect", &GeometryExEx::SampleVolumeIndirect)
        .def("DecodeFeatures", &GeometryExEx::DecodeFeatures)
        .def("IntegrateSamplesXRay", &GeometryExEx::IntegrateSamplesXRay)
        .def("FusedUnsupportedReason", &GeometryExEx::FusedUnsupportedReason)
        .def("CreateGeometryOptimizer", &GeometryExEx::CreateGeometryOptimizer)
        .def("OptimizerStep", &GeometryExEx::OptimizerStep)
        .def("UpdateLearningRate", &GeometryExEx::UpdateLearningRate)
        .def("VolumeRegularizer", &GeometryExEx::VolumeRegularizer, py::arg("seed") = -1)
        .def_readwrite("fused_edge_loss", &GeometryExEx::fused_edge_loss)
        .def("AccumulateSampleLossPerNode", &GeometryExEx::AccumulateSampleLossPerNode)
        .def("UniformSampledVolume", &GeometryExEx::UniformSampledVolume, py::arg("shape"), py::arg("num_channels"))
        .def_property_readonly("decoder", [](GeometryExEx& g) { return g.decoder.ptr(); })
        .def_property_readonly("grid", [](GeometryExEx& g) { return g.explicit_grid_generator->grid_data; });
}
