"""MI355X-native differentiable forward projector of NeAT (Neural Adaptive Tomography).

Layers:
  libneat_amd.so  HIP kernels for gfx950 behind the C ABI in include/neat_amd.h (ray sampling, indirect trilinear
                  grid sample, decoder MLP, per-ray integration, fused projector forward / backward)
  _neat_host.so   libtorch module layer with the reference's API (HyperTreeBase, IndirectGridSample3D, FCBlock,
                  ExplicitFeatureGrid, GeometryExEx::ComputeImage), loaded through neat_amd.host.load()
  ops / pipeline  torch-tensor wrappers of the C ABI used by the tests and the benchmark
"""
__all__ = ["ops", "pipeline", "workloads", "host"]
