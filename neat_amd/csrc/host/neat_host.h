// libtorch host layer over the C ABI: the module API of the reference's projector path (src/structure,
// src/modules, src/geometry), re-implemented on libneat_amd's HIP kernels. Every compute call goes through
// include/neat_amd.h; this layer only allocates tensors, picks the current HIP stream and builds autograd graphs.
#pragma once

#include <torch/torch.h>
#include <c10/hip/HIPStream.h>

#include <string>

#include "neat_amd.h"

namespace neat_host
{

inline neat_stream_t current_stream() { return reinterpret_cast<neat_stream_t>(c10::hip::getCurrentHIPStream().stream()); }

#define NEAT_CALL(expr)                                                                            \
    do                                                                                             \
    {                                                                                              \
        int st_ = (expr);                                                                          \
        TORCH_CHECK(st_ == NEAT_OK, #expr, " failed with status ", st_, ": ", neat_last_error()); \
    } while (0)

inline int act_id(const std::string& name)
{
    if (name == "id" || name == "identity" || name == "none") return NEAT_ACT_ID;
    if (name == "relu") return NEAT_ACT_RELU;
    if (name == "silu") return NEAT_ACT_SILU;
    if (name == "softplus") return NEAT_ACT_SOFTPLUS;
    if (name == "abs") return NEAT_ACT_ABS;
    TORCH_CHECK(false, "unknown activation '", name, "' (expected id, relu, silu, softplus, abs)");
    return -1;
}

// projector arithmetic by name; -1 for an unknown name
inline int precision_id(const std::string& name)
{
    if (name == "fp32") return NEAT_PREC_FP32;
    if (name == "bf16") return NEAT_PREC_BF16;
    if (name == "fp32_valu") return NEAT_PREC_FP32_VALU;
    return -1;
}

template <typename T>
inline T* dptr(const torch::Tensor& t)
{
    return t.defined() ? t.data_ptr<T>() : nullptr;
}

inline void check_cuda(const torch::Tensor& t, const char* what)
{
    TORCH_CHECK(t.defined() && t.is_cuda(), what, " must be a defined GPU tensor (this build has no CPU path)");
}

}  // namespace neat_host
