// Ray generation of the reference (SceneBase::GetRays, SceneBase.cpp:171-193) with its camera modules
// (CameraPoseModuleImpl, CameraModelModuleImpl, CameraPose.h / CameraPose.cu): pinhole unprojection with torch ops
// in FP64, rotation of the unprojected directions by the per-image pose on the FP64 rotate kernels, origin = the
// pose's translation. Gradients reach the pose's rotation tangent and translation and the intrinsics, as there.
#pragma once

#include <torch/torch.h>

#include "HelperStructs.h"

// uv float [N, 2] in [0, 1] -> pixel coordinates (CameraPose.h:20-35)
torch::Tensor UVToPixel(torch::Tensor uv, int w, int h, bool align_corners);
// the inverse (CameraPose.h:37-52)
torch::Tensor PixelToUV(torch::Tensor px, int w, int h, bool align_corners);

class CameraPoseModuleImpl : public torch::nn::Module
{
   public:
    // rotation: unit quaternions (x, y, z, w) [N, 4], translation [N, 3]; both FP64 (CameraPose.cu:133-160)
    CameraPoseModuleImpl(torch::Tensor rotation, torch::Tensor translation);

    // q[index] * p for p FP64 [M, 3], index int32 [M]; differentiable w.r.t. p and rotation_tangent (the forward does
    // not read the tangent: it is kept at zero between optimizer steps and folded in by ApplyTangent)
    torch::Tensor RotatePoint(torch::Tensor p, torch::Tensor index);
    // rotation <- exp(rotation_tangent) * rotation, rotation_tangent <- 0
    void ApplyTangent();

    torch::Tensor rotation;          // buffer, FP64 [N, 4]
    torch::Tensor rotation_tangent;  // parameter, FP64 [N, 3]
    torch::Tensor translation;       // parameter, FP64 [N, 3]
};
TORCH_MODULE(CameraPoseModule);

class CameraModelModuleImpl : public torch::nn::Module
{
   public:
    // pinhole intrinsics FP64 [N, 5]: fx, fy, cx, cy, skew (CameraPose.cu:240-260)
    CameraModelModuleImpl(int h, int w, torch::Tensor intrinsics);
    // ((px - (cx, cy)) / (fx, fy), 1) * depth (CameraPose.cu:262-275)
    torch::Tensor Unproject(torch::Tensor camera_index, torch::Tensor pixel_coords, torch::Tensor depth);

    int h, w;
    torch::Tensor intrinsics;  // parameter
};
TORCH_MODULE(CameraModelModule);

// SceneBase::GetRays: uv float [N, 2], image_id / camera_id int32 or int64 [N] -> float rays
RayList GetRays(CameraPoseModule pose, CameraModelModule camera, torch::Tensor uv, torch::Tensor image_id,
                torch::Tensor camera_id, bool uv_align_corners);
