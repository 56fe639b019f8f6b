#include "CameraPose.h"

#include <torch/csrc/autograd/custom_function.h>

#include "neat_host.h"

using namespace neat_host;
using torch::autograd::AutogradContext;

torch::Tensor UVToPixel(torch::Tensor uv, int w, int h, bool align_corners)
{
    auto u = uv.slice(1, 0, 1);
    auto v = uv.slice(1, 1, 2);
    if (align_corners)
    {
        u = u * (w - 1);
        v = v * (h - 1);
    }
    else
    {
        u = u * w - 0.5f;
        v = v * h - 0.5f;
    }
    return torch::cat({u, v}, 1);
}

torch::Tensor PixelToUV(torch::Tensor px, int w, int h, bool align_corners)
{
    auto x = px.slice(1, 0, 1);
    auto y = px.slice(1, 1, 2);
    if (align_corners)
    {
        x = x / (w - 1);
        y = y / (h - 1);
    }
    else
    {
        x = (x + 0.5f) / w;
        y = (y + 0.5f) / h;
    }
    return torch::cat({x, y}, 1);
}

namespace
{
// RotatePoint (CameraPose.cu:65-127) on neat_rotate_points_fwd / _bwd
struct RotatePointFn : public torch::autograd::Function<RotatePointFn>
{
    static torch::Tensor forward(AutogradContext* ctx, torch::Tensor rotation, torch::Tensor rotation_tangent,
                                 torch::Tensor points, torch::Tensor index)
    {
        check_cuda(points, "points");
        TORCH_CHECK(points.scalar_type() == torch::kFloat64 && points.dim() == 2 && points.size(1) == 3,
                    "RotatePoint: points must be FP64 [N, 3]");
        auto p   = points.contiguous();
        auto idx = index.to(torch::kInt32).contiguous();
        auto rot = rotation.contiguous();
        auto out = torch::empty_like(p);
        NEAT_CALL(neat_rotate_points_fwd(current_stream(), dptr<double>(p), dptr<int32_t>(idx), dptr<double>(rot),
                                         p.size(0), rot.size(0), dptr<double>(out)));
        ctx->save_for_backward({rot, p, idx});
        return out;
    }

    static torch::autograd::variable_list backward(AutogradContext* ctx, torch::autograd::variable_list grad_out)
    {
        auto sv  = ctx->get_saved_variables();
        auto rot = sv[0], p = sv[1], idx = sv[2];
        auto g   = grad_out[0].to(torch::kFloat64).contiguous();
        auto gt  = torch::zeros({rot.size(0), 3}, rot.options());
        auto gp  = torch::empty_like(p);
        NEAT_CALL(neat_rotate_points_bwd(current_stream(), dptr<double>(p), dptr<int32_t>(idx), dptr<double>(rot),
                                         p.size(0), rot.size(0), dptr<double>(g), dptr<double>(gt),
                                         dptr<double>(gp)));
        return {torch::Tensor(), gt, gp, torch::Tensor()};
    }
};
}  // namespace

CameraPoseModuleImpl::CameraPoseModuleImpl(torch::Tensor rotation_, torch::Tensor translation_)
{
    TORCH_CHECK(rotation_.dim() == 2 && rotation_.size(1) == 4, "CameraPoseModule: rotation must be [N, 4]");
    TORCH_CHECK(translation_.dim() == 2 && translation_.size(1) == 3 && translation_.size(0) == rotation_.size(0),
                "CameraPoseModule: translation must be [N, 3]");
    const int64_t N  = rotation_.size(0);
    auto o           = rotation_.options().dtype(torch::kFloat64);
    rotation_tangent = register_parameter("rotation_tangent", torch::zeros({N, 3}, o));
    translation      = register_parameter("translation", translation_.to(torch::kFloat64).clone());
    rotation         = register_buffer("rotation", rotation_.to(torch::kFloat64).clone());
}

torch::Tensor CameraPoseModuleImpl::RotatePoint(torch::Tensor p, torch::Tensor index)
{
    TORCH_CHECK(p.device() == rotation.device(), "RotatePoint: points and poses on different devices");
    return RotatePointFn::apply(rotation, rotation_tangent, p, index);
}

void CameraPoseModuleImpl::ApplyTangent()
{
    check_cuda(rotation, "the poses");
    torch::NoGradGuard ng;
    TORCH_CHECK(rotation.is_contiguous() && rotation_tangent.is_contiguous(), "ApplyTangent: contiguous poses");
    NEAT_CALL(neat_so3_apply_tangent(current_stream(), dptr<double>(rotation_tangent), dptr<double>(rotation),
                                     rotation.size(0)));
}

CameraModelModuleImpl::CameraModelModuleImpl(int h_, int w_, torch::Tensor intrinsics_) : h(h_), w(w_)
{
    TORCH_CHECK(intrinsics_.dim() == 2 && intrinsics_.size(1) == 5, "CameraModelModule: intrinsics must be [N, 5]");
    intrinsics = register_parameter("intrinsics", intrinsics_.to(torch::kFloat64).clone());
}

torch::Tensor CameraModelModuleImpl::Unproject(torch::Tensor camera_index, torch::Tensor pixel_coords,
                                               torch::Tensor depth)
{
    depth        = depth.unsqueeze(1);
    auto cameras = torch::index_select(intrinsics, 0, camera_index.to(torch::kLong));
    auto fxfy    = cameras.slice(1, 0, 2);
    auto cxcy    = cameras.slice(1, 2, 4);
    auto c       = (pixel_coords - cxcy) / fxfy;
    c            = torch::cat({c, torch::ones_like(depth)}, 1);
    return c * depth;
}

RayList GetRays(CameraPoseModule pose, CameraModelModule camera, torch::Tensor uv, torch::Tensor image_id,
                torch::Tensor camera_id, bool uv_align_corners)
{
    auto px     = UVToPixel(uv, camera->w, camera->h, uv_align_corners);
    auto unproj = camera->Unproject(camera_id, px, torch::ones({uv.size(0)}, torch::TensorOptions(uv.device())));
    unproj      = unproj / torch::norm(unproj, 2, 1, true);
    auto dir    = pose->RotatePoint(unproj, image_id);
    RayList r;
    r.origin    = torch::index_select(pose->translation, 0, image_id.to(torch::kLong)).to(torch::kFloat32);
    r.direction = dir.to(torch::kFloat32);
    return r;
}
