// Camera poses of the reference's ray generation (CameraPoseModuleImpl, CameraPose.cu:20-172), in FP64:
//
//   rotate forward  : out = q[index] * p, the unit quaternion rotation of each point (RotatePointGPUk, :20-34).
//   rotate backward : the point gradient R^T g, and the gradient with respect to a left tangent perturbation of the
//                     pose, exp(t) * q at t = 0 (the update ApplyTangent makes), which is (R p) x g summed per pose
//                     (RotatePointBackwardGPUk, :37-63; FP64 atomics per pose as there).
//   apply tangent   : q <- exp(t) * q, t <- 0 (ApplyTangent, :160-172).
//
// Quaternions are stored as (x, y, z, w), the coefficient order of the reference's rotation buffer.
#include "common.hpp"

namespace neat
{

struct Q
{
    double x, y, z, w;
};

__device__ __forceinline__ void cross(const double a[3], const double b[3], double out[3])
{
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}

// q * p for a unit quaternion: p + 2 w (v x p) + 2 v x (v x p), v = (x, y, z)
__device__ __forceinline__ void rotate(const Q& q, const double p[3], double out[3])
{
    const double v[3] = {q.x, q.y, q.z};
    double uv[3], uuv[3];
    cross(v, p, uv);
#pragma unroll
    for (int k = 0; k < 3; ++k) uv[k] += uv[k];
    cross(v, uv, uuv);
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = p[k] + q.w * uv[k] + uuv[k];
}

// R^T g = conj(q) * g
__device__ __forceinline__ void rotate_inverse(const Q& q, const double g[3], double out[3])
{
    rotate(Q{-q.x, -q.y, -q.z, q.w}, g, out);
}

__device__ __forceinline__ Q load_q(const double* rotation, int64_t i)
{
    return Q{rotation[4 * i], rotation[4 * i + 1], rotation[4 * i + 2], rotation[4 * i + 3]};
}

__global__ void __launch_bounds__(256) k_rotate_fwd(const double* points, const int32_t* index,
                                                    const double* rotation, int64_t n, double* out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double p[3] = {points[3 * i], points[3 * i + 1], points[3 * i + 2]};
    double r[3];
    rotate(load_q(rotation, index[i]), p, r);
#pragma unroll
    for (int k = 0; k < 3; ++k) out[3 * i + k] = r[k];
}

__global__ void __launch_bounds__(256) k_rotate_bwd(const double* points, const int32_t* index,
                                                    const double* rotation, int64_t n, const double* grad_out,
                                                    double* grad_tangent, double* grad_points)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t id  = index[i];
    const Q q         = load_q(rotation, id);
    const double p[3] = {points[3 * i], points[3 * i + 1], points[3 * i + 2]};
    const double g[3] = {grad_out[3 * i], grad_out[3 * i + 1], grad_out[3 * i + 2]};
    double r[3], gt[3], gp[3];
    rotate(q, p, r);
    cross(r, g, gt);  // d/dt of exp(t) R p = -[R p]x, transposed onto g: (R p) x g
    rotate_inverse(q, g, gp);
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        if (grad_tangent) atomicAdd(grad_tangent + 3 * id + k, gt[k]);
        if (grad_points) grad_points[3 * i + k] = gp[k];
    }
}

// exp(t) as a unit quaternion: (sin(|t|/2) t/|t|, cos(|t|/2)); near 0 the Taylor terms of sin(x/2)/x and cos(x/2)
__device__ __forceinline__ Q so3_exp(const double t[3])
{
    const double th2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
    const double th  = sqrt(th2);
    double s, c;
    if (th < 1e-10)
    {
        s = 0.5 - th2 / 48.0;
        c = 1.0 - th2 / 8.0;
    }
    else
    {
        s = sin(0.5 * th) / th;
        c = cos(0.5 * th);
    }
    return Q{s * t[0], s * t[1], s * t[2], c};
}

// a * b (Hamilton product)
__device__ __forceinline__ Q qmul(const Q& a, const Q& b)
{
    return Q{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
             a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}

__global__ void __launch_bounds__(128) k_apply_tangent(double* tangent, double* rotation, int64_t m)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const double t[3] = {tangent[3 * i], tangent[3 * i + 1], tangent[3 * i + 2]};
    Q q               = qmul(so3_exp(t), load_q(rotation, i));
    // keep it a unit quaternion, as the reference's SO3 multiplication does
    const double nrm = sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    rotation[4 * i]     = q.x / nrm;
    rotation[4 * i + 1] = q.y / nrm;
    rotation[4 * i + 2] = q.z / nrm;
    rotation[4 * i + 3] = q.w / nrm;
    tangent[3 * i] = tangent[3 * i + 1] = tangent[3 * i + 2] = 0.0;
}

}  // namespace neat

using namespace neat;

extern "C"
{
    int neat_rotate_points_fwd(neat_stream_t stream, const double* points, const int32_t* index,
                               const double* rotation, int64_t num_points, int64_t num_poses, double* out)
    {
        NEAT_REQUIRE(num_points >= 0 && num_poses >= 0, "neat_rotate_points_fwd: bad sizes");
        if (num_points == 0) return NEAT_OK;
        NEAT_REQUIRE(points && index && rotation && out, "neat_rotate_points_fwd: null pointer");
        k_rotate_fwd<<<grid_for(num_points, 256), 256, 0, as_stream(stream)>>>(points, index, rotation, num_points,
                                                                               out);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_rotate_points_bwd(neat_stream_t stream, const double* points, const int32_t* index,
                               const double* rotation, int64_t num_points, int64_t num_poses, const double* grad_out,
                               double* grad_tangent, double* grad_points)
    {
        NEAT_REQUIRE(num_points >= 0 && num_poses >= 0, "neat_rotate_points_bwd: bad sizes");
        if (num_points == 0) return NEAT_OK;
        NEAT_REQUIRE(points && index && rotation && grad_out, "neat_rotate_points_bwd: null pointer");
        k_rotate_bwd<<<grid_for(num_points, 256), 256, 0, as_stream(stream)>>>(points, index, rotation, num_points,
                                                                               grad_out, grad_tangent, grad_points);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }

    int neat_so3_apply_tangent(neat_stream_t stream, double* tangent, double* rotation, int64_t num_poses)
    {
        NEAT_REQUIRE(num_poses >= 0, "neat_so3_apply_tangent: bad sizes");
        if (num_poses == 0) return NEAT_OK;
        NEAT_REQUIRE(tangent && rotation, "neat_so3_apply_tangent: null pointer");
        k_apply_tangent<<<grid_for(num_poses, 128), 128, 0, as_stream(stream)>>>(tangent, rotation, num_poses);
        NEAT_LAUNCHED();
        return NEAT_OK;
    }
}
