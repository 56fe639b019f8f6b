"""CPU restatement of the reference's ray generation (SceneBase::GetRays, src/data/SceneBase.cpp:171-193) and its
camera modules (src/modules/CameraPose.h:20-52, src/modules/CameraPose.cu:20-275) in float64 -- TEST INFRASTRUCTURE
ONLY: tests/ use it as the checker of neat_amd/csrc/kernels/camera.hip and of the module layer's GetRays; the product
never imports it.

  rotate        q * p through the rotation matrix of the unit quaternion (x, y, z, w). Pinned: scipy's Rotation.
  rotate_bwd    R^T g and, per pose, (R p) x g, the gradient w.r.t. a left tangent perturbation exp(t) q at t = 0.
                Pinned: central finite differences of exp(t) R p and of R p.
  apply_tangent q <- exp(t) q. Pinned: scipy's Rotation.from_rotvec composed on the left.
  get_rays      UVToPixel, pinhole Unproject, normalise, rotate; origin = translation. Pinned: closed forms (the
                principal point's ray is the camera's optical axis R e_z; a ray's direction projects back onto its
                pixel).
"""
import numpy as np


def quat_matrix(q):
    """[N, 4] (x, y, z, w) -> [N, 3, 3]"""
    x, y, z, w = (q[:, i] for i in range(4))
    R = np.empty((q.shape[0], 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def rotate(rotation, points, index):
    R = quat_matrix(np.asarray(rotation, np.float64))[np.asarray(index)]
    return np.einsum("nij,nj->ni", R, np.asarray(points, np.float64))


def rotate_bwd(rotation, points, index, grad):
    index = np.asarray(index)
    R = quat_matrix(np.asarray(rotation, np.float64))[index]
    rp = np.einsum("nij,nj->ni", R, points)
    gp = np.einsum("nji,nj->ni", R, grad)
    gt = np.zeros((rotation.shape[0], 3))
    np.add.at(gt, index, np.cross(rp, grad))
    return gt, gp


def so3_exp(t):
    """[N, 3] rotation vectors -> [N, 4] unit quaternions (x, y, z, w)"""
    t = np.asarray(t, np.float64)
    th = np.linalg.norm(t, axis=1, keepdims=True)
    s = np.where(th < 1e-10, 0.5 - th * th / 48, np.sin(0.5 * th) / np.where(th == 0, 1, th))
    c = np.where(th < 1e-10, 1 - th * th / 8, np.cos(0.5 * th))
    return np.concatenate([s * t, c], axis=1)


def qmul(a, b):
    ax, ay, az, aw = (a[:, i] for i in range(4))
    bx, by, bz, bw = (b[:, i] for i in range(4))
    return np.stack([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz], axis=1)


def apply_tangent(tangent, rotation):
    q = qmul(so3_exp(tangent), rotation)
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def uv_to_pixel(uv, w, h, align_corners):
    uv = np.asarray(uv, np.float32)
    if align_corners:
        return np.stack([uv[:, 0] * np.float32(w - 1), uv[:, 1] * np.float32(h - 1)], axis=1)
    return np.stack([uv[:, 0] * np.float32(w) - np.float32(0.5), uv[:, 1] * np.float32(h) - np.float32(0.5)], axis=1)


def get_rays(rotation, translation, intrinsics, w, h, uv, image_id, camera_id, align_corners=False):
    px = uv_to_pixel(uv, w, h, align_corners).astype(np.float64)
    K = np.asarray(intrinsics, np.float64)[np.asarray(camera_id)]
    c = (px - K[:, 2:4]) / K[:, 0:2]
    un = np.concatenate([c, np.ones((c.shape[0], 1))], axis=1)
    un = un / np.linalg.norm(un, axis=1, keepdims=True)
    d = rotate(rotation, un, image_id)
    o = np.asarray(translation, np.float64)[np.asarray(image_id)]
    return o.astype(np.float32), d.astype(np.float32)
