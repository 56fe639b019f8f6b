"""Pins the CPU restatement of the ray generation (oracle/camera.py): the quaternion rotation and the tangent update
against scipy's Rotation, the rotation's gradients against central finite differences, GetRays against closed forms."""
import numpy as np
from scipy.spatial.transform import Rotation

from oracle import camera as CA


def random_poses(n, seed):
    rng = np.random.default_rng(seed)
    q = Rotation.random(n, random_state=seed).as_quat()  # (x, y, z, w)
    t = rng.normal(size=(n, 3))
    return q, t


def test_rotate_matches_scipy():
    q, _ = random_poses(5, 0)
    rng = np.random.default_rng(1)
    p = rng.normal(size=(40, 3))
    idx = rng.integers(0, 5, 40)
    want = np.stack([Rotation.from_quat(q[i]).apply(p[k]) for k, i in enumerate(idx)])
    assert np.allclose(CA.rotate(q, p, idx), want, rtol=0, atol=1e-14)


def test_apply_tangent_matches_scipy():
    q, _ = random_poses(6, 2)
    t = np.random.default_rng(3).normal(size=(6, 3)) * 0.3
    t[0] = 0.0
    t[1] = 1e-12
    got = CA.apply_tangent(t, q)
    want = (Rotation.from_rotvec(t) * Rotation.from_quat(q)).as_matrix()
    assert np.allclose(CA.quat_matrix(got), want, atol=1e-14)


def test_rotate_bwd_matches_finite_differences():
    q, _ = random_poses(3, 4)
    rng = np.random.default_rng(5)
    p = rng.normal(size=(12, 3))
    idx = rng.integers(0, 3, 12)
    g = rng.normal(size=(12, 3))
    gt, gp = CA.rotate_bwd(q, p, idx, g)

    def loss(qq, pp):
        return float((CA.rotate(qq, pp, idx) * g).sum())

    h = 1e-6
    fd_t = np.zeros((3, 3))
    for pose in range(3):
        for k in range(3):
            e = np.zeros((3, 3))
            e[pose, k] = h
            fd_t[pose, k] = (loss(CA.apply_tangent(e, q), p) - loss(CA.apply_tangent(-e, q), p)) / (2 * h)
    fd_p = np.zeros_like(p)
    for i in range(p.shape[0]):
        for k in range(3):
            e = np.zeros_like(p)
            e[i, k] = h
            fd_p[i, k] = (loss(q, p + e) - loss(q, p - e)) / (2 * h)
    assert np.allclose(gt, fd_t, atol=1e-7) and np.allclose(gp, fd_p, atol=1e-7)


def test_get_rays_closed_forms():
    q, t = random_poses(2, 6)
    w, h = 64, 48
    K = np.array([[50.0, 55.0, 31.5, 23.5, 0.0], [40.0, 40.0, 30.0, 20.0, 0.0]])
    # the principal point of camera 0 (pixel (31.5, 23.5)) with align_corners = False
    uv = np.array([[(31.5 + 0.5) / w, (23.5 + 0.5) / h], [0.1, 0.9], [0.7, 0.2]], np.float32)
    img = np.array([0, 1, 1])
    cam = np.array([0, 1, 0])
    o, d = CA.get_rays(q, t, K, w, h, uv, img, cam)
    assert np.allclose(o, t[img].astype(np.float32))
    assert np.allclose(d[0], CA.quat_matrix(q)[0][:, 2], atol=1e-6)  # the optical axis R e_z
    # each direction, rotated back into the camera and projected, lands on its pixel
    for k in range(3):
        dc = CA.quat_matrix(q)[img[k]].T @ d[k].astype(np.float64)
        px = K[cam[k], 0:2] * dc[:2] / dc[2] + K[cam[k], 2:4]
        want = uv[k].astype(np.float64) * np.array([w, h]) - 0.5
        assert np.allclose(px, want, atol=1e-4)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
