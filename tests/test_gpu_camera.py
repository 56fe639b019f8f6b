"""GPU checks of the ray generation (camera.hip and the module layer's CameraPoseModule / CameraModelModule / GetRays)
against oracle/camera.py: the FP64 rotation, its gradients and the tangent update within 1e-12; the rays within
1e-6; the pose gradients through GetRays against the oracle's closed form; and the pose gradient reaching the
translation through the projector (rays that require grad keep the fused projector, whose backward returns the
positional gradient; the kernel chain gives the same ray gradients)."""
import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from neat_amd import ops, workloads as W
from oracle import camera as CA

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.fixture(scope="module")
def H():
    import neat_amd.host as host
    return host.load()


def poses(n, seed):
    return Rotation.random(n, random_state=seed).as_quat(), np.random.default_rng(seed).normal(size=(n, 3))


def test_rotate_kernels_vs_oracle():
    print("case: rotate fwd/bwd, apply tangent", flush=True)
    q, _ = poses(7, 0)
    rng = np.random.default_rng(1)
    p = rng.normal(size=(5000, 3))
    idx = rng.integers(0, 7, 5000).astype(np.int32)
    g = rng.normal(size=(5000, 3))
    qt = torch.from_numpy(q).to(DEV)
    out = ops.rotate_points(qt, torch.from_numpy(p).to(DEV), torch.from_numpy(idx).to(DEV)).cpu().numpy()
    assert np.allclose(out, CA.rotate(q, p, idx), rtol=0, atol=1e-12)
    gt, gp = ops.rotate_points_bwd(qt, torch.from_numpy(p).to(DEV), torch.from_numpy(idx).to(DEV),
                                   torch.from_numpy(g).to(DEV))
    rgt, rgp = CA.rotate_bwd(q, p, idx, g)
    assert np.allclose(gp.cpu().numpy(), rgp, rtol=0, atol=1e-12)
    assert np.allclose(gt.cpu().numpy(), rgt, rtol=1e-12, atol=1e-10)
    t = rng.normal(size=(7, 3)) * 0.2
    t[0] = 0.0
    tt = torch.from_numpy(t.copy()).to(DEV)
    ops.so3_apply_tangent(tt, qt)
    assert np.allclose(qt.cpu().numpy(), CA.apply_tangent(t, q), rtol=0, atol=1e-14)
    assert bool((tt == 0).all())


@pytest.mark.parametrize("align", [False, True])
def test_get_rays_module_vs_oracle(H, align):
    print("case: GetRays align_corners", align, flush=True)
    q, t = poses(4, 3)
    K = np.array([[120.0, 118.0, 63.5, 47.5, 0.0], [90.0, 92.0, 60.0, 50.0, 0.0]])
    pose = H.CameraPoseModule(torch.from_numpy(q), torch.from_numpy(t))
    pose.to(DEV)
    cam = H.CameraModelModule(96, 128, torch.from_numpy(K))
    cam.to(DEV)
    rng = np.random.default_rng(4)
    n = 3000
    uv = rng.uniform(size=(n, 2)).astype(np.float32)
    img = rng.integers(0, 4, n).astype(np.int32)
    cid = rng.integers(0, 2, n).astype(np.int32)
    rays = H.GetRays(pose, cam, torch.from_numpy(uv).to(DEV), torch.from_numpy(img).to(DEV),
                     torch.from_numpy(cid).to(DEV), align)
    o, d = CA.get_rays(q, t, K, 128, 96, uv, img, cid, align)
    assert np.allclose(rays.origin.detach().cpu().numpy(), o, atol=1e-6)
    assert np.allclose(rays.direction.detach().cpu().numpy(), d, atol=1e-6)
    # gradients of sum(a * o + b * d): translation gets sum a per image; the tangent sum (R u) x b per image
    a = rng.normal(size=(n, 3)).astype(np.float32)
    b = rng.normal(size=(n, 3)).astype(np.float32)
    loss = (rays.origin * torch.from_numpy(a).to(DEV)).sum() + (rays.direction * torch.from_numpy(b).to(DEV)).sum()
    loss.backward()
    gtr = np.zeros((4, 3))
    np.add.at(gtr, img, a.astype(np.float64))
    assert np.allclose(pose.translation.grad.cpu().numpy(), gtr, rtol=1e-6, atol=1e-6)
    px = CA.uv_to_pixel(uv, 128, 96, align).astype(np.float64)
    c = (px - K[cid, 2:4]) / K[cid, 0:2]
    un = np.concatenate([c, np.ones((n, 1))], axis=1)
    un /= np.linalg.norm(un, axis=1, keepdims=True)
    rgt, _ = CA.rotate_bwd(q, un, img, b.astype(np.float64))
    assert np.allclose(pose.rotation_tangent.grad.cpu().numpy(), rgt, rtol=1e-6, atol=1e-6)
    assert pose.translation.grad is not None and cam.intrinsics.grad is not None


def test_pose_gradient_through_projector(H):
    """Rays from GetRays with pose gradients: ComputeImage keeps the fused projector, its ray gradients equal the
    kernel chain's (use_fused_projector off) within 1e-5, and the translation gradient equals the per-image sum of the
    ray-origin gradients."""
    print("case: pose gradient through the projector", flush=True)
    from test_gpu_host import make_geometry
    cfg = W.C1
    sc, tree, params, geo = make_geometry(H, cfg, seed=2)
    q, _ = poses(3, 5)
    t = np.array([[0.0, 0.0, -2.5], [0.1, 0.0, -2.4], [0.0, -0.1, -2.6]])
    # point the cameras at the origin: rotate e_z onto -t
    R = []
    for k in range(3):
        z = -t[k] / np.linalg.norm(t[k])
        x = np.cross([0.0, 1.0, 0.0], z)
        x /= np.linalg.norm(x)
        y = np.cross(z, x)
        R.append(np.stack([x, y, z], axis=1))
    q = Rotation.from_matrix(np.stack(R)).as_quat()
    pose = H.CameraPoseModule(torch.from_numpy(q), torch.from_numpy(t))
    pose.to(DEV)
    cam = H.CameraModelModule(32, 32, torch.from_numpy(np.array([[40.0, 40.0, 15.5, 15.5, 0.0]])))
    cam.to(DEV)
    rng = np.random.default_rng(6)
    n = 2048
    uv = torch.from_numpy(rng.uniform(0.2, 0.8, size=(n, 2)).astype(np.float32)).to(DEV)
    img = torch.from_numpy(rng.integers(0, 3, n).astype(np.int32)).to(DEV)
    rays = H.GetRays(pose, cam, uv, img, torch.zeros(n, dtype=torch.int32, device=DEV), False)
    o_leaf = rays.origin.detach().clone().requires_grad_(True)
    d_leaf = rays.direction.detach().clone().requires_grad_(True)
    w = torch.from_numpy(rng.normal(size=(1, n)).astype(np.float32)).to(DEV)

    def project(o, d):
        smp = tree.CreateSamplesForRays(H.RayList(o, d), cfg.max_spn, False)
        return (geo.ComputeImage(smp, 1, n) * w).sum()

    project(rays.origin, rays.direction).backward()
    project(o_leaf, d_leaf).backward()
    go = o_leaf.grad.double().cpu().numpy()
    gdir = d_leaf.grad.double().cpu().numpy()
    # the kernel chain (grid-sample backward with d uv, local-coordinate backward, index_select backward)
    params.train_params.use_fused_projector = False
    o2 = rays.origin.detach().clone().requires_grad_(True)
    d2 = rays.direction.detach().clone().requires_grad_(True)
    project(o2, d2).backward()
    params.train_params.use_fused_projector = True
    gc_o, gc_d = o2.grad.double().cpu().numpy(), d2.grad.double().cpu().numpy()
    eo = np.linalg.norm(go - gc_o) / np.linalg.norm(gc_o)
    ed = np.linalg.norm(gdir - gc_d) / np.linalg.norm(gc_d)
    print("fused vs chain ray gradients: origin %.3e, direction %.3e" % (eo, ed), flush=True)
    assert eo <= 1e-5 and ed <= 1e-5
    want = np.zeros((3, 3))
    np.add.at(want, img.cpu().numpy(), go)
    got = pose.translation.grad.cpu().numpy()
    print("translation grad", got.ravel(), "from ray grads", want.ravel())
    assert np.abs(go).sum() > 0 and np.allclose(got, want, rtol=1e-5, atol=1e-7)
    assert pose.rotation_tangent.grad is not None and bool(torch.isfinite(pose.rotation_tangent.grad).all())
