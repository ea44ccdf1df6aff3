"""Finite-difference pins of the tracking residual's Jacobians (CPU only).

The reduction's [1..6] sums are J^T W e with J = -de/dxi for a left
perturbation T <- exp(xi) T, xi = (v, w); with the correspondences of every
row locally fixed, the derivative of the robust loss L = sum rho(e) is
therefore -J^T W e (robust weight w = psi(e)/e).  FP64 central differences of
the FP64 oracle's loss pin each row kind where that holds:

- camera ICP picks the reference pixel by truncating the projection, so the
  correspondence is piecewise constant: with h = 1e-9 no row changes pixel;
- photometric rows interpolate the live intensity and its central-difference
  gradients bilinearly, which is the exact derivative for an image linear in
  x (an 8-bit ramp: R = G = B = column index), away from the one-pixel
  border strip whose gradients the reference leaves at zero (that strip is
  why the camera case is held to the reference's own 1% criterion,
  apps/kfusion/test/tracking/jacobians_unittest.cpp:224-295).

The object ICP rows interpolate the model maps bilinearly and J ignores the
motion of that interpolated correspondence (the reference's Gauss-Newton
linearisation), so no finite difference of L pins them; their rows are pinned
against the GPU bit for bit instead.
"""
import numpy as np
import pytest
from scipy.linalg import expm

from conftest import scene, oracle

W, H = 160, 120
K = np.array([131.25, 131.25, 79.5, 59.5], np.float32)
RAMP = np.repeat(np.tile(np.arange(W, dtype=np.uint8)[None, :, None], (H, 1, 1)), 3, axis=2)


def twist(x):
    v, w = x[:3], x[3:]
    M = np.zeros((4, 4))
    M[:3, 3] = v
    M[:3, :3] = [[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]]
    return expm(M)


def big_box():
    c = 2.56
    T0 = np.eye(4)
    T0[:3, 3] = np.array([c, c, c - 1.5]) + np.array([0.05, 0.05, 1.0]) - 0.256
    return scene.MovingBox([0.15, 0.12, 0.1], T0, [0.003, -0.002, 0.002, 0.01, 0.012, -0.008], [0.8, 0.3, 0.2],
                           0.512, 128)


def prepared(use_icp, use_rgb, obj, ramp, seed=1):
    """FP64 oracle after frame 0 and frame 1's camera phase; returns it, the
    instance id and a pose perturbed by ~1e-3 from the tracked one."""
    mov = [big_box()] if obj else []
    sc = scene.Scene(seed=seed, width=W, height=H, K=tuple(K), moving=mov)
    o = oracle.Oracle(W, H, sc.camera_pose(0).astype(np.float32), use_icp=use_icp, use_rgb=use_rgb, robust='huber',
                      robust_k=0.1, fp64=True)
    o.add_object(5.12, 256, np.eye(4))
    for m in mov:
        o.add_object(m.vol_size, m.vol_res, m.pose(0).astype(np.float32))
        o.set_moving(1, True)
    for f in range(2):
        d, rgb, lab = sc.render(f)
        for i, m in enumerate(scene.masks_for_labels(lab, len(mov))):
            o.set_mask(i, m)
        o.preprocess(d, RAMP if ramp else rgb)
        o.track_camera(K, f)
        if f == 1:
            break
        o.track_objects(K, f)
        o.integrate(K, f)
        o.raycast(K, f)
    oid = 1 if obj else -1
    T = np.linalg.inv(o.pose(-1)) @ o.pose(1) if obj else o.pose(-1)
    rng = np.random.default_rng(seed + 2)
    return o, oid, twist(np.r_[rng.normal(0, 1e-3, 3), rng.normal(0, 1e-3, 3)]) @ T


def fd_gradient(o, oid, level, T, h):
    g = np.zeros(6)
    for k in range(6):
        e = np.zeros(6)
        e[k] = h
        g[k] = (o.eval(oid, level, twist(e) @ T)[0] - o.eval(oid, level, twist(-e) @ T)[0]) / (2 * h)
    return g


@pytest.mark.parametrize('use_icp,use_rgb,obj,ramp,h,tol', [
    (True, False, False, False, 1e-9, 1e-5),    # camera ICP
    (False, True, False, True, 1e-6, 1e-2),     # camera photometric
    (False, True, True, True, 1e-6, 1e-6),      # object photometric
])
def test_adjoint_matches_fp64_central_differences(use_icp, use_rgb, obj, ramp, h, tol):
    o, oid, T = prepared(use_icp, use_rgb, obj, ramp)
    s = o.eval(oid, 0, T)
    assert s[28] > 500
    g = -s[1:7]
    fd = fd_gradient(o, oid, 0, T, h)
    rel = np.abs(fd - g).max() / np.abs(g).max()
    print('adjoint vs FD', use_icp, use_rgb, obj, rel)
    assert rel < tol


def test_adjoint_sign_convention_breaks_fd():
    # the pin is sensitive: the opposite sign (or a dropped rotational term)
    # fails the same comparison
    o, oid, T = prepared(True, False, False, False)
    s = o.eval(oid, 0, T)
    fd = fd_gradient(o, oid, 0, T, 1e-9)
    assert np.abs(fd - s[1:7]).max() / np.abs(fd).max() > 0.5
    g = -s[1:7].copy()
    g[3:] = 0
    assert np.abs(fd - g).max() / np.abs(fd).max() > 0.05
