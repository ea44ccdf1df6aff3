"""Pins of the oracle's motion residuals and moving-object decision
(Tracking::compute_residuals, tracking.cpp:1535-1747; check_static_state,
rendering.cpp:907-947; kfusion.cpp:314-354), from closed-form cases.

Scene: a fronto-parallel plane at 0.3 m seen by a 64x48 camera with
fx = fy = 600 at the identity pose, the same depth in the reference and the
live frame, and a horizontal grey ramp whose reference copy is brighter by
`delta` grey levels.  Then, in closed form:

* ICP: the live vertex moved by T_w_l = translation(0, 0, dz) lands on the
  plane's model vertex map with diff = -dz along the normal, so
  e = dz / sqrt(P_z) with the variance P_z = (d^2 / (f * 0.075))^2 = 4e-6
  (calc_icp_cov), i.e. e = 500 * dz: dz = 0.002 -> e = 1 (result 1),
  dz = 0.004 -> e = 2 > 1.5 (-4 from the threshold, while |diff| < 0.1),
  dz = 0.12 -> |diff| > 0.1 (-4 from the distance gate).
* Normals: model normals tilted by 45 degrees -> dot < 0.8 -> -5, a gate
  that comes after the distance gate (dz = 0.12 still gives -4); INVALID
  model normals -> -3; projections outside [1, W-2] x [1, H-2] -> -2.
* Photometric: residual = I_live - I_ref = -delta/255, e = residual /
  (10/255): delta = 10 -> |e| = 1 (result 1), delta = 20 -> |e| = 2 (-4).
  A vertex map before fusion 0.2 m behind the live one -> occluded (-5); a
  missing one (z = 0) -> -1.
* The decision: per-object pixel and inlier counts recounted from the
  oracle's own rows and labels, and moving iff 0.05 < ratio < 0.8.
"""
import numpy as np
import pytest

from conftest import scene, oracle, K320

W, H = 64, 48
K = np.array([600.0, 600.0, 31.5, 23.5], np.float32)
D = 0.3


def frames(delta):
    depth = np.full((H, W), int(D * 1000), np.uint16)
    ramp = (60 + 2 * np.arange(W))[None, :].repeat(H, 0)
    live = np.repeat(ramp[..., None], 3, 2).astype(np.uint8)
    ref = np.repeat((ramp + delta)[..., None], 3, 2).astype(np.uint8)
    return depth, ref, live


def setup(delta):
    o = oracle.Oracle(W, H, np.eye(4, dtype=np.float32), iters=(0, 0, 0), use_rgb=True)
    o.add_object(5.12, 64, np.eye(4))
    depth, ref, live = frames(delta)
    o.preprocess(depth, ref)
    o.track_camera(K, 0)          # seeds the reference frame
    o.preprocess(depth, live)
    o.track_camera(K, 1)          # no GN iterations: the live pyramid only
    return o


def Tz(dz):
    T = np.eye(4, dtype=np.float32)
    T[2, 3] = dz
    return T


def interior(a, m=3):
    return a[m:H - m, m:W - m]


def icp(codes):
    return codes[:W * H].reshape(H, W)


def rgb(codes):
    return codes[W * H:].reshape(H, W)


@pytest.fixture(scope='module')
def ctx():
    o = setup(10)
    V, N = o.level(0, 2), o.level(0, 3)
    # the live maps are the plane: z = 0.3 and normals along -z
    assert np.allclose(V[..., 2], D, atol=1e-6)
    assert np.allclose(np.abs(N[..., 2]), 1.0, atol=1e-6)
    return o, V, N


def test_icp_threshold_and_covariance(ctx):
    o, V, N = ctx
    c = icp(o.motion_residuals(K, np.eye(4, dtype=np.float32), V, N, V))
    assert (interior(c) == 1).all()
    assert (c[:, 0] == -2).all() and (c[:, -1] == -2).all() and (c[0] == -2).all() and (c[-1] == -2).all()
    assert (interior(icp(o.motion_residuals(K, Tz(0.002), V, N, V)), 8) == 1).all()
    assert (interior(icp(o.motion_residuals(K, Tz(0.004), V, N, V)), 8) == -4).all()
    assert (interior(icp(o.motion_residuals(K, Tz(0.12), V, N, V)), 8) == -4).all()


def test_icp_normal_gates(ctx):
    o, V, N = ctx
    c45 = np.cos(np.pi / 4)
    tilted = np.zeros_like(N)
    tilted[..., 0], tilted[..., 2] = c45 * N[..., 2], c45 * N[..., 2]
    assert (interior(icp(o.motion_residuals(K, np.eye(4, dtype=np.float32), V, tilted, V))) == -5).all()
    assert (interior(icp(o.motion_residuals(K, Tz(0.002), V, tilted, V)), 8) == -5).all()
    assert (interior(icp(o.motion_residuals(K, Tz(0.12), V, tilted, V)), 8) == -4).all()
    inval = N.copy()
    inval[..., 0] = -2.0
    assert (interior(icp(o.motion_residuals(K, np.eye(4, dtype=np.float32), V, inval, V))) == -3).all()


@pytest.mark.parametrize('delta,expect', [(10, 1), (20, -4)])
def test_rgb_threshold_and_scale(delta, expect):
    o = setup(delta)
    V, N = o.level(0, 2), o.level(0, 3)
    c = rgb(o.motion_residuals(K, np.eye(4, dtype=np.float32), V, N, V))
    assert (interior(c) == expect).all(), np.unique(interior(c), return_counts=True)


def test_rgb_occlusion_and_missing_reference(ctx):
    o, V, N = ctx
    behind = V.copy()
    behind[..., 2] += 0.2
    assert (interior(rgb(o.motion_residuals(K, np.eye(4, dtype=np.float32), V, N, behind))) == -5).all()
    assert (rgb(o.motion_residuals(K, np.eye(4, dtype=np.float32), V, N, np.zeros_like(V))) == -1).all()


def test_decision_recount():
    # two boxes, one moving: the oracle's counts equal a recount of its own
    # rows over the labels before fusion, and the decision follows the
    # thresholds of kernels.h:95-97
    c = 2.56
    T0 = np.eye(4)
    T0[:3, 3] = np.array([c, c, c - 1.5]) + np.array([0.05, 0.05, 1.0]) - 0.256
    T1 = T0.copy()
    T1[:3, 3] += np.array([-0.45, 0.1, 0.2])
    mov = [scene.MovingBox([0.15, 0.12, 0.1], T0, [0.006, -0.004, 0.003, 0.02, 0.02, -0.01], [0.8, 0.3, 0.2],
                           0.512, 128),
           scene.MovingBox([0.1, 0.1, 0.1], T1, [0] * 6, [0.2, 0.7, 0.3], 0.512, 128)]
    Wc, Hc = 320, 240
    sc = scene.Scene(seed=2, width=Wc, height=Hc, K=tuple(K320), moving=mov)
    o = oracle.Oracle(Wc, Hc, sc.camera_pose(0).astype(np.float32), use_rgb=True)
    o.add_object(5.12, 256, np.eye(4))
    for m in mov:
        o.add_object(m.vol_size, m.vol_res, m.pose(0).astype(np.float32))
    o.set_motion_auto(True)
    decided = []
    for f in range(4):
        d, rgb_, lab = sc.render(f)
        for i, m in enumerate(scene.masks_for_labels(lab, 2)):
            o.set_mask(i, m)
        o.preprocess(d, rgb_)
        moving_before = o.motion_result()[2].copy()
        o.track_camera(K320, f)
        if f > 0:
            rows, _ = o.motion_rows()
            _, _, lbf = o.maps_bf()
            lbf = lbf.reshape(-1)
            n = Wc * Hc
            inl = (rows[:n] > -3) & (rows[n:] > -3)
            px, il, mv = o.motion_result()
            for j in (1, 2):
                sel = lbf == j
                assert px[j] == int(sel.sum()) and il[j] == int((inl & sel).sum())
                if px[j]:
                    r = np.float32(il[j]) / np.float32(px[j])
                    assert mv[j] == int(0.05 < r < 0.8), (f, j, r)
                else:
                    assert mv[j] == moving_before[j]
            decided.append(mv.tolist())
        o.track_objects(K320, f)
        o.integrate(K320, f)
        o.raycast(K320, f)
    assert decided[-1] == [0, 1, 0]
