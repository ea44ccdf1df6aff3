"""Failure paths and input variants of the dense path, GPU against the FP32
oracle on identical inputs (model maps and poses copied from the oracle
before each tracking call).

Reference behaviour pinned here (tracking.cpp:242-335, 624-690;
preprocessing.cpp:132-159; kfusion.cpp:210-235):
- an all-zero depth frame: with ICP only every row is invalid, the camera's
  error is 0/0 = NaN, no comparison with it holds, the solve runs on zero
  sums and the pose is left as it was;
- a frame whose integration masks are all -1 allocates and updates nothing;
- an object with no model surface has zero inliers and breaks at once on
  every level (tracking.cpp:674-676);
- a NaN in the model vertex map enters the sums through the rows whose
  distance and normal tests it defeats, and the GN step turns the pose NaN;
- a step-back restores the previous pose; a converged iteration breaks
  without applying its update;
- compute_size_ratio 2 subsamples the input; an input size that is not a
  whole multiple of the computation size is an error (the reference exits);
- the -F filter bilateral-filters the integration depth.
"""
import numpy as np
import pytest

from conftest import scene, midfusion, oracle, K320, K640

pytestmark = pytest.mark.gpu
W, H = 320, 240


def mover():
    T0 = np.eye(4)
    T0[:3, 3] = np.array([2.56, 2.56, 1.06]) + np.array([0.05, 0.05, 1.0]) - 0.256
    return scene.MovingBox([0.15, 0.12, 0.1], T0, [0.003, -0.002, 0.002, 0.01, 0.012, -0.008], [0.8, 0.3, 0.2],
                           0.512, 128)


class Pair:
    """A GPU context and an FP32 oracle fed identical inputs."""

    def __init__(self, movers=(), use_rgb=True, seed=1, w=W, h=H, k=K320):
        self.w, self.h, self.k = w, h, k
        self.movers = list(movers)
        self.sc = scene.Scene(seed=seed, width=w, height=h, K=tuple(k), moving=self.movers)
        P0 = self.sc.camera_pose(0).astype(np.float32)
        self.gpu = midfusion.Kfusion(w, h, P0, bg_size=5.12, bg_res=256, use_rgb=use_rgb)
        self.orc = oracle.Oracle(w, h, P0, use_rgb=use_rgb)
        self.orc.add_object(5.12, 256, np.eye(4))
        for j, m in enumerate(self.movers):
            T = m.pose(0).astype(np.float32)
            for x in (self.gpu, self.orc):
                x.add_object(m.vol_size, m.vol_res, T)
                x.set_moving(j + 1, True)

    def inputs(self, f, depth=None, masks=None):
        d, rgb, lab = self.sc.render(f)
        if depth is not None:
            d = depth
        if masks is None:
            masks = scene.masks_for_labels(lab, len(self.movers))
        for x in (self.gpu, self.orc):
            for i, m in enumerate(masks):
                x.set_mask(i, m)
        self.gpu.preprocessing(d, rgb)
        self.orc.preprocess(d, rgb)

    def sync_state(self):
        vtx, nrm, lb = self.orc.maps()
        self.gpu.set_maps(vtx, nrm, lb)
        for i in range(1, 1 + len(self.movers)):
            self.gpu.set_object_maps(i, *self.orc.object_maps(i))
        for i in [-1] + list(range(1, 1 + len(self.movers))):
            self.gpu.setPose(i, self.orc.pose(i).astype(np.float32))

    def track(self, f):
        for x in (self.gpu, self.orc):
            x.trace_clear()
        self.gpu.track_camera(self.k, f)
        self.orc.track_camera(self.k, f)
        self.gpu.setPose(-1, self.orc.pose().astype(np.float32))
        self.gpu.track_objects(self.k, f)
        self.orc.track_objects(self.k, f)
        self.gpu.sync()
        return self.gpu.trace(), self.orc.trace()

    def map(self, f):
        for i in [-1] + list(range(1, 1 + len(self.movers))):
            self.gpu.setPose(i, self.orc.pose(i).astype(np.float32))
        for i in range(1 + len(self.movers)):
            self.gpu.set_trackresult(i, self.orc.trackresult(i)[0])
        self.gpu.integration(self.k, f)
        self.orc.integrate(self.k, f)
        self.gpu.raycasting(self.k, f)
        self.orc.raycast(self.k, f)
        self.gpu.sync()

    def close(self):
        self.gpu.close()


def flow(tr):
    """(instance, level, iteration, flag) of every executed GN iteration."""
    return [tuple(int(v) for v in r[[0, 1, 2, 5]]) for r in tr]


def check_flow(tg, to, ties=True):
    """Identical GN control flow.  With ties, a differing decision is
    accepted only at a near-tie of the oracle's own numbers (error change
    within 1e-4 of the error, or the residual test within 1e-4 of its 2e-3
    bound), after which that instance's records are not compared further.
    Returns the number of records compared."""
    fg, fo = flow(tg), flow(to)
    stopped, n = set(), 0
    i = 0
    while i < min(len(fg), len(fo)):
        inst = fo[i][0]
        if inst in stopped:
            i += 1
            continue
        if fg[i] != fo[i]:
            assert ties, (i, fg[i], fo[i])
            e, ep = to[i, 3], to[i - 1, 3] if i > 0 else np.inf
            near = abs(e - ep) <= 1e-4 * abs(ep) or abs(np.sqrt(e) - 2e-3) <= 2e-7
            assert near, ('control flow differs away from a tie', i, fg[i], fo[i], e, ep)
            stopped.add(inst)
            break
        n += 1
        i += 1
    if not stopped:
        assert len(fg) == len(fo), (len(fg), len(fo))
    return n


def run_invariants(tr, nlev=3):
    """Within each GN run of the trace: a converged iteration (flag 1) leaves
    the pose of the previous iteration; a step-back (flag 2) restores the
    pose of the iteration before that.  Returns the flags seen."""
    seen = set()
    start = 0
    for n in range(len(tr)):
        inst, lvl, it, flag = (int(v) for v in tr[n, [0, 1, 2, 5]])
        if lvl == nlev - 1 and it == 0:
            start = n
        seen.add(flag)
        P = tr[n, 6:22]
        if flag == 1 and n - 1 >= start:
            assert np.array_equal(P, tr[n - 1, 6:22]), ('converged iteration changed the pose', n)
        if flag == 2 and n - 2 >= start:
            assert np.array_equal(P, tr[n - 2, 6:22]), ('step-back did not restore the previous pose', n)
    return seen


def test_zero_depth_frame_icp_only():
    p = Pair(use_rgb=False)
    for f in range(4):
        zero = np.zeros((H, W), np.uint16) if f == 2 else None
        p.inputs(f, depth=zero)
        if f > 0:
            p.sync_state()
        P_before = p.orc.pose()
        tg, to = p.track(f)
        if f > 0:
            assert check_flow(tg, to, ties=(f != 2)) > 0
        if f == 2:
            # every iteration: no inliers, NaN error, pose unchanged
            for t in (tg, to):
                assert np.all(t[:, 4] == 0) and np.all(np.isnan(t[:, 3])), t[:, :6]
                assert np.all(t[:, 5] == 0)
            assert np.array_equal(p.gpu.getPose(), P_before.astype(np.float32))
            assert np.array_equal(p.orc.pose(), P_before)
            keys_before = [p.orc.keys(0)]
        p.map(f)
        assert np.array_equal(p.gpu.keys(0), p.orc.keys(0)), f
        if f == 2:
            assert np.array_equal(p.orc.keys(0), keys_before[0])
    p.close()


def test_zero_depth_frame_icp_rgb():
    # photometric rows still hold (the reference frame has depth, the live
    # frame has intensity): control flow and pose follow the oracle
    p = Pair(use_rgb=True)
    for f in range(4):
        p.inputs(f, depth=np.zeros((H, W), np.uint16) if f == 2 else None)
        if f > 0:
            p.sync_state()
        tg, to = p.track(f)
        if f > 0:
            check_flow(tg, to)
            d = np.abs(p.gpu.getPose() - p.orc.pose()).max()
            assert d < 5e-5, (f, d)
        p.map(f)
        assert np.array_equal(p.gpu.keys(0), p.orc.keys(0)), f
    p.close()


def test_fully_masked_frame_allocates_and_updates_nothing():
    p = Pair(movers=[mover()])
    for f in range(4):
        masks = [np.full((H, W), -1.0, np.float32)] * 2 if f == 2 else None
        p.inputs(f, masks=masks)
        if f > 0:
            p.sync_state()
        p.track(f)
        if f == 2:
            keys = [p.gpu.keys(i) for i in range(2)]
            vox = [p.gpu.voxels(i, keys[i])[0] for i in range(2)]
        p.map(f)
        for i in range(2):
            assert np.array_equal(p.gpu.keys(i), p.orc.keys(i)), (f, i)
        if f == 2:
            for i in range(2):
                assert np.array_equal(p.gpu.keys(i), keys[i])
                assert np.array_equal(p.gpu.voxels(i, keys[i])[0], vox[i])
    p.close()


def test_object_without_surface_breaks_on_zero_inliers():
    """Object 2 is never integrated (its mask is -1 everywhere), so its model
    maps hold no surface and every one of its rows is invalid: each level
    breaks at its first iteration with no inliers, and the pose stays at the
    static guess T_w_c2 * (T_w_c2^-1 * T_w_o1)."""
    empty = mover()
    empty.T0 = empty.T0.copy()
    empty.T0[:3, 3] += [-0.6, 0.0, 0.3]
    p = Pair(movers=[mover(), empty])
    for f in range(3):
        d, rgb, lab = p.sc.render(f)
        masks = scene.masks_for_labels(lab, 2)
        masks[2] = np.full((H, W), -1.0, np.float32)
        p.inputs(f, masks=masks)
        if f > 0:
            p.sync_state()
        T_before = p.orc.pose(2)
        tg, to = p.track(f)
        if f > 0:
            for t in (tg, to):
                rec = t[t[:, 0] == 2] if len(t) else t
                if len(rec):
                    assert np.all(rec[:, 5] == 3) and np.all(rec[:, 4] == 0), rec[:, :6]
                    assert sorted(rec[:, 1].astype(int)) == [0, 1, 2]
            assert np.abs(p.gpu.getPose(2) - p.orc.pose(2)).max() < 1e-6
            assert np.abs(p.orc.pose(2) - T_before).max() < 1e-5
        p.map(f)
    p.close()


def test_nan_model_vertex_propagates_as_in_the_reference():
    p = Pair(use_rgb=True)
    for f in range(3):
        p.inputs(f)
        if f > 0:
            p.sync_state()
        if f == 2:
            vtx, nrm, lb = p.orc.maps()
            ok = np.argwhere(nrm[..., 0] != -2)
            for y, x in ok[:: max(1, len(ok) // 7)][:7]:
                vtx[y, x, 0] = np.nan
            p.orc.set_maps(vtx, nrm, lb)
            p.gpu.set_maps(vtx, nrm, lb)
            for x in (p.gpu, p.orc):
                x.trace_clear()
            p.gpu.track_camera(K320, f)
            p.orc.track_camera(K320, f)
            p.gpu.sync()
            tg, to = p.gpu.trace(), p.orc.trace()
            assert flow(tg) == flow(to)
            assert np.isnan(p.orc.pose()).any() and np.isnan(p.gpu.getPose()).any()
            assert np.array_equal(np.isnan(p.gpu.getPose()), np.isnan(p.orc.pose()))
            break   # a NaN pose ends the sequence
        p.track(f)
        p.map(f)
    p.close()


def test_converged_break_without_update():
    """A constant-depth frame (a fronto-parallel plane 1.5 m ahead, camera
    rotation the identity) tracked against model maps holding the plane's
    exact points and normals: every ICP row's residual n . (v_ref - p) is 0
    (both points lie on the plane, the normal is +-z), so the sums give x = 0
    and sqrt(s0 / s28) = 0: each level converges at its first iteration and
    breaks without applying the update (tracking.cpp:326-331, 1365)."""
    P0 = np.eye(4, dtype=np.float32)
    P0[:3, 3] = [2.56, 2.56, 1.0]
    gpu = midfusion.Kfusion(W, H, P0, bg_size=5.12, bg_res=256, use_rgb=False)
    orc = oracle.Oracle(W, H, P0, use_rgb=False)
    orc.add_object(5.12, 256, np.eye(4))
    depth = np.full((H, W), 1500, np.uint16)
    rgb = np.zeros((H, W, 3), np.uint8)
    mask = np.ones((H, W), np.float32)
    for f in range(2):
        for x in (gpu, orc):
            x.set_mask(0, mask)
        gpu.preprocessing(depth, rgb)
        orc.preprocess(depth, rgb)
        if f == 1:
            v = orc.level(0, 2)       # the plane's vertices (camera frame) and normals
            n = orc.level(0, 3)
            assert np.all(np.abs(n[1:-1, 1:-1, :2]) == 0) and np.all(np.abs(n[1:-1, 1:-1, 2]) > 0.99)
            vw = (v + P0[:3, 3]).astype(np.float32)
            lab = np.zeros((H, W), np.int32)
            for x in (gpu, orc):
                x.set_maps(vw, n, lab)
                x.trace_clear()
        gpu.track_camera(K320, f)
        orc.track_camera(K320, f)
        gpu.sync()
        if f == 1:
            tg, to = gpu.trace(), orc.trace()
            for t in (tg, to):
                assert flow(t) == [(0, 2, 0, 1), (0, 1, 0, 1), (0, 0, 0, 1)], flow(t)
                assert np.all(t[:, 3] == 0)
            assert np.array_equal(gpu.getPose(), P0)
            assert np.array_equal(orc.pose(), P0.astype(np.float64))
        gpu.track_objects(K320, f)
        orc.track_objects(K320, f)
    gpu.close()


def test_step_back_restores_previous_pose():
    p = Pair(movers=[mover()])
    seen_g, seen_o, compared = set(), set(), 0
    for f in range(6):
        p.inputs(f)
        if f > 0:
            p.sync_state()
        tg, to = p.track(f)
        if f > 0:
            compared += check_flow(tg, to)
            seen_g |= run_invariants(tg)
            seen_o |= run_invariants(to)
        p.map(f)
    print('flags seen: gpu', seen_g, 'oracle', seen_o, 'records compared', compared)
    assert 2 in seen_o and 2 in seen_g
    assert compared > 50
    p.close()


def test_size_ratio_two_and_bad_ratio():
    """A 640x480 input into a 320x240 computation (compute_size_ratio 2):
    every second pixel of every second row (preprocessing.cpp:132-159,
    73-102); frame buffers bit-exact and block keys bit-exact after two
    tracked frames.  A 500x480 input is not a whole multiple: MF_E_ARG."""
    sc = scene.Scene(seed=1, width=640, height=480, K=tuple(K640))
    P0 = sc.camera_pose(0).astype(np.float32)
    gpu = midfusion.Kfusion(W, H, P0, bg_size=5.12, bg_res=256, use_rgb=True)
    orc = oracle.Oracle(W, H, P0, use_rgb=True)
    orc.add_object(5.12, 256, np.eye(4))
    for f in range(3):
        d, rgb, lab = sc.render(f)
        m = scene.masks_for_labels(lab[::2, ::2], 0)[0]
        for x in (gpu, orc):
            x.set_mask(0, m)
        gpu.preprocessing(d, rgb)
        assert orc.preprocess(d, rgb) == 0
        for a, b in zip(gpu.frame_buffers(), orc.frame_buffers()):
            assert np.array_equal(a, b), f
        assert np.array_equal(gpu.frame_buffers()[0], (d[::2, ::2].astype(np.float32) / np.float32(1000.0)))
        if f > 0:
            vtx, nrm, lb = orc.maps()
            gpu.set_maps(vtx, nrm, lb)
            gpu.setPose(-1, orc.pose().astype(np.float32))
        gpu.tracking(K320, f)
        orc.track(K320, f)
        gpu.setPose(-1, orc.pose().astype(np.float32))
        gpu.set_trackresult(0, orc.trackresult(0)[0])
        gpu.integration(K320, f)
        orc.integrate(K320, f)
        gpu.raycasting(K320, f)
        orc.raycast(K320, f)
        assert np.array_equal(gpu.keys(0), orc.keys(0)), f
    bad_d = np.zeros((480, 500), np.uint16)
    bad_rgb = np.zeros((480, 500, 3), np.uint8)
    with pytest.raises(midfusion.MFError):
        gpu.preprocessing(bad_d, bad_rgb)
    assert orc.preprocess(bad_d, bad_rgb) != 0
    gpu.close()


def test_filtered_integration_depth():
    """-F: the integration depth is bilateral filtered (kfusion.cpp:218-221),
    so allocation and integration see the filtered depth."""
    sc = scene.Scene(seed=1, width=W, height=H, K=tuple(K320))
    P0 = sc.camera_pose(0).astype(np.float32)
    gpu = midfusion.Kfusion(W, H, P0, bg_size=5.12, bg_res=256, use_rgb=True)
    orc = oracle.Oracle(W, H, P0, use_rgb=True)
    orc.add_object(5.12, 256, np.eye(4))
    for f in range(3):
        d, rgb, lab = sc.render(f)
        m = scene.masks_for_labels(lab, 0)[0]
        for x in (gpu, orc):
            x.set_mask(0, m)
        gpu.preprocessing(d, rgb, filt=True)
        orc.preprocess(d, rgb, filt=True)
        dg, do_ = gpu.frame_buffers()[0], orc.frame_buffers()[0]
        assert np.array_equal(dg, do_), (f, float(np.abs(dg - do_).max()))
        assert not np.array_equal(dg, d.astype(np.float32) / np.float32(1000.0))
        if f > 0:
            vtx, nrm, lb = orc.maps()
            gpu.set_maps(vtx, nrm, lb)
            gpu.setPose(-1, orc.pose().astype(np.float32))
        gpu.tracking(K320, f)
        orc.track(K320, f)
        gpu.setPose(-1, orc.pose().astype(np.float32))
        gpu.set_trackresult(0, orc.trackresult(0)[0])
        gpu.integration(K320, f)
        orc.integrate(K320, f)
        gpu.raycasting(K320, f)
        orc.raycast(K320, f)
        assert np.array_equal(gpu.keys(0), orc.keys(0)), f
    gpu.close()
