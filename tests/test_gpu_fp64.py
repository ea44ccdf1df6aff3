"""Pose error of the FP32 GPU path against the FP64 oracle.

The GPU accumulates residual rows in FP32 per thread and per workgroup and
in FP64 across workgroups, then solves the 6x6 system in FP64; the FP32
oracle follows the reference's own 8-block sequential FP32 sums
(tracking.cpp:1113-1176).  Both are held against the FP64 oracle on the
same frame and the same starting state: the model maps and poses of every
context are copied from the FP32 oracle before each tracking call, so the
only differences are arithmetic precision and summation order.

The first coarse-level GN steps are sensitive to that arithmetic: the FP32
oracle's own iterates differ from the FP64 oracle's by up to ~6 mm after the
first step and converge to within ~5e-6 m.  So the per-frame result is held
to 1e-5, and each intermediate iterate to the FP32 oracle's own error
envelope (twice it, plus 1e-5).
"""
import numpy as np
import pytest

from conftest import scene, midfusion, oracle, K320
from lockstep import condition, check_poses

pytestmark = pytest.mark.gpu
W, H = 320, 240


def rot_err(A, B):
    R = A[:3, :3].astype(np.float64).T @ B[:3, :3].astype(np.float64)
    s = 0.5 * np.linalg.norm([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
    return float(np.arctan2(s, (np.trace(R) - 1) / 2))


def pose_err(A, B):
    return float(np.linalg.norm(A[:3, 3] - B[:3, 3])), rot_err(A, B)


def big_box():
    c = 2.56
    T0 = np.eye(4)
    T0[:3, 3] = np.array([c, c, c - 1.5]) + np.array([0.05, 0.05, 1.0]) - 0.256
    return scene.MovingBox([0.15, 0.12, 0.1], T0, [0.003, -0.002, 0.002, 0.01, 0.012, -0.008], [0.8, 0.3, 0.2],
                           0.512, 128)


def sphere():
    """A uniform-albedo sphere: rotations about its centre change neither its
    ideal depth nor its intensity."""
    c = 2.56
    T0 = np.eye(4)
    T0[:3, 3] = np.array([c, c, c - 1.5]) + np.array([0.05, 0.05, 1.0]) - 0.256
    return scene.MovingBox([0.1, 0.1, 0.1], T0, [0.003, -0.002, 0.002, 0.01, 0.012, -0.008], [0.7, 0.6, 0.5],
                           0.512, 128, shape='ellipsoid')


def tsdf_vs(gpu, orc, i, max_blocks=3000):
    """tsdf and weight of instance i's allocated blocks (up to max_blocks,
    evenly spaced over the sorted keys), GPU against an oracle: worst
    differences and the fraction of voxels with |dtsdf| > 1e-4 or dw != 0."""
    L = oracle.lib(False)
    keys = orc.keys(i)
    keys = keys[np.linspace(0, len(keys) - 1, min(len(keys), max_blocks)).astype(int)] if len(keys) else keys
    xyz = np.zeros(3, np.uint32)
    wt = ww = 0.0
    bad = n = 0
    bulk = gpu.voxels(i, keys)[0] if hasattr(gpu, 'voxels') else None
    for kk, k in enumerate(keys):
        L.orc_unpack(int(k) & ~0x1FF, xyz)
        bg = bulk[kk] if bulk is not None else gpu.block(i, *map(int, xyz))
        bo = orc.block(i, *map(int, xyz))
        if bg is None or bo is None:
            bad += 512
            n += 512
            continue
        dt, dw = np.abs(bg[:, 0] - bo[:, 0]), np.abs(bg[:, 1] - bo[:, 1])
        wt, ww = max(wt, float(dt.max())), max(ww, float(dw.max()))
        bad += int(np.sum((dt > 1e-4) | (dw > 0)))
        n += 512
    return {'blocks': len(keys), 'tsdf_max': wt, 'w_max': ww, 'outside': bad / max(n, 1)}


def run(n_frames, moving, W=W, H=H, K=K320, bg_res=256, invalid_frac=0.0, seed=1, tsdf=False):
    mov = moving if isinstance(moving, list) else ([big_box()] if moving else [])
    sc = scene.Scene(seed=seed, width=W, height=H, K=tuple(K), moving=mov, invalid_frac=invalid_frac)
    P0 = sc.camera_pose(0).astype(np.float32)
    gpu = midfusion.Kfusion(W, H, P0, bg_size=5.12, bg_res=bg_res, use_rgb=True)
    o32 = oracle.Oracle(W, H, P0, use_rgb=True)
    o64 = oracle.Oracle(W, H, P0, use_rgb=True, fp64=True)
    for o in (o32, o64):
        o.add_object(5.12, bg_res, np.eye(4))
    for j, m in enumerate(mov):
        T = m.pose(0).astype(np.float32)
        for x in (gpu, o32, o64):
            x.add_object(m.vol_size, m.vol_res, T)
            x.set_moving(j + 1, True)
    report = []
    for f in range(n_frames):
        d, rgb, lab = sc.render(f)
        masks = scene.masks_for_labels(lab, len(mov))
        for x in (gpu, o32, o64):
            for i, m in enumerate(masks):
                x.set_mask(i, m)
        gpu.preprocessing(d, rgb)
        o32.preprocess(d, rgb)
        o64.preprocess(d, rgb)
        if f > 0:
            # identical model maps and starting poses everywhere
            vtx, nrm, lb = o32.maps()
            o64.set_maps(vtx, nrm, lb)
            gpu.set_maps(vtx, nrm, lb)
            for i in range(1, 1 + len(mov)):
                ov, on = o32.object_maps(i)
                o64.set_object_maps(i, ov, on)
                gpu.set_object_maps(i, ov, on)
            for i in [-1] + list(range(1, 1 + len(mov))):
                T = o32.pose(i).astype(np.float32)
                o64.set_pose(i, T)
                gpu.setPose(i, T)
        for x in (gpu, o32, o64):
            x.trace_clear()
        gpu.track_camera(K, f)
        o32.track_camera(K, f)
        o64.track_camera(K, f)
        gpu.sync()
        r = {'frame': f}
        P64 = o64.pose()
        r['cam_gpu'] = pose_err(gpu.getPose(), P64)
        r['cam_o32'] = pose_err(o32.pose(), P64)
        if f > 0:
            tg, t64, t32 = gpu.trace(), o64.trace(), o32.trace()
            r['iters_gpu'], r['iters_o64'] = len(tg), len(t64)
            # per-iteration poses while the control flow agrees
            n = 0
            worst = (0.0, 0.0)
            # columns: instance, level, iteration, error, inliers, flag, pose
            key = [0, 1, 2, 5]
            while n < min(len(tg), len(t64)) and np.array_equal(tg[n, key], t64[n, key]):
                e = pose_err(tg[n, 6:22].reshape(4, 4), t64[n, 6:22].reshape(4, 4))
                worst = (max(worst[0], e[0]), max(worst[1], e[1]))
                n += 1
            r['iter_agree'] = n
            r['iter_worst'] = worst
            r['iter_t_err'] = [round(pose_err(tg[q, 6:22].reshape(4, 4), t64[q, 6:22].reshape(4, 4))[0], 7)
                               for q in range(n)]
            r['iter_envelope_ok'] = True
            for q in range(min(n, len(t32))):
                if not np.array_equal(t32[q, key], t64[q, key]):
                    break
                e_g = pose_err(tg[q, 6:22].reshape(4, 4), t64[q, 6:22].reshape(4, 4))
                e_o = pose_err(t32[q, 6:22].reshape(4, 4), t64[q, 6:22].reshape(4, 4))
                if e_g[0] > 2 * e_o[0] + 1e-5 or e_g[1] > 2 * e_o[1] + 1e-5:
                    r['iter_envelope_ok'] = False
        # objects start from the FP64 camera pose everywhere
        gpu.setPose(-1, P64.astype(np.float32))
        o32.set_pose(-1, P64.astype(np.float32))
        if f > 0:
            for i in range(1, 1 + len(mov)):
                s = o64.eval(i, 0, np.linalg.inv(P64) @ o64.pose(i))
                r['obj%d_cond' % i] = condition(s)
                r['obj%d_inliers' % i] = int(s[28])
        for x in (gpu, o32, o64):
            x.trace_clear()
        gpu.track_objects(K, f)
        o32.track_objects(K, f)
        o64.track_objects(K, f)
        gpu.sync()
        tg, t32, t64 = gpu.trace(), o32.trace(), o64.trace()
        for i in range(1, 1 + len(mov)):
            fg, f32, fo = ([tuple(int(v) for v in q[[1, 2, 5]]) for q in t if int(q[0]) == i] for t in (tg, t32, t64))
            r['obj%d_flow_same' % i] = fg == fo
            r['obj%d_flow_same_o32' % i] = f32 == fo
            Q64 = o64.pose(i)
            r['obj%d_gpu' % i] = pose_err(gpu.getPose(i), Q64)
            r['obj%d_o32' % i] = pose_err(o32.pose(i), Q64)
            # column 24: 1 certified Gauss-Jordan, 2 Jacobi pseudo-inverse
            r['obj%d_pinv' % i] = int(np.sum((tg[:, 0] == i) & (tg[:, 24] == 2)))
        for x in (gpu, o32, o64):
            x.integrate(K, f) if x is not gpu else x.integration(K, f)
        if tsdf:
            gpu.sync()
            r['tsdf_bg_vs_o64'] = tsdf_vs(gpu, o64, 0)
            r['tsdf_bg_o32_vs_o64'] = tsdf_vs(o32, o64, 0)
            r['tsdf_bg_vs_o32'] = tsdf_vs(gpu, o32, 0)
        for x in (gpu, o32, o64):
            x.raycast(K, f) if x is not gpu else x.raycasting(K, f)
        print(r)
        report.append(r)
    gpu.close()
    return report


def test_camera_pose_vs_fp64_oracle():
    rep = run(4, moving=False)
    for r in rep[1:]:
        assert r['cam_gpu'][0] < 1e-5 and r['cam_gpu'][1] < 1e-5, r
        assert r['iter_agree'] > 0, r
        assert r['iter_envelope_ok'], r


def test_camera_and_object_pose_vs_fp64_oracle():
    rep = run(4, moving=True)
    check_poses(rep, 1)


def test_weakly_observable_object_vs_fp64_oracle():
    """A uniform sphere's rotations about its centre are observable only
    through the TSDF's discretisation and the depth noise, so its 6x6 systems
    are ill-conditioned and both FP32 paths drift from the FP64 result along
    them.  Its condition number is far past the bound of the per-object
    contract (lockstep.check_poses), so every frame is an exception there,
    held in aggregate: the GPU follows the FP64 control flow as often as the
    FP32 oracle does, with a median error no larger than the FP32 oracle's.
    Systems past the solve's certificate (the pseudo-inverse branch) are
    checked directly in test_gpu_solve."""
    rep = run(4, moving=[sphere()])
    exc = check_poses(rep, 1)
    # rotations about the centre are unobservable but for discretisation:
    # the sphere's systems are far past the conditioning bound
    assert all(r['obj1_cond'] > 1e4 for r in rep[1:]), [r['obj1_cond'] for r in rep[1:]]
    assert len(exc) == len(rep) - 1


def second_box():
    c = 2.56
    T0 = np.eye(4)
    T0[:3, 3] = np.array([c, c, c - 1.5]) + np.array([-0.35, -0.05, 1.1]) - 0.256
    return scene.MovingBox([0.1, 0.08, 0.07], T0, [-0.002, 0.003, 0.001, -0.012, 0.008, 0.01], [0.3, 0.7, 0.4],
                           0.512, 128)


def test_two_objects_batched_vs_fp64_oracle():
    """Two textured boxes tracked in one launch per GN iteration, each object
    under the per-object contract of lockstep.check_poses (1e-5 of the FP64
    oracle when well conditioned with the FP64 control flow, else held in
    aggregate against the FP32 oracle).  (A near-degenerate
    object, such as a small ellipsoid of ~1000 pixels, is no test of this:
    there the FP32 paths' errors at the finer levels differ by ~1% from the
    FP64 oracle's, and a step-back taken one iteration earlier or later
    moves the result by millimetres, on the CPU oracle as well.)"""
    rep = run(4, moving=[big_box(), second_box()])
    check_poses(rep, 2)


def test_precision_scale_1280x720_16_objects():
    """The precision + scale configuration at full size: 1280x720 (fx = fy =
    1050), a 1024^3 background at 5 mm, 16 moving boxes and ellipsoids of
    0.05-0.4 m, 2 % random invalid depth pixels, three frames.  Camera pose
    within 1e-5 m / 1e-5 rad of the FP64 oracle.  Background TSDF and weights:
    on the first frame (identical poses everywhere) bit-identical to the FP32
    oracle's; against the FP64 oracle's, about 1 % of the voxels differ in
    both FP32 paths, because a projection rounded in FP32 picks another pixel
    at half-pixel boundaries (an invalid pixel or the far side of a depth
    edge, so one side updates the voxel and the other does not), and the
    GPU's share stays within 1.5x the FP32 oracle's.  Object poses under the
    per-object contract of lockstep.check_poses."""
    K1050 = np.array([1050.0, 1050.0, 639.5, 359.5], np.float32)
    rep = run(3, moving=scene.batch_objects(16, 5000), W=1280, H=720, K=K1050, bg_res=1024, invalid_frac=0.02,
              seed=5000, tsdf=True)
    for r in rep[1:]:
        errs = [(r['obj%d_gpu' % i], r['obj%d_o32' % i]) for i in range(1, 17)]
        print('objects (gpu, o32) vs o64:', errs)
    assert rep[0]['tsdf_bg_vs_o32']['outside'] == 0.0, rep[0]['tsdf_bg_vs_o32']
    for r in rep:
        g, o = r['tsdf_bg_vs_o64'], r['tsdf_bg_o32_vs_o64']
        assert g['outside'] <= 1.5 * o['outside'] + 1e-4, (g, o)
    check_poses(rep, 16)
