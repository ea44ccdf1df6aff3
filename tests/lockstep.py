"""Lockstep runs of the GPU path beside the FP32 and FP64 oracles.

Every frame starts all three from identical inputs: the model maps and the
poses of the FP32 oracle are copied into the FP64 oracle and the GPU context
before tracking.  Tracking results are compared against the FP64 oracle (the
pose contract: 1e-5 m / 1e-5 rad); the FP32 oracle's own error is recorded
beside it.  Integration and raycasting then run on identical inputs (the GPU
takes the FP32 oracle's poses and track rows), so block keys compare bit for
bit and voxels and maps within their tolerances against the FP32 oracle.

The conditioning of each object's GN problem is measured before it is
tracked: the FP64 oracle's level-0 J^T W J at the object's starting pose
(T_c2_o2 = T_w_c2^-1 T_w_o1), through its residual evaluation.
"""
import numpy as np

from conftest import scene, midfusion, oracle


def rot_err(A, B):
    R = A[:3, :3].astype(np.float64).T @ B[:3, :3].astype(np.float64)
    s = 0.5 * np.linalg.norm([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
    return float(np.arctan2(s, (np.trace(R) - 1) / 2))


def pose_err(A, B):
    return float(np.linalg.norm(A[:3, 3] - B[:3, 3])), rot_err(A, B)


def jtj(sums):
    """6x6 J^T W J from the 32 reduction sums (upper triangle, row-major, at 7..27)."""
    A = np.zeros((6, 6))
    c = 7
    for a in range(6):
        for b in range(a, 6):
            A[a, b] = A[b, a] = sums[c]
            c += 1
    return A


def condition(sums):
    """lambda_max / lambda_min of J^T W J (inf when singular or empty)."""
    if sums[28] <= 0:
        return float('inf')
    ev = np.linalg.eigvalsh(jtj(sums))
    return float(ev[-1] / ev[0]) if ev[0] > 0 else float('inf')


def fp32_floor(sums, T):
    """Rounding floor of a GN solution in FP32: the standard deviation of
    the twist that an independent error of one FP32 ulp of the point
    coordinates (|t| of the pose plus 1 m) in every residual induces,
    sigma^2 (J^T W J)^-1 with sigma = 2^-24 * (|t| + 1) * sqrt(s28 / trace
    W) ~ 2^-24 * (|t| + 1) in the rows' weighted units; returned as the
    largest translation and rotation standard deviations (m, rad)."""
    A = jtj(sums)
    if sums[28] <= 0:
        return (float('inf'), float('inf'))
    try:
        C = np.linalg.pinv(A, rcond=1e-12)
    except np.linalg.LinAlgError:
        return (float('inf'), float('inf'))
    # the rows carry the weight w = 1/sigma_icp, so a coordinate error d
    # enters e as w * d: scale by the mean weight from the Gauss-Newton
    # curvature of the translation block
    wmean = np.sqrt(max(np.trace(A[:3, :3]) / max(sums[28], 1.0), 1e-30))
    sig = 2.0 ** -24 * (np.linalg.norm(np.asarray(T)[:3, 3]) + 1.0) * wmean
    d = np.sqrt(np.maximum(np.diag(C), 0.0)) * sig
    return (float(d[:3].max()), float(d[3:].max()))


def voxels_vs(gpu, orc, i, max_blocks=None):
    """Voxels of instance i's blocks (the oracle's key set, or max_blocks of
    it evenly spaced), GPU against an oracle: worst differences per field and
    the fraction of voxels outside tsdf/rgb 1e-4, weight and counts exact,
    fg 1e-6."""
    L = oracle.lib(False)
    keys = orc.keys(i)
    if max_blocks and len(keys) > max_blocks:
        keys = keys[np.linspace(0, len(keys) - 1, max_blocks).astype(int)]
    vg, found = gpu.voxels(i, keys)
    xyz = np.zeros(3, np.uint32)
    worst = np.zeros(8)
    bad = 0
    tol = np.array([1e-4, 0, 1e-4, 1e-4, 1e-4, 0, 1e-6, 0])
    for kk, k in enumerate(keys):
        L.orc_unpack(int(k) & ~0x1FF, xyz)
        bo = orc.block(i, int(xyz[0]), int(xyz[1]), int(xyz[2]))
        d = np.abs(vg[kk] - bo)
        worst = np.maximum(worst, d.max(axis=0))
        bad += int(np.any(d > tol, axis=1).sum())
    n = 512 * max(len(keys), 1)
    return {'blocks': len(keys), 'found': int(found), 'tsdf_max': float(worst[0]), 'w_max': float(worst[1]),
            'rgb_max': float(worst[2:5].max()), 'fg_max': float(worst[6]), 'count_max': float(max(worst[5], worst[7])),
            'outside': bad / n}


def lockstep(movers, W, H, K, n_frames, seed, bg_res, batch=True, invalid_frac=0.0, voxel_blocks=None,
             per_object_calls=False, tuning=None):
    """Runs n_frames of the full path on the GPU and both oracles; returns
    one report dict per frame.  per_object_calls: the GPU tracks each moving
    object with mf_track_object before mf_track_objects."""
    sc = scene.Scene(seed=seed, width=W, height=H, K=tuple(K), moving=movers, invalid_frac=invalid_frac)
    P0 = sc.camera_pose(0).astype(np.float32)
    gpu = midfusion.Kfusion(W, H, P0, bg_size=5.12, bg_res=bg_res, use_rgb=True, batch_objects=batch)
    for knob, v in (tuning or {}).items():
        gpu.set_tuning(knob, v)
    o32 = oracle.Oracle(W, H, P0, use_rgb=True)
    o64 = oracle.Oracle(W, H, P0, use_rgb=True, fp64=True)
    for o in (o32, o64):
        o.add_object(5.12, bg_res, np.eye(4))
    for j, m in enumerate(movers):
        T = m.pose(0).astype(np.float32)
        for x in (gpu, o32, o64):
            x.add_object(m.vol_size, m.vol_res, T)
            x.set_moving(j + 1, True)
    n_obj = len(movers)
    report = []
    for f in range(n_frames):
        d, rgb, lab = sc.render(f)
        masks = scene.masks_for_labels(lab, n_obj)
        for x in (gpu, o32, o64):
            for i, m in enumerate(masks):
                x.set_mask(i, m)
        gpu.preprocessing(d, rgb)
        o32.preprocess(d, rgb)
        o64.preprocess(d, rgb)
        if f > 0:
            vtx, nrm, lb = o32.maps()
            o64.set_maps(vtx, nrm, lb)
            gpu.set_maps(vtx, nrm, lb)
            for i in range(1, 1 + n_obj):
                ov, on = o32.object_maps(i)
                o64.set_object_maps(i, ov, on)
                gpu.set_object_maps(i, ov, on)
            for i in [-1] + list(range(1, 1 + n_obj)):
                T = o32.pose(i).astype(np.float32)
                o64.set_pose(i, T)
                gpu.setPose(i, T)
        gpu.track_camera(K, f)
        o32.track_camera(K, f)
        o64.track_camera(K, f)
        gpu.sync()
        r = {'frame': f}
        P64 = o64.pose()
        r['cam_gpu'] = pose_err(gpu.getPose(), P64)
        r['cam_o32'] = pose_err(o32.pose(), P64)
        # objects start from the FP64 camera pose everywhere
        gpu.setPose(-1, P64.astype(np.float32))
        o32.set_pose(-1, P64.astype(np.float32))
        if f > 0:
            for i in range(1, 1 + n_obj):
                T0 = np.linalg.inv(P64) @ o64.pose(i)
                s = o64.eval(i, 0, T0)
                r['obj%d_cond' % i] = condition(s)
                r['obj%d_inliers' % i] = int(s[28])
                r['obj%d_floor' % i] = fp32_floor(s, T0)
        for x in (gpu, o32, o64):
            x.trace_clear()
        if per_object_calls and f > 0:
            for i in range(1, 1 + n_obj):
                gpu.track_object(i, K)
        gpu.track_objects(K, f)
        o32.track_objects(K, f)
        o64.track_objects(K, f)
        gpu.sync()
        tg, t32, t64 = gpu.trace(), o32.trace(), o64.trace()
        for i in range(1, 1 + n_obj):
            fg, f32, fo = ([tuple(int(v) for v in q[[1, 2, 5]]) for q in t if int(q[0]) == i] for t in (tg, t32, t64))
            r['obj%d_flow_same' % i] = fg == fo
            r['obj%d_flow_same_o32' % i] = f32 == fo
            Q64 = o64.pose(i)
            r['obj%d_gpu' % i] = pose_err(gpu.getPose(i), Q64)
            r['obj%d_o32' % i] = pose_err(o32.pose(i), Q64)
            r['obj%d_pose' % i] = gpu.getPose(i)
        # integration and raycast on identical inputs (the FP32 oracle's poses
        # and rows)
        gpu.setPose(-1, o32.pose().astype(np.float32))
        for i in range(1, 1 + n_obj):
            gpu.setPose(i, o32.pose(i).astype(np.float32))
        for i in range(1 + n_obj):
            cg, (co, _) = gpu.trackresult(i), o32.trackresult(i)
            r['rows%d_mismatch' % i] = float(np.mean(cg != co))
            gpu.set_trackresult(i, co)
        gpu.integration(K, f)
        o32.integrate(K, f)
        o64.integrate(K, f)
        gpu.sync()
        for i in range(1 + n_obj):
            kg, ko = gpu.keys(i), o32.keys(i)
            r['keys%d' % i] = (len(kg), len(ko), int(len(np.setxor1d(kg, ko))))
            r['vox%d' % i] = voxels_vs(gpu, o32, i, voxel_blocks)
        gpu.raycasting(K, f)
        o32.raycast(K, f)
        o64.raycast(K, f)
        vg, ng, lg = gpu.maps()
        vo, no, lo = o32.maps()
        both = (lg == lo) & (lg >= 0)
        dv = np.abs(vg - vo)[both]
        r['label_mismatch'] = float(np.mean(lg != lo))
        r['vtx_max'] = float(dv.max()) if dv.size else 0.0
        r['vtx_p999'] = float(np.percentile(dv, 99.9)) if dv.size else 0.0
        r['vtx_frac_over_1e-4'] = float(np.mean(dv > 1e-4)) if dv.size else 0.0
        print({k: v for k, v in r.items() if not k.endswith('_pose')}, flush=True)
        report.append(r)
    gpu.close()
    return report


# Per-object pose contract.  An object is held to 1e-5 m / 1e-5 rad of the
# FP64 oracle when its GN problem is well conditioned (lambda_max /
# lambda_min of its level-0 J^T W J at the starting pose below COND_BOUND,
# where the FP32 reference arithmetic itself still meets 1e-5) and the GPU's
# GN control flow (every level's iterations, step-backs and breaks) is the
# FP64 oracle's.  Every other object is an exception, printed with its
# condition number: the FP32 rows round the iterates by about 1e-7 m, which
# can flip one row's distance or normal gate or a near-tied step-back test,
# and the GN then takes another path (the FP32 oracle does the same on other
# objects).  Exceptions are held in aggregate: the GPU follows the FP64
# control flow as often as the FP32 oracle does (within one object in
# seven), and its median error is no larger than the FP32 oracle's.
COND_BOUND = 5e3


def check_poses(rep, n_obj):
    """Camera within 1e-5 of the FP64 oracle on every tracked frame; objects
    under the per-object contract above.  Returns the exceptions as (frame,
    id, condition, gpu err, FP32 oracle err, same flow as FP64)."""
    excepted, bad, flows = [], [], []
    for r in rep[1:]:
        assert r['cam_gpu'][0] < 1e-5 and r['cam_gpu'][1] < 1e-5, r
        for i in range(1, 1 + n_obj):
            eg, eo, kappa = r['obj%d_gpu' % i], r['obj%d_o32' % i], r['obj%d_cond' % i]
            same = r['obj%d_flow_same' % i]
            flows.append((same, r['obj%d_flow_same_o32' % i]))
            row = (r['frame'], i, round(kappa), eg, eo, same, r['obj%d_inliers' % i])
            print('object', row)
            if kappa < COND_BOUND and same:
                if not (eg[0] < 1e-5 and eg[1] < 1e-5):
                    bad.append(row)
            else:
                excepted.append(row)
    print('exceptions (frame, id, cond, gpu err, fp32 oracle err, same flow as fp64, inliers):', excepted)
    assert not bad, bad
    f = np.array(flows, float)
    if len(f):
        assert f[:, 0].mean() >= f[:, 1].mean() - 1.0 / 7.0, ('control flow', f.mean(axis=0))
    if excepted:
        g = np.array([e[3] for e in excepted])
        o = np.array([e[4] for e in excepted])
        assert np.all(np.median(g, axis=0) <= np.median(o, axis=0) + 1e-6), (np.median(g, axis=0), np.median(o, axis=0))
    return excepted


def check_maps(rep, n_obj):
    """Keys bit-exact, voxels within tolerance, raycast maps within 1e-4 m."""
    for r in rep:
        for i in range(1 + n_obj):
            assert r['keys%d' % i][2] == 0, (r['frame'], i, r['keys%d' % i])
            v = r['vox%d' % i]
            assert v['found'] == v['blocks'], (r['frame'], i, v)
            assert v['outside'] < 1e-4, (r['frame'], i, v)
        assert r['label_mismatch'] < 1e-3, r
        assert r['vtx_frac_over_1e-4'] < 1e-3, r
