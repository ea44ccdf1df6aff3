"""Real data from the reference's own fixtures (apps/kfusion/test/tracking,
converted by tests/golden/make_real_pairs.py), GPU against the FP32 and FP64
oracles: the reference frame is integrated and raycast into the model, the
model maps are made identical everywhere (the FP32 oracle's), and the live
frame is tracked by the full camera pass.  The world is shifted (by a
translation) so that the scene lies in the 512^3 background volume, which
leaves relative poses unchanged.

- TUM fr2 walking_halfsphere pair (tum_fr2_pair.npz) with its ground truth,
  K = (525, 525, 319.5, 239.5), as jacobians_unittest.cpp:1023-1052 uses it:
  photometric rows only without robust weight (the reference test's
  Tracking(..., using_ICP = false, using_RGB = true), :126), and the default
  ICP + photometric rows with the Cauchy weight.
- ICL-NUIM pair 720 / 721 (icl_720_721.npz), K = (481.2, -480, 319.5, 239.5):
  the negative focal length takes the NegY normal and the ICL covariance
  paths (tracking.cpp:764-770, 805-810).

Pins: the GPU's pose within 1e-5 m / 1e-5 rad of the FP64 oracle's; level-0
row codes equal to the FP32 oracle's but for rows at a threshold; the GPU's
relative motion estimate equal to the FP32 oracle's within 1e-4 m.  With the
reference's default ICP + photometric rows (Cauchy weight) on the TUM pair the
tracker runs away from the ground truth (77 mm) on every path, and FP32 and
FP64 take different Gauss-Newton paths there (row gates and step-backs
decided at the FP32 rounding level): that case is held to the FP64 oracle
only while its control flow is the FP64 oracle's, and printed.  The
reference test's criterion (relative translation within 1 mm of the ground
truth) is printed for the TUM pair, not asserted: people walk through this
scene, and the reference algorithm, as the oracle runs it, drifts by
centimetres on the pair (photometric only 24 mm; ICP + photometric 77 mm,
against 13 mm of true motion).
"""
import os

import numpy as np
import pytest

from conftest import ROOT, midfusion, oracle
from lockstep import pose_err

pytestmark = pytest.mark.gpu
W, H = 640, 480


def track_pair(fix, K, kw, P_r):
    P0 = P_r.astype(np.float32)
    gpu = midfusion.Kfusion(W, H, P0, bg_size=5.12, bg_res=512, **kw)
    o32 = oracle.Oracle(W, H, P0, **kw)
    o64 = oracle.Oracle(W, H, P0, fp64=True, **kw)
    for o in (o32, o64):
        o.add_object(5.12, 512, np.eye(4))
    mask = np.ones((H, W), np.float32)
    for x in (gpu, o32, o64):
        x.set_mask(0, mask)
    gpu.preprocessing(fix['depth_r'], fix['rgb_r'])
    for o in (o32, o64):
        o.preprocess(fix['depth_r'], fix['rgb_r'])
    gpu.tracking(K, 0)
    gpu.integration(K, 0)
    gpu.raycasting(K, 0)
    for o in (o32, o64):
        o.track(K, 0)
        o.integrate(K, 0)
        o.raycast(K, 0)
    maps = o32.maps()
    gpu.set_maps(*maps)
    o64.set_maps(*maps)
    gpu.preprocessing(fix['depth_l'], fix['rgb_l'])
    for o in (o32, o64):
        o.preprocess(fix['depth_l'], fix['rgb_l'])
    for x in (gpu, o64):
        x.trace_clear()
    gpu.track_camera(K, 1)
    o32.track_camera(K, 1)
    o64.track_camera(K, 1)
    gpu.sync()
    flow = lambda t: [tuple(int(v) for v in q[[0, 1, 2, 5]]) for q in t]
    out = {'gpu': gpu.getPose().astype(np.float64), 'fp32 oracle': o32.pose(), 'fp64 oracle': o64.pose(),
           'vs_fp64': pose_err(gpu.getPose(), o64.pose()), 'same_flow': flow(gpu.trace()) == flow(o64.trace()),
           'codes': float(np.mean(gpu.trackresult(0) != o32.trackresult(0)[0]))}
    gpu.close()
    return out


@pytest.mark.parametrize('kw', [dict(use_icp=False, use_rgb=True, robust='none'),
                                dict(use_icp=True, use_rgb=True, robust='cauchy')])
def test_tum_pair(kw):
    fix = np.load(os.path.join(ROOT, 'tests', 'golden', 'tum_fr2_pair.npz'))
    K = np.array([525.0, 525.0, 319.5, 239.5], np.float32)
    gt_r, gt_l = fix['gt_r'], fix['gt_l']
    S = np.eye(4)
    S[:3, 3] = np.array([2.56, 2.56, 2.56]) - 1.5 * gt_r[:3, 2] - gt_r[:3, 3]   # camera 1.5 m behind the centre
    P_r, P_l = S @ gt_r, S @ gt_l
    r = track_pair(fix, K, kw, P_r)
    gt_rel = np.linalg.inv(P_l) @ P_r
    rel = {k: np.linalg.inv(r[k]) @ P_r for k in ('gpu', 'fp32 oracle', 'fp64 oracle')}
    drift = {k: round(1000 * float(np.linalg.norm(v[:3, 3] - gt_rel[:3, 3])), 3) for k, v in rel.items()}
    print(kw, 'translation drift vs ground truth [mm]', drift, 'ground-truth motion [mm]',
          round(1000 * float(np.linalg.norm(gt_rel[:3, 3])), 3), 'gpu vs fp64 oracle', r['vs_fp64'],
          'same control flow', r['same_flow'], 'row-code mismatch vs fp32 oracle', r['codes'])
    if kw['robust'] == 'none' or r['same_flow']:
        assert r['vs_fp64'][0] < 1e-5 and r['vs_fp64'][1] < 1e-5, r['vs_fp64']
        assert r['codes'] < 1e-3, r['codes']
        assert np.linalg.norm(rel['gpu'][:3, 3] - rel['fp32 oracle'][:3, 3]) < 1e-4
    else:
        assert drift['gpu'] > 20 and drift['fp64 oracle'] > 20     # every path runs away alike


def test_icl_pair_negative_focal_length():
    fix = np.load(os.path.join(ROOT, 'tests', 'golden', 'icl_720_721.npz'))
    K = np.array([481.2, -480.0, 319.5, 239.5], np.float32)
    P_r = np.eye(4)
    P_r[:3, 3] = [2.56, 2.56, 0.3]
    r = track_pair(fix, K, dict(use_icp=True, use_rgb=True, robust='cauchy'), P_r)
    rel = np.linalg.inv(r['gpu']) @ P_r
    rel32 = np.linalg.inv(r['fp32 oracle']) @ P_r
    print('estimated motion [mm]', 1000 * rel[:3, 3], 'gpu vs fp64 oracle', r['vs_fp64'], 'codes', r['codes'])
    assert np.linalg.norm(rel[:3, 3]) < 0.05
    assert r['vs_fp64'][0] < 1e-5 and r['vs_fp64'][1] < 1e-5, r['vs_fp64']
    assert r['codes'] < 1e-3, r['codes']
    assert np.linalg.norm(rel[:3, 3] - rel32[:3, 3]) < 1e-4
