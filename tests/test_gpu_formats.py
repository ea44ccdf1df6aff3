"""GPU side of the on-disk formats: frames streamed from a .raw sequence
through pinned staging (mf_preprocess_raw), the ground-truth camera mode fed
by a TUM ground-truth file (kfusion.cpp:262-268, parseGTFile), and the
trajectory files of Kfusion::save_poses (kfusion.cpp:1591-1626)."""
import os

import numpy as np
import pytest

from conftest import scene, midfusion, oracle, K320
import formats

pytestmark = pytest.mark.gpu
W, H = 320, 240


def test_preprocess_raw_matches_host_frames(tmp_path):
    sc = scene.Scene(seed=3, width=W, height=H, K=tuple(K320))
    fr = [sc.render(f)[:2] for f in range(3)]
    p = str(tmp_path / 'scene.raw')
    formats.write_raw(p, fr)
    raw = midfusion.RawSequence(p)
    P0 = sc.camera_pose(0).astype(np.float32)
    a = midfusion.Kfusion(W, H, P0, bg_res=256)
    b = midfusion.Kfusion(W, H, P0, bg_res=256)
    for f in (0, 1, 2, 1):
        print('frame', f, flush=True)
        a.preprocessing_raw(raw, f)
        b.preprocessing(*fr[f])
        for x, y in zip(a.frame_buffers(), b.frame_buffers()):
            assert np.array_equal(x, y)
    with pytest.raises(midfusion.MFError):
        a.preprocessing_raw(raw, 3)
    raw.close()
    a.close()
    b.close()


def test_ground_truth_camera_mode(tmp_path):
    sc = scene.Scene(seed=1, width=W, height=H, K=tuple(K320))
    n = 4
    # a TUM ground-truth file of the scene's camera, read back by the library
    lines = []
    for f in range(n):
        T = sc.camera_pose(f).astype(np.float32)
        q = formats.rot_quat(T)
        lines.append(' '.join('%.9g' % float(v) for v in [f * 0.033, T[0, 3], T[1, 3], T[2, 3]] + list(q)))
    p = str(tmp_path / 'groundtruth.txt')
    open(p, 'w').write('\n'.join(lines) + '\n')
    G = midfusion.parse_gt(p)
    assert len(G) == n
    P0 = np.eye(4, dtype=np.float32)
    P0[:3, 3] = sc.camera_pose(0)[:3, 3]
    P0[:3, :3] = G[0][:3, :3]
    gpu = midfusion.Kfusion(W, H, P0, bg_size=5.12, bg_res=256, use_rgb=True)
    orc = oracle.Oracle(W, H, P0, use_rgb=True)
    orc.add_object(5.12, 256, np.eye(4))
    gpu.set_gt_poses(G)
    orc.set_gt_poses(G)
    m = np.ones((H, W), np.float32)
    for f in range(n):
        print('frame', f, flush=True)
        d, rgb, _ = sc.render(f)
        for x in (gpu, orc):
            x.set_mask(0, m)
        gpu.preprocessing(d, rgb)
        orc.preprocess(d, rgb)
        gpu.tracking(K320, f)
        orc.track(K320, f)
        gpu.sync()
        Pg, Po = gpu.getPose(), orc.pose().astype(np.float32)
        # the re-based ground-truth pose: poses[f] moved by P0 - poses[0]
        want = G[f].copy()
        want[:3, 3] = (G[f][:3, 3] - G[0][:3, 3]) + P0[:3, 3]
        assert np.array_equal(Pg, want) and np.array_equal(Po, want), (f, Pg, want)
        gpu.integration(K320, f)
        orc.integrate(K320, f)
        gpu.sync()
        kg, ko = gpu.keys(0), orc.keys(0)
        assert np.array_equal(kg, ko), f
        gpu.raycasting(K320, f)
        orc.raycast(K320, f)
        _, _, lg = gpu.maps()
        _, _, lo = orc.maps()
        assert np.mean(lg != lo) < 1e-3
    with pytest.raises(midfusion.MFError):
        gpu.tracking(K320, n)            # no ground-truth pose for this frame
    gpu.close()


def test_save_poses(tmp_path):
    P0 = np.eye(4, dtype=np.float32)
    P0[:3, 3] = [2.5, 2.4, 1.0]
    g = midfusion.Kfusion(W, H, P0, bg_res=64)
    T1 = np.eye(4, dtype=np.float32)
    T1[:3, 3] = [0.1, 0.2, 0.3]
    g.add_object(0.5, 64, T1)
    prefix = str(tmp_path / 'traj')
    g.save_poses(prefix, 0)
    T1b = T1.copy()
    T1b[:3, :3] = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], np.float32)
    g.setPose(1, T1b)
    g.save_poses(prefix, 1)
    cam = open(prefix + '_camera').read()
    obj = open(prefix + '_1').read()
    assert cam == formats.tum_line(0, P0) + formats.tum_line(1, P0)
    # each instance's own translation
    assert obj == formats.tum_line(0, T1) + formats.tum_line(1, T1b)
    g.save_poses(prefix, 2)          # appends
    assert open(prefix + '_1').read().count('\n') == 3
    g2 = midfusion.Kfusion(W, H, P0, bg_res=64)
    g2.save_poses(prefix, 5)         # a new context starts the files again
    assert open(prefix + '_camera').read() == formats.tum_line(5, P0)
    assert os.path.exists(prefix + '_1')
    g.close()
    g2.close()
