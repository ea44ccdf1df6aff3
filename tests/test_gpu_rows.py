"""GPU parity of every per-pixel tracking row (TrackData: result code, error,
Jacobian) against the CPU oracle on identical inputs.

One GN iteration is run at a single pyramid level (the others get zero
iterations), so the rows are computed from the same starting pose on both
sides.  Covers the four row kinds: camera ICP and photometric
(tracking.cpp:923-1111) and object ICP and photometric (tracking.cpp:392-578).
Result codes must agree exactly; error and Jacobian within 1e-5 relative.
"""
import numpy as np
import pytest

from conftest import scene, midfusion, oracle, K320

pytestmark = pytest.mark.gpu
W, H = 320, 240


def big_box():
    c = 2.56
    T0 = np.eye(4)
    T0[:3, 3] = np.array([c, c, c - 1.5]) + np.array([0.05, 0.05, 1.0]) - 0.256
    return scene.MovingBox([0.15, 0.12, 0.1], T0, [0.003, -0.002, 0.002, 0.01, 0.012, -0.008], [0.8, 0.3, 0.2],
                           0.512, 128)


def compare_rows(a, b, tag):
    ra, ea, Ja = a
    rb, eb, Jb = b
    mism = np.nonzero(ra != rb)[0]
    print(tag, 'valid', int((ra == 1).sum()), 'code mismatches', len(mism), ra[mism[:5]], rb[mism[:5]])
    assert len(mism) == 0, tag
    ok = ra == 1
    assert ok.sum() > 50, tag
    scale_e = max(1e-6, np.abs(eb[ok]).max())
    scale_J = max(1e-6, np.abs(Jb[ok]).max())
    de = np.abs(ea[ok] - eb[ok]).max() / scale_e
    dJ = np.abs(Ja[ok] - Jb[ok]).max() / scale_J
    print(tag, 'max rel err diff %.2e, J diff %.2e' % (de, dJ))
    assert de < 1e-5 and dJ < 1e-5, tag


@pytest.mark.parametrize('level', [2, 1, 0])
@pytest.mark.parametrize('use_rgb', [False, True])
def test_camera_rows(level, use_rgb):
    iters = [0, 0, 0]
    iters[level] = 1
    sc = scene.Scene(seed=1, width=W, height=H, K=tuple(K320))
    P0 = sc.camera_pose(0).astype(np.float32)
    g = midfusion.Kfusion(W, H, P0, bg_res=256, use_rgb=use_rgb, iters=tuple(iters))
    o = oracle.Oracle(W, H, P0, use_rgb=use_rgb, iters=tuple(iters))
    o.add_object(5.12, 256, np.eye(4))
    g.debug_rows(True)
    m = np.ones((H, W), np.float32)
    for f in range(2):
        d, rgb, _ = sc.render(f)
        for x in (g, o):
            x.set_mask(0, m)
        g.preprocessing(d, rgb)
        o.preprocess(d, rgb)
        g.tracking(K320, f)
        o.track(K320, f)
        g.sync()
        if f == 1:
            compare_rows(g.rows(0, level), o.rows(level), 'camera L%d rgb=%d' % (level, use_rgb))
            if use_rgb:
                # T1: the rendered reference depth pyramid the photometric
                # rows read, every level bit-exact
                for lv in range(3):
                    rg, ro = g.level(lv, 7), o.level(lv, 7)
                    assert np.array_equal(rg, ro), (lv, int(np.count_nonzero(rg != ro)))
                    assert np.count_nonzero(rg) > 100
        g.setPose(-1, o.pose().astype(np.float32))
        g.integration(K320, f)
        o.integrate(K320, f)
        g.raycasting(K320, f)
        o.raycast(K320, f)
    g.close()


@pytest.mark.parametrize('level', [2, 1, 0])
def test_object_rows(level):
    iters = [0, 0, 0]
    iters[level] = 1
    mb = big_box()
    sc = scene.Scene(seed=1, width=W, height=H, K=tuple(K320), moving=[mb])
    P0 = sc.camera_pose(0).astype(np.float32)
    g = midfusion.Kfusion(W, H, P0, bg_res=256, use_rgb=True, iters=tuple(iters))
    o = oracle.Oracle(W, H, P0, use_rgb=True, iters=tuple(iters))
    o.add_object(5.12, 256, np.eye(4))
    T = mb.pose(0).astype(np.float32)
    g.add_object(mb.vol_size, mb.vol_res, T)
    o.add_object(mb.vol_size, mb.vol_res, T)
    g.set_moving(1, True)
    o.set_moving(1, True)
    g.debug_rows(True)
    for f in range(2):
        d, rgb, lab = sc.render(f)
        for i, m in enumerate(scene.masks_for_labels(lab, 1)):
            g.set_mask(i, m)
            o.set_mask(i, m)
        g.preprocessing(d, rgb)
        o.preprocess(d, rgb)
        g.track_camera(K320, f)
        o.track_camera(K320, f)
        g.sync()
        g.setPose(-1, o.pose().astype(np.float32))
        g.track_objects(K320, f)
        o.track_objects(K320, f)
        g.sync()
        if f == 1:
            compare_rows(g.rows(1, level), o.rows(level), 'object L%d' % level)
        g.setPose(1, o.pose(1).astype(np.float32))
        g.integration(K320, f)
        o.integrate(K320, f)
        g.raycasting(K320, f)
        o.raycast(K320, f)
    g.close()
