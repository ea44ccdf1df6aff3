"""GPU parity of the display renderers and the label projection
(rendering/rendering.cpp; Kfusion::renderTrack / renderDepth /
renderIntensity / renderVolume / volume2mask, kfusion.cpp:679-686,
1284-1475) against the CPU oracle.

Inputs are made identical first (the oracle's poses, rows, row errors and
model maps are copied into the GPU context), so that:
* track, depth and intensity images are bit-exact;
* volume2mask labels equal the oracle's except at grazing pixels (< 0.1%);
* renderVolume, from the last raycast's maps or by raycasting at another
  view or size, with palette or voxel colours, is bit-exact.
"""
import numpy as np
import pytest

from conftest import scene, midfusion, oracle, K320

pytestmark = pytest.mark.gpu
W, H = 320, 240
PALETTE = np.array([[(37 * i) % 256, (91 * i + 40) % 256, (53 * i + 90) % 256, 0] for i in range(91)], np.uint8)


@pytest.fixture(scope='module')
def pair():
    c = 2.56
    T0 = np.eye(4)
    T0[:3, 3] = np.array([c, c, c - 1.5]) + np.array([0.05, 0.05, 1.0]) - 0.256
    mov = [scene.MovingBox([0.15, 0.12, 0.1], T0, [0.003, -0.002, 0.002, 0.01, 0.012, -0.008], [0.8, 0.3, 0.2],
                           0.512, 128)]
    sc = scene.Scene(seed=1, width=W, height=H, K=tuple(K320), moving=mov)
    P0 = sc.camera_pose(0).astype(np.float32)
    gpu = midfusion.Kfusion(W, H, P0, bg_size=5.12, bg_res=256, use_rgb=True)
    orc = oracle.Oracle(W, H, P0, use_rgb=True)
    orc.add_object(5.12, 256, np.eye(4))
    T = mov[0].pose(0).astype(np.float32)
    gpu.add_object(0.512, 128, T)
    orc.add_object(0.512, 128, T)
    for x in (gpu, orc):
        x.set_moving(1, True)
    for f in range(3):
        print('frame', f, flush=True)
        d, rgb, lab = sc.render(f)
        for i, m in enumerate(scene.masks_for_labels(lab, 1)):
            gpu.set_mask(i, m)
            orc.set_mask(i, m)
        gpu.preprocessing(d, rgb)
        orc.preprocess(d, rgb)
        gpu.track_camera(K320, f)
        orc.track_camera(K320, f)
        gpu.setPose(-1, orc.pose().astype(np.float32))
        gpu.track_objects(K320, f)
        orc.track_objects(K320, f)
        gpu.setPose(1, orc.pose(1).astype(np.float32))
        gpu.setPose(-1, orc.pose().astype(np.float32))
        for i in (0, 1):
            res, err = orc.trackresult(i)
            gpu.set_trackresult(i, res)
            gpu.set_track_errors(i, err[W * H:])
        gpu.integration(K320, f)
        orc.integrate(K320, f)
        gpu.raycasting(K320, f)
        orc.raycast(K320, f)
        gpu.sync()
    # identical model maps for the renders that read them
    vtx, nrm, lb = orc.maps()
    gpu.set_maps(vtx, nrm, lb)
    yield sc, gpu, orc
    gpu.close()


def test_track_depth_intensity_renders_bit_exact(pair):
    sc, gpu, orc = pair
    for oid in (0, 1):
        for mode in (0, 1, 2):
            a, b = gpu.render_track(oid, mode), orc.render_track(oid, mode)
            assert np.array_equal(a, b), (oid, mode, int(np.sum(np.any(a != b, axis=2))))
    assert np.array_equal(gpu.render_depth(), orc.render_depth())
    assert np.array_equal(gpu.render_intensity(), orc.render_intensity())
    # the RGBD render shows every class of the legend on this frame
    img = orc.render_track(0, 2)
    assert len(np.unique(img.reshape(-1, 4), axis=0)) > 4


def test_volume2mask(pair):
    sc, gpu, orc = pair
    a, b = gpu.volume2mask(K320), orc.volume2mask(K320)
    assert np.mean(a != b) < 1e-3
    assert (b == 1).sum() > 1000 and (b == 0).sum() > 10000


def test_render_volume_from_maps(pair):
    sc, gpu, orc = pair
    P = orc.pose().astype(np.float32)
    for color in (False, True):
        a, ra = gpu.render_volume(K320, P, render_color=color, palette=PALETTE)
        b, rb = orc.render_volume(K320, P, render_color=color, palette=PALETTE)
        assert not ra and not rb            # the raycast pose at the computation size: maps
        d = np.abs(a.astype(int) - b.astype(int)).max(axis=2)
        print('maps color=%d: max %d, frac %.5f' % (color, d.max(), np.mean(d > 0)))
        assert d.max() == 0


def test_render_volume_raycast_at_another_view(pair):
    sc, gpu, orc = pair
    P = orc.pose().astype(np.float32).copy()
    P[:3, 3] += np.array([0.05, -0.03, 0.02], np.float32)
    for color in (False, True):
        a, ra = gpu.render_volume(K320, P, largestep=0.075, render_color=color, palette=PALETTE)
        b, rb = orc.render_volume(K320, P, largestep=0.075, render_color=color, palette=PALETTE)
        assert ra and rb
        d = np.abs(a.astype(int) - b.astype(int)).max(axis=2)
        print('ray color=%d: frac>1 %.5f frac>0 %.5f' % (color, np.mean(d > 1), np.mean(d > 0)))
        assert d.max() == 0
        assert (b.reshape(-1, 4)[:, :3].sum(1) > 0).mean() > 0.5   # mostly surface
    # half size (k halved): raycast path too
    k2 = np.array(K320, np.float32) / 2
    a, ra = gpu.render_volume(k2, P, W // 2, H // 2, render_color=False, palette=PALETTE)
    b, rb = orc.render_volume(k2, P, W // 2, H // 2, render_color=False, palette=PALETTE)
    assert ra and rb and a.shape == (H // 2, W // 2, 4)
    assert np.array_equal(a, b)
