"""The benchmark configurations at their stated sizes, GPU against the FP32
and FP64 oracles (lockstep protocol: tests/lockstep.py).

- Plain method on the GPU: 640x480 (fx = fy = 525), background 512^3 at 1 cm,
  one textured box 0.2 x 0.15 x 0.1 m in its own 128^3 volume at 2 mm on a
  constant twist, ICP + photometric {10, 5, 4}, one object per launch.
- Batching: 640x480, camera + K seeded boxes/ellipsoids of 0.05-0.4 m in
  64^3-256^3 volumes at 2-5 mm (scene seed 4, objects seed 4000 + K), all
  objects per launch; K = 4 and K = 8.
Reference sequencing: Kfusion::tracking / integration / raycasting
(kfusion.cpp:237-677), trackEachObject (tracking.cpp:580-708).
"""
import numpy as np
import pytest

from conftest import scene, midfusion, K640
from lockstep import lockstep, check_poses, check_maps

pytestmark = pytest.mark.gpu
W, H = 640, 480


def plain_box():
    """The single-object workload's box: 0.2 x 0.15 x 0.1 m, 128^3 at 2 mm,
    twist of at most 4 mm / 0.9 degrees per frame."""
    rng = np.random.default_rng(2)
    T0 = np.eye(4)
    T0[:3, 3] = [2.30 - 0.128, 2.45 - 0.128, 2.45]
    tw = np.r_[rng.uniform(-0.004, 0.004, 3), rng.uniform(-0.015, 0.015, 3)]
    return scene.MovingBox([0.1, 0.075, 0.05], T0, tw, rng.uniform(0.3, 0.9, 3), 0.256, 128)


def test_plain_one_object_per_launch_640x480_bg512():
    rep = lockstep([plain_box()], W, H, K640, 4, seed=2, bg_res=512, batch=False, voxel_blocks=4000)
    check_maps(rep, 1)
    check_poses(rep, 1)


@pytest.mark.parametrize('k', [4, 8])
def test_batched_objects_640x480(k):
    movers = scene.batch_objects(k, 4000 + k)
    rep = lockstep(movers, W, H, K640, 4, seed=4, bg_res=512, batch=True, voxel_blocks=4000)
    check_maps(rep, k)
    check_poses(rep, k)


def test_batched_plain_and_per_object_calls_agree():
    """The batched launches, the one-object-per-launch form and per-object
    mf_track_object calls compute every object's rows and FP64 sums in the
    same order at the same launch shape, so poses, rows and keys are
    bit-identical across the three.  The objects' rows per thread are set
    alike in the three contexts: by default a batch of objects takes another
    shape than a single object, which changes the FP32 partial sums' order."""
    movers = scene.batch_objects(4, 4004)
    sc = scene.Scene(seed=4, width=W, height=H, K=tuple(K640), moving=movers)
    P0 = sc.camera_pose(0).astype(np.float32)
    ctx = []
    for batch in (True, False, True):
        g = midfusion.Kfusion(W, H, P0, bg_size=5.12, bg_res=512, use_rgb=True, batch_objects=batch)
        for lvl, v in enumerate((8, 8, 4)):
            g.set_tuning(g.TUNE_OBJ_ROWS_PER_THREAD_L0 + lvl, v)
        for i, m in enumerate(movers):
            g.add_object(m.vol_size, m.vol_res, m.pose(0).astype(np.float32))
            g.set_moving(i + 1, True)
        ctx.append(g)
    for f in range(4):
        d, rgb, lab = sc.render(f)
        masks = scene.masks_for_labels(lab, len(movers))
        for n, g in enumerate(ctx):
            for i, m in enumerate(masks):
                g.set_mask(i, m)
            g.preprocessing(d, rgb)
            g.track_camera(K640, f)
            if n == 2 and f > 0:
                for i in range(1, 1 + len(movers)):
                    g.track_object(i, K640)
            g.track_objects(K640, f)
            g.integration(K640, f)
            g.raycasting(K640, f)
            g.sync()
        a = ctx[0]
        for b in ctx[1:]:
            for oid in [-1] + list(range(1, 1 + len(movers))):
                assert np.array_equal(a.getPose(oid), b.getPose(oid)), (f, oid)
            for oid in range(1 + len(movers)):
                assert np.array_equal(a.trackresult(oid), b.trackresult(oid)), (f, oid)
                assert np.array_equal(a.keys(oid), b.keys(oid)), (f, oid)
    for g in ctx:
        g.close()


def test_track_object_arguments():
    movers = scene.batch_objects(2, 4002)
    sc = scene.Scene(seed=4, width=320, height=240, K=(262.5, 262.5, 159.5, 119.5), moving=movers)
    K = np.array([262.5, 262.5, 159.5, 119.5], np.float32)
    g = midfusion.Kfusion(320, 240, sc.camera_pose(0).astype(np.float32), bg_res=256)
    for i, m in enumerate(movers):
        g.add_object(m.vol_size, m.vol_res, m.pose(0).astype(np.float32))
    with pytest.raises(midfusion.MFError):
        g.track_object(1, K)          # no camera phase pending
    for f in range(2):
        d, rgb, lab = sc.render(f)
        for i, m in enumerate(scene.masks_for_labels(lab, 2)):
            g.set_mask(i, m)
        g.preprocessing(d, rgb)
        g.track_camera(K, f)
        if f == 1:
            with pytest.raises(midfusion.MFError):
                g.track_object(0, K)  # the background is not an object
            with pytest.raises(midfusion.MFError):
                g.track_object(3, K)  # no such object
            g.track_object(2, K)      # static objects too: the call tracks regardless
            with pytest.raises(midfusion.MFError):
                g.track_object(2, K)  # once per frame
        g.track_objects(K, f)
        g.integration(K, f)
        g.raycasting(K, f)
    g.sync()
    g.close()


def test_batched_device_masks_match_per_object_calls():
    """mf_objects_set_masks_device copies every listed object's mask in one
    launch: equal to mf_object_set_mask_device per object, float4 and scalar
    paths (a pointer 4 bytes off alignment), a null entry clears that
    object's mask, bad ids are MF_E_ARG."""
    import os
    from conftest import load_module, PKG
    hiprt = load_module('mf_hiprt', os.path.join(PKG, 'hiprt.py'))
    w, h = 160, 120
    rng = np.random.default_rng(7)
    masks = [rng.choice([-1.0, 0.0, 0.5, 1.0], size=(h, w)).astype(np.float32) for _ in range(4)]
    bufs = [hiprt.DeviceBuffer(np.concatenate([[0.0], m.reshape(-1)]).astype(np.float32)) for m in masks]
    g = midfusion.Kfusion(w, h, np.eye(4, dtype=np.float32), bg_res=64)
    for _ in range(3):
        g.add_object(0.256, 64, np.eye(4, dtype=np.float32))
    # element 0 of each buffer is padding: the mask starts 4 bytes in, off
    # 16-byte alignment, so the launch takes its scalar path
    g.set_masks_device([0, 1, 2, 3], [b.ptr.value + 4 for b in bufs])
    g.sync()
    for i in range(4):
        assert np.array_equal(g.get_mask(i).reshape(h, w), masks[i]), i
    al = [hiprt.DeviceBuffer(m) for m in masks]
    g.set_masks_device([3, 2, 1, 0], [b.ptr.value for b in al])   # aligned: float4 copies, permuted
    g.sync()
    for i in range(4):
        assert np.array_equal(g.get_mask(i).reshape(h, w), masks[3 - i]), i
    g.set_masks_device([2], [0])                                  # clears object 2's mask
    with pytest.raises(midfusion.MFError):
        g.set_masks_device([4], [al[0].ptr.value])
    g.close()
