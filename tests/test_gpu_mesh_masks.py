"""GPU parity of the integration-mask generalisation (generalize_label,
segmentation.cpp:10-47) and of the marching-cubes mesh export
(algorithms::marching_cube, meshing.hpp:157-208; Kfusion::dump_mesh,
kfusion.cpp:1555-1588) against the CPU oracle.

* Masks are bit-exact (Mask R-CNN style and rendered masks either side of
  the 3% size threshold, with and without a model mask).
* After a few frames of the pipeline on identical inputs, each object's
  triangles equal the oracle's as a set, bit for bit; the VTK file holds them.
"""
import os

import numpy as np
import pytest

from conftest import scene, midfusion, oracle, K320
import masks

pytestmark = pytest.mark.gpu
W, H = 320, 240


def blobs(seed, n, rmax):
    rng = np.random.default_rng(seed)
    m = np.zeros((H, W), np.uint8)
    yy, xx = np.mgrid[0:H, 0:W]
    for _ in range(n):
        cy, cx, r = rng.integers(0, H), rng.integers(0, W), rng.integers(2, rmax)
        m[(yy - cy) ** 2 + (xx - cx) ** 2 <= r * r] = 255
    m[0, :7] = 255          # pixels on the border
    return m


def test_generalized_masks_bit_exact():
    g = midfusion.Kfusion(W, H, np.eye(4, dtype=np.float32), bg_res=64)
    model = np.zeros((H, W), np.uint8)
    model[200:240, 0:100] = 255
    for seed, n, rmax in ((1, 6, 30), (2, 2, 6), (3, 30, 40)):
        m = blobs(seed, n, rmax)
        for rendered in (False, True):
            for mm in (None, model):
                print(seed, rendered, mm is not None, flush=True)
                g.set_instance_mask(0, m, rendered=rendered, model_mask=mm)
                got = g.get_mask(0)
                want = masks.generalize_label(m, rendered=rendered, model_mask=mm)
                assert np.array_equal(got, want), (seed, rendered, int((got != want).sum()))
    g.close()


def sorted_tris(t):
    t = np.asarray(t, np.float32).reshape(-1, 9)
    return t[np.lexsort(t.T[::-1])]


def test_mesh_matches_oracle(tmp_path):
    c = 2.56
    T0 = np.eye(4)
    T0[:3, 3] = np.array([c, c, c - 1.5]) + np.array([0.05, 0.05, 1.0]) - 0.256
    mov = [scene.MovingBox([0.15, 0.12, 0.1], T0, [0.0] * 6, [0.8, 0.3, 0.2], 0.512, 128)]
    sc = scene.Scene(seed=1, width=W, height=H, K=tuple(K320), moving=mov)
    P0 = sc.camera_pose(0).astype(np.float32)
    gpu = midfusion.Kfusion(W, H, P0, bg_size=5.12, bg_res=256, use_rgb=True)
    orc = oracle.Oracle(W, H, P0, use_rgb=True)
    orc.add_object(5.12, 256, np.eye(4))
    T = mov[0].pose(0).astype(np.float32)
    gpu.add_object(0.512, 128, T)
    orc.add_object(0.512, 128, T)
    for f in range(6):
        print('frame', f, flush=True)
        d, rgb, lab = sc.render(f)
        for i, m in enumerate(scene.masks_for_labels(lab, 1)):
            gpu.set_mask(i, m)
            orc.set_mask(i, m)
        gpu.preprocessing(d, rgb)
        orc.preprocess(d, rgb)
        gpu.tracking(K320, f)
        orc.track(K320, f)
        gpu.sync()
        gpu.setPose(-1, orc.pose().astype(np.float32))
        for i in (0, 1):
            gpu.set_trackresult(i, orc.trackresult(i)[0])
        gpu.integration(K320, f)
        orc.integrate(K320, f)
        gpu.raycasting(K320, f)
        orc.raycast(K320, f)
    gpu.sync()
    for oid in (0, 1):
        a, b = sorted_tris(gpu.mesh(oid)), sorted_tris(orc.mesh(oid))
        print('object', oid, 'triangles', len(a), len(b))
        assert len(b) > 100
        assert a.shape == b.shape and np.array_equal(a, b)
    prefix = str(tmp_path / 'mesh')
    gpu.dump_mesh(prefix)
    txt = open(prefix + '_1.vtk').read().split('\n')
    assert txt[:4] == ['# vtk DataFile Version 1.0', 'vtk mesh generated from KFusion', 'ASCII', 'DATASET POLYDATA']
    npts = int(txt[4].split()[1])
    assert npts == 3 * len(b)
    pts = np.array([[float(v) for v in l.split()] for l in txt[5:5 + npts]], np.float32)
    want = np.array([[float('%g' % v) for v in row] for row in b], np.float32)   # iostream's 6 digits
    assert np.array_equal(sorted_tris(pts.reshape(-1, 9)), sorted_tris(want))
    assert txt[5 + npts].startswith('POLYGONS %d %d' % (len(b), 4 * len(b)))
    assert os.path.exists(prefix + '_0.vtk')
    gpu.close()
