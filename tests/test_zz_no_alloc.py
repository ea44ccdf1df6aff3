"""mf_object_voxels_device (the export entry point that allocates nothing)
against mf_object_voxels, on two calls in a row, the second smaller, through
the C ABI on device buffers."""
import importlib.util
import os

import numpy as np
import pytest

from conftest import scene, midfusion, K320, PKG

pytestmark = pytest.mark.gpu


def _hiprt():
    spec = importlib.util.spec_from_file_location('mf_hiprt', os.path.join(PKG, 'hiprt.py'))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_no_alloc_voxel_export_matches_allocating_export():
    hiprt = _hiprt()
    W, H = 320, 240
    T0 = np.eye(4)
    T0[:3, 3] = np.array([2.56, 2.56, 1.06]) + np.array([0.05, 0.05, 1.0]) - 0.256
    mov = [scene.MovingBox([0.15, 0.12, 0.1], T0, [0.003, -0.002, 0.002, 0.01, 0.012, -0.008], [0.8, 0.3, 0.2],
                           0.512, 128)]
    sc = scene.Scene(seed=3, width=W, height=H, K=tuple(K320), moving=mov)
    g = midfusion.Kfusion(W, H, sc.camera_pose(0).astype(np.float32), bg_res=256)
    g.add_object(mov[0].vol_size, mov[0].vol_res, mov[0].pose(0).astype(np.float32))
    g.set_moving(1, True)
    for f in range(3):
        d, rgb, lab = sc.render(f)
        for i, m in enumerate(scene.masks_for_labels(lab, 1)):
            g.set_mask(i, m)
        g.preprocessing(d, rgb)
        g.tracking(K320, f)
        g.integration(K320, f)
        g.raycasting(K320, f)
    g.sync()
    bg, ob = g.keys(0), g.keys(1)
    assert len(bg) > 1000 and len(ob) > 50
    # first call: background keys plus keys of unallocated blocks; second, on
    # the same buffers: fewer keys of the object's volume
    absent = np.array([(1 << 9) | 5, (7 << 9) | 5], np.uint64)
    calls = [(0, np.concatenate([bg[:3000], absent])), (1, ob[: max(1, len(ob) // 2)])]
    n_max = max(len(k) for _, k in calls)
    d_keys = hiprt.DeviceBuffer(np.zeros(n_max, np.uint64))
    d_out = hiprt.DeviceBuffer(np.zeros(n_max * 512 * 8))
    for oid, keys in calls:
        n = len(keys)
        hiprt.hip().hipMemcpy(d_keys.ptr, keys.ctypes.data, keys.nbytes, 1)
        found = g.voxels_device(oid, d_keys.ptr, n, d_out.ptr)
        want, want_found = g.voxels(oid, keys)
        got = d_out.download(np.float64, (n, 512, 8))
        assert found == want_found, (oid, found, want_found)
        assert np.array_equal(got, want), oid
    with pytest.raises(midfusion.MFError):
        g.voxels_device(0, d_keys.ptr, 65537, d_out.ptr)
    g.close()
