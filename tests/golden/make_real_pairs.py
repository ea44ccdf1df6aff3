"""Writes the real-data fixtures of tests/test_gpu_realdata.py from the
reference's own tracking fixtures (apps/kfusion/test/tracking):

tests/golden/tum_fr2_pair.npz: the TUM fr2 walking_halfsphere pair
1341846436.889894 / .922139 and their ground truth in "tum pose", as used by
jacobians_unittest.cpp:55-110, 1023-1052.

Contents: rgb_r, rgb_l (480x640x3 uint8); depth_r, depth_l (480x640 uint16,
millimetres: the TUM 16-bit depth is in 1/5000 m, rounded here to the
1/1000 m the preprocessing input takes); gt_r, gt_l (4x4 float64 world poses
from the "tum pose" lines: tx ty tz qx qy qz qw).  Run in a tree holding the
reference at /root/reference; the test reads only the .npz.
"""
import os

import numpy as np
from PIL import Image

SRC = '/root/reference/apps/kfusion/test/tracking'
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, 'tum_fr2_pair.npz')
OUT_ICL = os.path.join(HERE, 'icl_720_721.npz')


def depth_mm(path):
    """16-bit depth in 1/5000 m, rounded to millimetres."""
    return np.round(np.array(Image.open(path), np.int64) / 5.0).astype(np.uint16)


def rgb(path):
    return np.array(Image.open(path).convert('RGB'), np.uint8)


def quat_pose(t, q):
    x, y, z, w = q
    n = np.sqrt(x * x + y * y + z * z + w * w)
    x, y, z, w = x / n, y / n, z / n, w / n
    T = np.eye(4)
    T[:3, :3] = [[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                 [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                 [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]]
    T[:3, 3] = t
    return T


def main():
    lines = [l.split() for l in open(os.path.join(SRC, 'tum pose')) if l.strip() and not l.startswith('tum')]
    poses = {}
    for v in lines:
        # rgb_ts rgb_file depth_ts depth_file gt_ts tx ty tz qx qy qz qw
        poses[v[0]] = (v[1].split('/')[-1], v[3].split('/')[-1], quat_pose([float(a) for a in v[5:8]],
                                                                          [float(a) for a in v[8:12]]))
    out = {}
    for tag, ts in (('r', '1341846436.889894'), ('l', '1341846436.922139')):
        rgb_f, depth_f, T = poses[ts]
        out['rgb_' + tag] = rgb(os.path.join(SRC, rgb_f))
        out['depth_' + tag] = depth_mm(os.path.join(SRC, depth_f))
        out['gt_' + tag] = T
    np.savez_compressed(OUT, **out)
    print('wrote', OUT, os.path.getsize(OUT), 'bytes')
    icl = {'rgb_r': rgb(os.path.join(SRC, '720.png')), 'depth_r': depth_mm(os.path.join(SRC, '720_depth.png')),
           'rgb_l': rgb(os.path.join(SRC, '721.png')), 'depth_l': depth_mm(os.path.join(SRC, '721_depth.png'))}
    np.savez_compressed(OUT_ICL, **icl)
    print('wrote', OUT_ICL, os.path.getsize(OUT_ICL), 'bytes')


if __name__ == '__main__':
    main()
