"""CPU restatement of the on-disk formats on either side of the dense path.

Test infrastructure only (the checker for the library's readers and
writers); nothing here is part of the product.

* SLAMBench .raw sequences: the frame layout RawDepthReader reads
  (apps/kfusion/include/interface.h:236-304) and scene2raw writes
  (apps/kfusion/thirdparty/scene2raw.cpp:170-176): per frame
  {uint32 w, uint32 h, uint16 depth[w*h], uint32 w, uint32 h, uint8 rgb[3*w*h]}.
* scene2raw's ICL-NUIM depth conversion (scene2raw.cpp:75-113), to predict
  the content of a file the reference writer produces.
* TUM trajectory lines of Kfusion::save_poses (kfusion.cpp:1591-1626) and the
  rotation quaternion of getQuaternion (preprocessing.cpp:312-320), by
  Shepperd's trace / largest-diagonal method in float32.
* TUM ground truth parsing of parseGTFile (include/vtk-io.h:304-330) with
  toMatrix4(quat, trans) (utils/math_utils.h:137-146) in float32, its
  diagonal terms in float64 as the reference's `1.0 - a - b` expressions.
"""
import struct

import numpy as np

f32 = np.float32


def write_raw(path, frames):
    """frames: iterable of (depth uint16 [h, w], rgb uint8 [h, w, 3])."""
    with open(path, 'wb') as f:
        for depth, rgb in frames:
            h, w = depth.shape
            f.write(struct.pack('<II', w, h))
            f.write(np.ascontiguousarray(depth, '<u2').tobytes())
            f.write(struct.pack('<II', w, h))
            f.write(np.ascontiguousarray(rgb, np.uint8).tobytes())


def read_raw(path):
    """All whole frames of a .raw file: list of (depth [h, w], rgb [h, w, 3])."""
    data = open(path, 'rb').read()
    if len(data) < 8:
        return []
    w, h = struct.unpack_from('<II', data, 0)
    fb = 16 + 5 * w * h
    out = []
    for i in range(len(data) // fb):
        o = i * fb
        assert struct.unpack_from('<II', data, o) == (w, h)
        depth = np.frombuffer(data, '<u2', w * h, o + 8).reshape(h, w)
        assert struct.unpack_from('<II', data, o + 8 + 2 * w * h) == (w, h)
        rgb = np.frombuffer(data, np.uint8, 3 * w * h, o + 16 + 2 * w * h).reshape(h, w, 3)
        out.append((depth.copy(), rgb.copy()))
    return out


def icl_depth_to_mm(d):
    """scene2raw readDepthFile: ray distance in metres [480, 640] (float64) to
    uint16 z-depth in mm; (u - u0) / fx is float arithmetic (the constants are
    floats), the rest double, truncated on the uint16 store."""
    h, w = d.shape
    u = np.arange(w, dtype=np.int64)[None, :].repeat(h, 0)
    v = np.arange(h, dtype=np.int64)[:, None].repeat(w, 1)
    a = ((u.astype(f32) - f32(319.5)) / f32(481.2)).astype(np.float64)
    b = ((v.astype(f32) - f32(239.5)) / f32(-480.0)).astype(np.float64)
    z = (d * 1000.0) / np.sqrt(a * a + b * b + 1.0)
    return z.astype(np.uint16)


def rot_quat(T):
    """(x, y, z, w) of the rotation block of T, float32 throughout."""
    m = np.asarray(T, f32).reshape(4, 4)
    t = f32(m[0, 0] + m[1, 1] + m[2, 2])
    half = f32(0.5)
    if t > 0:
        t = f32(np.sqrt(f32(t + f32(1))))
        w = f32(half * t)
        t = f32(half / t)
        x = f32(f32(m[2, 1] - m[1, 2]) * t)
        y = f32(f32(m[0, 2] - m[2, 0]) * t)
        z = f32(f32(m[1, 0] - m[0, 1]) * t)
        return [x, y, z, w]
    i = 0
    if m[1, 1] > m[0, 0]:
        i = 1
    if m[2, 2] > m[i, i]:
        i = 2
    j, k = (i + 1) % 3, (i + 2) % 3
    q = [f32(0)] * 3
    t = f32(np.sqrt(f32(f32(f32(m[i, i] - m[j, j]) - m[k, k]) + f32(1))))
    q[i] = f32(half * t)
    t = f32(half / t)
    w = f32(f32(m[k, j] - m[j, k]) * t)
    q[j] = f32(f32(m[j, i] + m[i, j]) * t)
    q[k] = f32(f32(m[k, i] + m[i, k]) * t)
    return [q[0], q[1], q[2], w]


def tum_line(frame, T):
    """Kfusion::save_poses line with the pose's own translation; %g is the
    iostream default for floats."""
    m = np.asarray(T, f32).reshape(4, 4)
    q = rot_quat(m)
    vals = [m[0, 3], m[1, 3], m[2, 3]] + q
    return '%d\t' % frame + ''.join('%g\t' % float(v) for v in vals) + '\n'


def gt_pose(quat, trans):
    q = [f32(v) for v in quat]
    n = f32(f32(f32(f32(q[0] * q[0]) + f32(q[1] * q[1])) + f32(q[2] * q[2])) + f32(q[3] * q[3]))
    s = f32(np.sqrt(f32(2.0 / float(n))))
    a = [f32(v * s) for v in q]
    M = [[f32(a[i] * a[j]) for j in range(4)] for i in range(4)]

    def diag(p, r):
        return f32(1.0 - float(p) - float(r))
    T = np.array([[diag(M[1][1], M[2][2]), M[0][1] - M[2][3], M[0][2] + M[1][3], trans[0]],
                  [M[0][1] + M[2][3], diag(M[0][0], M[2][2]), M[1][2] - M[0][3], trans[1]],
                  [M[0][2] - M[1][3], M[1][2] + M[0][3], diag(M[0][0], M[1][1]), trans[2]],
                  [0, 0, 0, 1]], f32)
    return T


def parse_gt(path):
    """Poses of every line with exactly 8 single-space-separated fields."""
    poses = []
    for line in open(path).read().split('\n'):
        f = line.split(' ')
        if len(f) != 8:
            continue
        v = [f32(x) for x in f[1:]]
        poses.append(gt_pose(v[3:7], v[0:3]))
    return poses
