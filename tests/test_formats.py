"""On-disk formats on either side of the path, through the library's C ABI
(host code, no GPU): SLAMBench .raw frames (RawDepthReader,
interface.h:207-333), TUM trajectory lines (Kfusion::save_poses,
kfusion.cpp:1591-1626) and TUM ground truth (parseGTFile,
vtk-io.h:304-330), each against oracle/formats.py, which is pinned here too:

* .raw: against a file written by the reference's own writer, scene2raw
  (apps/kfusion/thirdparty/scene2raw.cpp), built from its sources into
  oracle/_ref (skipped where that build is absent), whose content the oracle
  predicts byte for byte from the inputs;
* quaternions: against scipy's rotation conversions, on both branches of the
  trace / largest-diagonal method;
* ground-truth poses: against scipy's quaternion-to-matrix conversion,
  normalised and unnormalised quaternions.
"""
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from conftest import ROOT, midfusion
import formats

SCENE2RAW = os.path.join(ROOT, 'oracle', '_ref', 'scene2raw')


def frames(n, w=40, h=30, seed=0):
    rng = np.random.default_rng(seed)
    return [(rng.integers(0, 65535, (h, w), dtype=np.uint16), rng.integers(0, 255, (h, w, 3), dtype=np.uint8))
            for _ in range(n)]


def test_raw_roundtrip(tmp_path):
    fr = frames(3)
    p = str(tmp_path / 'scene.raw')
    formats.write_raw(p, fr)
    raw = midfusion.RawSequence(p)
    assert (raw.width, raw.height, raw.frames) == (40, 30, 3)
    assert np.allclose(raw.k, [531.15, 531.15, 320, 240])
    for i, (d, r) in enumerate(fr):
        dg, rg = raw.read(i)
        assert np.array_equal(dg, d) and np.array_equal(rg, r)
    for i, (d, r) in enumerate(formats.read_raw(p)):
        assert np.array_equal(d, fr[i][0]) and np.array_equal(r, fr[i][1])
    with pytest.raises(midfusion.MFError):
        raw.read(3)
    raw.close()


def test_raw_truncated_and_inconsistent(tmp_path):
    fr = frames(3)
    p = str(tmp_path / 'scene.raw')
    formats.write_raw(p, fr)
    data = open(p, 'rb').read()
    open(p, 'wb').write(data[:-10])          # last frame cut short: two whole frames
    raw = midfusion.RawSequence(p)
    assert raw.frames == 2
    raw.read(1)
    with pytest.raises(midfusion.MFError):
        raw.read(2)
    raw.close()
    bad = bytearray(data)
    fb = 16 + 5 * 40 * 30
    bad[fb:fb + 8] = struct.pack('<II', 41, 30)   # frame 1 header disagrees with frame 0
    open(p, 'wb').write(bytes(bad))
    raw = midfusion.RawSequence(p)
    raw.read(0)
    with pytest.raises(midfusion.MFError):
        raw.read(1)
    raw.close()
    open(p, 'wb').write(b'\x01\x00')
    with pytest.raises(midfusion.MFError):
        midfusion.RawSequence(p)


def write_png(path, rgba):
    h, w, _ = rgba.shape
    rows = b''.join(b'\x00' + rgba[y].tobytes() for y in range(h))

    def chunk(t, d):
        c = struct.pack('>I', len(d)) + t + d
        return c + struct.pack('>I', zlib.crc32(t + d) & 0xffffffff)
    png = b'\x89PNG\r\n\x1a\n' + chunk(b'IHDR', struct.pack('>IIBBBBB', w, h, 8, 6, 0, 0, 0))
    png += chunk(b'IDAT', zlib.compress(rows)) + chunk(b'IEND', b'')
    open(path, 'wb').write(png)


@pytest.mark.skipif(not os.path.exists(SCENE2RAW), reason='reference writer not built (oracle/_ref)')
def test_raw_from_reference_writer(tmp_path):
    d = tmp_path / 'scene'
    d.mkdir()
    rng = np.random.default_rng(3)
    truth = []
    for i in range(2):
        yy, xx = np.mgrid[0:480, 0:640]
        dist = 1.0 + 0.5 * np.sin(xx / 97.0 + i) * np.cos(yy / 61.0) + 0.001 * i
        dist = np.round(dist, 6)
        with open(d / ('scene_00_%04d.depth' % i), 'w') as f:
            f.write(' '.join('%.6f' % v for v in dist.reshape(-1)))
        rgba = rng.integers(0, 255, (480, 640, 4), dtype=np.uint8)
        write_png(str(d / ('scene_00_%04d.png' % i)), rgba)
        truth.append((formats.icl_depth_to_mm(dist), rgba[..., :3]))
    out = str(tmp_path / 'scene.raw')
    subprocess.run([SCENE2RAW, str(d), out], check=True, capture_output=True, timeout=120)
    ref = formats.read_raw(out)
    assert len(ref) == 2
    raw = midfusion.RawSequence(out)
    assert (raw.width, raw.height, raw.frames) == (640, 480, 2)
    for i in range(2):
        assert np.array_equal(ref[i][0], truth[i][0]) and np.array_equal(ref[i][1], truth[i][1])
        dg, rg = raw.read(i)
        assert np.array_equal(dg, ref[i][0]) and np.array_equal(rg, ref[i][1])
    raw.close()


def poses(n, seed=1):
    rng = np.random.default_rng(seed)
    out = []
    rots = list(Rotation.random(n, random_state=seed).as_matrix())
    # rotations with trace <= 0 (each largest-diagonal branch) and near pi
    for axis in np.eye(3):
        rots.append(Rotation.from_rotvec(axis * np.pi).as_matrix())
        rots.append(Rotation.from_rotvec(axis * 2.9).as_matrix())
    rots.append(Rotation.from_rotvec([0.6, -2.1, 1.4]).as_matrix())
    for R in rots:
        T = np.eye(4, dtype=np.float32)
        T[:3, :3] = R
        T[:3, 3] = rng.uniform(-3, 3, 3)
        out.append(T)
    return out


def test_tum_lines_and_quaternion():
    for f, T in enumerate(poses(40)):
        line = midfusion.format_tum(T, f)
        assert line == formats.tum_line(f, T), (line, formats.tum_line(f, T))
        fields = line.split('\t')
        assert len(fields) == 9 and fields[-1] == '\n' and int(fields[0]) == f
        q = np.array([float(v) for v in fields[4:8]])
        ref = Rotation.from_matrix(T[:3, :3].astype(np.float64)).as_quat()
        assert min(np.abs(q - ref).max(), np.abs(q + ref).max()) < 2e-5, (q, ref)
        assert np.allclose(Rotation.from_quat(q).as_matrix(), T[:3, :3], atol=5e-5)
        assert np.allclose([float(v) for v in fields[1:4]], T[:3, 3], rtol=1e-5, atol=1e-5)
    # trace > 0 branch keeps w >= 0; the largest-diagonal branch picks the
    # largest component of that axis positive
    q = formats.rot_quat(poses(0)[0])
    assert q[0] > 0


def test_parse_gt(tmp_path):
    rng = np.random.default_rng(5)
    lines = ['# ground truth trajectory', '# timestamp tx ty tz qx qy qz qw']
    want = []
    for i in range(25):
        q = rng.normal(size=4) * (1.0 if i % 2 else 3.0)     # unnormalised too
        t = rng.uniform(-2, 2, 3)
        vals = [1305031449.7996 + 0.01 * i] + list(t) + list(q)
        lines.append(' '.join('%.4f' % v for v in vals))
        want.append((q, t))
    lines.insert(5, '1305031450.0000  0.1 0.2 0.3 0 0 0 1')     # double space: 9 fields, skipped
    lines.append('1305031451.0 0.1 0.2')                         # short line, skipped
    p = str(tmp_path / 'groundtruth.txt')
    open(p, 'w').write('\n'.join(lines) + '\n')
    P = midfusion.parse_gt(p)
    O = formats.parse_gt(p)
    assert len(P) == len(O) == 25
    for i, (q, t) in enumerate(want):
        assert np.array_equal(P[i], O[i])
        qr = np.array([float('%.4f' % v) for v in q])
        assert np.allclose(P[i][:3, :3], Rotation.from_quat(qr).as_matrix(), atol=2e-6)
        assert np.allclose(P[i][:3, 3], [float('%.4f' % v) for v in t], atol=1e-6)
    with pytest.raises(midfusion.MFError):
        midfusion.parse_gt(str(tmp_path / 'missing.txt'))


def test_tum_fixture_ground_truth_convention():
    """The real-data fixture's poses follow the TUM line order tx ty tz qx qy
    qz qw ("tum pose"): rotation from the quaternion (x, y, z, w), as scipy
    reads it, and the two frames 33 ms apart about 1.2 cm apart."""
    import os
    from scipy.spatial.transform import Rotation
    from conftest import ROOT
    d = np.load(os.path.join(ROOT, 'tests', 'golden', 'tum_fr2_pair.npz'))
    R = Rotation.from_quat([-0.7091, -0.2017, 0.2146, 0.6407]).as_matrix()
    assert np.allclose(d['gt_r'][:3, :3], R, atol=1e-4)
    assert np.allclose(d['gt_r'][:3, 3], [0.0068, -2.7651, 1.6077])
    assert 0.005 < np.linalg.norm(d['gt_l'][:3, 3] - d['gt_r'][:3, 3]) < 0.02
    assert d['depth_r'].dtype == np.uint16 and d['depth_r'].shape == (480, 640) and d['rgb_l'].shape == (480, 640, 3)
