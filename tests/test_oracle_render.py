"""Pins of the oracle's display renderers (rendering/rendering.cpp).

* gs2rgb (commons.h:87-147) is the HLS colour wheel at lightness 0.5 and
  saturation 0.5 (v = 0.75, m = 0.25; the reference rounds 2/3 to 0.6667):
  renderDepthKernel's colours are held to colorsys.hls_to_rgb within one
  level, and near/far pixels to white/black.
* renderIntensityKernel: grey = trunc(255 * I).
* renderVolume's shading from maps, in closed form: for a fronto-parallel
  plane at z = 1 seen from the origin with normals (0, 0, -1), the Lambert
  term is dir = 1/|v|, so a label's pixel is trunc(255 * min(1, dir * c/255
  + 0.1)) with its palette colour c, and the background's trunc(255 *
  min(1, dir + 0.1)).
"""
import colorsys

import numpy as np

from conftest import oracle

W, H = 64, 48
K = np.array([60.0, 60.0, 31.5, 23.5], np.float32)


def make():
    o = oracle.Oracle(W, H, np.eye(4, dtype=np.float32), iters=(0, 0, 0))
    o.add_object(5.12, 64, np.eye(4))
    o.add_object(0.5, 64, np.eye(4))
    return o


def test_depth_and_intensity_renders():
    o = make()
    depth = (np.linspace(0.05, 4.3, W * H).reshape(H, W) * 1000).astype(np.uint16)
    grey = (np.arange(W * H) % 256).reshape(H, W).astype(np.uint8)
    o.preprocess(depth, np.repeat(grey[..., None], 3, 2))
    img = o.render_depth()
    d, g, _ = o.frame_buffers()
    for (y, x) in [(yy, xx) for yy in range(0, H, 5) for xx in range(0, W, 7)]:
        z = d[y, x]
        if z < 0.1:
            assert list(img[y, x]) == [255, 255, 255, 0]
        elif z > 4.0:
            assert list(img[y, x]) == [0, 0, 0, 0]
        else:
            h = np.float32((np.float32(z) - np.float32(0.1)) * (np.float32(1) / np.float32(3.9)))
            ref = np.array(colorsys.hls_to_rgb(float(h) % 1.0, 0.5, 0.5)) * 255
            assert np.all(np.abs(img[y, x, :3].astype(float) - ref) <= 1.01), (z, img[y, x], ref)
    gi = o.render_intensity()
    assert np.array_equal(gi[..., 0], np.trunc(np.float32(255) * g).astype(np.uint8))
    assert (gi[..., 3] == 255).all()


def test_volume_shading_closed_form():
    o = make()
    o.raycast(K, 0)       # records the raycast pose (identity)
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float32)
    v = np.stack([(xs - K[2]) / K[0], (ys - K[3]) / K[1], np.ones_like(xs)], -1).astype(np.float32)
    n = np.zeros_like(v)
    n[..., 2] = -1
    lab = np.zeros((H, W), np.int32)
    lab[10:30, 20:40] = 1
    o.set_maps(v, n, lab)
    pal = np.array([[0, 0, 0, 0], [200, 100, 50, 0]], np.uint8)
    img, ray = o.render_volume(K, np.eye(4, dtype=np.float32), render_color=False, palette=pal)
    assert not ray
    # closed form in float32: dir = dot(n, normalize(-v)) = 1/|v|
    vx, vy, vz = v[..., 0], v[..., 1], v[..., 2]
    dd = (vx * vx + vy * vy) + vz * vz          # dot(-v, -v), float32, left to right
    inv = (np.float32(1) / np.sqrt(dd)).astype(np.float32)
    dirv = np.maximum(np.float32(-1) * ((np.float32(0) - vz) * inv), np.float32(0))
    amb = np.float32(0.1)
    for c in range(3):
        base = np.where(lab == 1, dirv * (np.float32(pal[1, c]) / np.float32(255)) + amb, dirv + amb)
        want = np.trunc(np.clip(base, 0, 1).astype(np.float32) * np.float32(255)).astype(np.uint8)
        assert np.array_equal(img[..., c], want), c
