"""Pins of the oracle's generalize_label (segmentation.cpp:10-47): the 5x5
elliptic dilation against scipy's binary dilation with the same structuring
element (border pixels take no part: border_value 0), the foreground values
of rendered masks either side of the 3% threshold, the model-mask OR, and
the package's host helper used to build test masks."""
import numpy as np
from scipy import ndimage

from conftest import scene
import masks


def blobs(h, w, seed, n=6):
    rng = np.random.default_rng(seed)
    m = np.zeros((h, w), np.uint8)
    yy, xx = np.mgrid[0:h, 0:w]
    for _ in range(n):
        cy, cx, r = rng.integers(0, h), rng.integers(0, w), rng.integers(2, 12)
        m[(yy - cy) ** 2 + (xx - cx) ** 2 <= r * r] = 255
    return m


def test_dilation_matches_scipy():
    for seed in range(5):
        m = blobs(48, 64, seed)
        m[0, 5] = m[47, 63] = m[20, 0] = 255        # pixels on the border
        ref = ndimage.binary_dilation(m != 0, structure=masks.ELLIPSE5, border_value=0)
        assert np.array_equal(masks.dilate(m), ref)


def test_generalize_values():
    m = blobs(48, 64, 1)
    g = masks.generalize_label(m)
    d = ndimage.binary_dilation(m != 0, structure=masks.ELLIPSE5)
    assert np.all(g[m != 0] == 1.0) and np.all(g[d & (m == 0)] == 0.0) and np.all(g[~d] == -1.0)
    assert np.array_equal(g, scene.generalize_label(m != 0))
    # rendered: 0.1 under 3% of the image, 0.5 from 3% on
    small = np.zeros((48, 64), np.uint8)
    small[10:13, 10:13] = 255
    assert set(np.unique(masks.generalize_label(small, rendered=True))) == {-1.0, 0.0, np.float32(0.1)}
    n3 = int(np.ceil(0.03 * 48 * 64))
    big = np.zeros(48 * 64, np.uint8)
    big[:n3] = 255
    big = big.reshape(48, 64)
    assert np.float32(0.5) in masks.generalize_label(big, rendered=True)
    big.reshape(-1)[n3 - 1] = 0
    assert np.float32(0.1) in masks.generalize_label(big, rendered=True)
    # the model mask joins the background layer
    model = np.zeros((48, 64), np.uint8)
    model[40:48, 0:20] = 1
    g2 = masks.generalize_label(m, model_mask=model)
    assert np.all(g2[(model != 0) & (m == 0)] == 0.0)
