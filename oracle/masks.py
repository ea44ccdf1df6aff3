"""CPU restatement of instance_seg::generalize_label
(apps/kfusion/src/cpp/segmentation/segmentation.cpp:10-47, constants
segmentation.h:110-113).  Test infrastructure only.

The instance mask (0 / non-zero) is dilated by OpenCV's 5x5 elliptic
structuring element (getStructuringElement(MORPH_ELLIPSE, Size(5, 5)): the
centre column only in rows 0 and 4, full rows 1-3), pixels outside the image
taking no part; the model mask, if any, is OR-ed in.  Outside the result:
-1; on it but off the instance: 0 (the background layer); on the instance:
1, or for a mask rendered from the volumes 0.1 if it covers under 3% of the
image, else 0.5.
"""
import numpy as np

ELLIPSE5 = np.array([[0, 0, 1, 0, 0],
                     [1, 1, 1, 1, 1],
                     [1, 1, 1, 1, 1],
                     [1, 1, 1, 1, 1],
                     [0, 0, 1, 0, 0]], bool)


def dilate(mask, element=ELLIPSE5):
    m = np.asarray(mask) != 0
    h, w = m.shape
    r = element.shape[0] // 2
    pad = np.pad(m, r)
    out = np.zeros_like(m)
    for dy in range(element.shape[0]):
        for dx in range(element.shape[1]):
            if element[dy, dx]:
                out |= pad[dy:dy + h, dx:dx + w]
    return out


def generalize_label(mask, rendered=False, model_mask=None):
    m = np.asarray(mask) != 0
    h, w = m.shape
    fg = np.float32(1.0)
    if rendered:
        thr = np.float32(np.float32(0.03) * np.float32(w)) * np.float32(h)
        fg = np.float32(0.1) if np.float32(m.sum()) < thr else np.float32(0.5)
    d = dilate(m)
    if model_mask is not None:
        d |= np.asarray(model_mask) != 0
    out = np.full((h, w), -1.0, np.float32)
    out[d & ~m] = 0.0
    out[m] = fg
    return out
