"""Seeded synthetic RGB-D scenes for the per-frame dense path.

Stands in for the reference's dataset readers (apps/kfusion/include/interface.h
RawDepthReader): produces uint16 depth in millimetres (scale 1000, 0 = invalid),
uint8 RGB, ground-truth camera/object poses and per-object generalised
integration masks (instance_seg::generalize_label semantics,
apps/kfusion/src/cpp/segmentation/segmentation.cpp:10-47).

Scene: a 4 m box room centred in a cubic world volume of side V, eight
random spheres/boxes (radius 0.1-0.5 m, Lambertian albedo with a procedural
texture), optional independently moving rigid boxes.  The camera follows a
smooth random trajectory.  Seeds: scene = s, trajectory = s+1, noise = s+2.
"""
import numpy as np

LIGHT = np.array([1.0, 1.0, -1.0]) / np.sqrt(3.0)


def se3_exp(xi):
    """Rotation-vector + translation twist (t first) to a 4x4 matrix (float64)."""
    u, w = np.asarray(xi[:3], float), np.asarray(xi[3:], float)
    th = np.linalg.norm(w)
    W = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    if th < 1e-9:
        A, B, C = 1.0, 0.5, 1.0 / 6.0
    else:
        A = np.sin(th) / th
        B = (1 - np.cos(th)) / th ** 2
        C = (1 - A) / th ** 2
    T = np.eye(4)
    T[:3, :3] = np.eye(3) + A * W + B * W @ W
    T[:3, 3] = (np.eye(3) + B * W + C * W @ W) @ u
    return T


def _rot(axis_angle):
    return se3_exp(np.r_[0, 0, 0, axis_angle])[:3, :3]


class MovingBox:
    """Rigid box (or ellipsoid, ``shape='ellipsoid'``, with semi-axes
    ``half``) moving with a constant body twist per frame.  ``T0`` is the
    pose of its volume's origin corner; the shape is centred in the volume."""

    def __init__(self, half, T0, twist, albedo, vol_size, vol_res, shape='box'):
        self.shape = shape
        self.half = np.asarray(half, float)
        self.T0 = np.asarray(T0, float)
        self.twist = np.asarray(twist, float)
        self.albedo = np.asarray(albedo, float)
        self.vol_size = float(vol_size)
        self.vol_res = int(vol_res)

    def pose(self, f):
        return self.T0 @ se3_exp(self.twist * f)


class Scene:
    def __init__(self, seed=1, width=640, height=480, K=(525.0, 525.0, 319.5, 239.5), world=5.12,
                 n_prims=8, noise=True, invalid_frac=0.0, moving=()):
        self.seed, self.w, self.h = seed, width, height
        self.K = tuple(float(v) for v in K)
        self.V = float(world)
        self.noise, self.invalid_frac = noise, invalid_frac
        rng = np.random.Generator(np.random.MT19937(seed))
        c = self.V / 2
        self.room_lo, self.room_hi = np.full(3, c - 2.0), np.full(3, c + 2.0)
        self.prims = []
        for i in range(n_prims):
            r = rng.uniform(0.1, 0.5)
            # keep primitives inside the room, away from the camera start
            while True:
                p = rng.uniform(c - 1.9 + r, c + 1.9 - r, 3)
                if np.linalg.norm(p - np.array([c, c, c - 1.5])) > r + 0.7:
                    break
            kind = 'sphere' if i % 2 == 0 else 'box'
            self.prims.append((kind, p, r, rng.uniform(0.2, 0.9, 3)))
        self.room_albedo = rng.uniform(0.4, 0.9, 3)
        self.moving = list(moving)
        trng = np.random.Generator(np.random.MT19937(seed + 1))
        self.traj_phase = trng.uniform(0, 2 * np.pi, 6)
        self.traj_freq = trng.uniform(0.02, 0.05, 6)
        self.nrng_seed = seed + 2

    # ---------------------------------------------------------- trajectory
    def camera_pose(self, f):
        """Smooth bounded motion: <= ~2 cm and <= ~1 degree per frame."""
        c = self.V / 2
        ph, fr = self.traj_phase, self.traj_freq
        t = np.array([c, c, c - 1.5]) + 0.3 * np.sin(fr[:3] * f + ph[:3]) - 0.3 * np.sin(ph[:3])
        ang = 0.25 * (np.sin(fr[3:] * f + ph[3:]) - np.sin(ph[3:]))
        T = np.eye(4)
        T[:3, :3] = _rot(ang)
        T[:3, 3] = t
        return T

    # ----------------------------------------------------------- rendering
    def _rays(self, Twc):
        fx, fy, cx, cy = self.K
        xs, ys = np.meshgrid(np.arange(self.w), np.arange(self.h))
        d_cam = np.stack([(xs - cx) / fx, (ys - cy) / fy, np.ones_like(xs, float)], -1).reshape(-1, 3)
        d_w = d_cam @ Twc[:3, :3].T
        return Twc[:3, 3], d_w, d_cam

    @staticmethod
    def _slab(o, d, lo, hi):
        with np.errstate(divide='ignore', invalid='ignore'):
            inv = 1.0 / d
            t0 = (lo - o) * inv
            t1 = (hi - o) * inv
        tmin = np.minimum(t0, t1)
        tmax = np.maximum(t0, t1)
        return tmin.max(-1), tmax.min(-1), tmin.argmax(-1), tmax.argmin(-1), t0, t1

    def render(self, f):
        """Returns depth_mm (uint16 HxW), rgb (uint8 HxWx3), labels (int32 HxW:
        0 background, i+1 moving object i)."""
        Twc = self.camera_pose(f)
        o, d, d_cam = self._rays(Twc)
        n = d.shape[0]
        best_t = np.full(n, np.inf)
        best_n = np.zeros((n, 3))
        best_alb = np.zeros((n, 3))
        best_lab = np.zeros(n, np.int32)
        best_p = np.zeros((n, 3))
        # room: ray starts inside, hits the exit face
        _, tmax, _, amin, t0, t1 = self._slab(o[None], d, self.room_lo, self.room_hi)
        face_t = tmax
        nrm = np.zeros((n, 3))
        idx = np.arange(n)
        sgn = np.where(d[idx, amin] > 0, -1.0, 1.0)
        nrm[idx, amin] = sgn
        best_t, best_n = face_t, nrm
        best_alb[:] = self.room_albedo
        for kind, p, r, alb in self.prims:
            if kind == 'sphere':
                oc = o - p
                b = d @ oc
                cc = oc @ oc - r * r
                dd = d2 = (d * d).sum(-1)
                disc = b * b - dd * cc
                ok = disc > 0
                t = np.where(ok, (-b - np.sqrt(np.maximum(disc, 0))) / d2, np.inf)
                hit = ok & (t > 1e-6) & (t < best_t)
                pts = o + d[hit] * t[hit, None]
                best_n[hit] = (pts - p) / r
            else:
                tmin, tmx, amax, _, _, _ = self._slab(o[None], d, p - r, p + r)
                t = np.where((tmin < tmx) & (tmin > 1e-6), tmin, np.inf)
                hit = t < best_t
                nn = np.zeros((hit.sum(), 3))
                ii = np.arange(hit.sum())
                ax = amax[hit]
                nn[ii, ax] = np.where(d[hit][ii, ax] > 0, -1.0, 1.0)
                best_n[hit] = nn
            best_t = np.where(hit, t, best_t)
            best_alb[hit] = alb
            best_lab[hit] = 0
        for mi, mb in enumerate(self.moving):
            Two = mb.pose(f)
            Tow = np.linalg.inv(Two)
            oo = Tow[:3, :3] @ o + Tow[:3, 3]
            do = d @ Tow[:3, :3].T
            ctr = np.full(3, mb.vol_size / 2)
            if mb.shape == 'ellipsoid':
                # unit sphere in coordinates scaled by the semi-axes
                os_ = (oo - ctr) / mb.half
                ds = do / mb.half
                a = np.einsum('ij,ij->i', ds, ds)
                b = ds @ os_
                cc = os_ @ os_ - 1.0
                disc = b * b - a * cc
                t = np.where(disc > 0, (-b - np.sqrt(np.maximum(disc, 0))) / a, np.inf)
                t = np.where(t > 1e-6, t, np.inf)
                hit = t < best_t
                ph = oo + do[hit] * t[hit, None] - ctr
                nn = ph / (mb.half * mb.half)
                nn /= np.linalg.norm(nn, axis=1, keepdims=True)
            else:
                tmin, tmx, amax, _, _, _ = self._slab(oo[None], do, ctr - mb.half, ctr + mb.half)
                t = np.where((tmin < tmx) & (tmin > 1e-6), tmin, np.inf)
                hit = t < best_t
                nn = np.zeros((hit.sum(), 3))
                ii = np.arange(hit.sum())
                ax = amax[hit]
                nn[ii, ax] = np.where(do[hit][ii, ax] > 0, -1.0, 1.0)
            best_n[hit] = nn @ Two[:3, :3].T
            best_t = np.where(hit, t, best_t)
            # texture in the object frame so it moves with the object
            po = oo + do[hit] * t[hit, None]
            tex = 0.75 + 0.25 * np.sign(np.sin(po[:, 0] * 60) * np.sin(po[:, 1] * 60) * np.sin(po[:, 2] * 60))
            best_alb[hit] = mb.albedo * tex[:, None]
            best_lab[hit] = mi + 1
        pts = o + d * best_t[:, None]
        # world-space texture for static surfaces
        stat = best_lab == 0
        tex = 0.8 + 0.2 * np.sin(pts[stat, 0] * 11) * np.sin(pts[stat, 1] * 13) * np.sin(pts[stat, 2] * 7)
        best_alb[stat] *= tex[:, None]
        shade = 0.1 + np.maximum(0.0, best_n @ -LIGHT) * 0.9
        rgb = np.clip(best_alb * shade[:, None] * 255.0, 0, 255).astype(np.uint8)
        z = best_t * d_cam[:, 2]
        rng = np.random.Generator(np.random.MT19937(self.nrng_seed * 100003 + f))
        if self.noise:
            z = z + rng.normal(0, 1, n) * (0.0012 * z * z)
        depth = np.round(z * 1000.0)
        depth = np.where(np.isfinite(depth) & (depth > 0) & (depth < 65535), depth, 0)
        if self.invalid_frac > 0:
            depth[rng.uniform(0, 1, n) < self.invalid_frac] = 0
        return (depth.reshape(self.h, self.w).astype(np.uint16), rgb.reshape(self.h, self.w, 3),
                best_lab.reshape(self.h, self.w))


def batch_objects(k, seed, world=5.12):
    """K rigid objects for the multi-object workload: boxes and ellipsoids of
    seeded extent 0.05-0.4 m, each in its own volume of 64^3-256^3 voxels at
    2-5 mm, on constant twists, laid out in front of the camera's start."""
    rng = np.random.Generator(np.random.MT19937(seed))
    c = world / 2
    cols = int(np.ceil(np.sqrt(k)))
    out = []
    for i in range(k):
        ext = rng.uniform(0.05, 0.4)
        half = 0.5 * ext * rng.uniform(0.6, 1.0, 3)
        voxel = rng.choice([0.002, 0.003, 0.004, 0.005])
        res = 64
        while res * voxel < 1.3 * ext and res < 256:
            res *= 2
        voxel = max(voxel, 1.3 * ext / res)
        vs = res * voxel
        row, col = divmod(i, cols)
        centre = np.array([c - 0.45 + 0.9 * (col + 0.5) / cols, c - 0.35 + 0.7 * (row + 0.5) / cols,
                           c - 0.2 + rng.uniform(-0.2, 0.2)])
        T0 = np.eye(4)
        T0[:3, 3] = centre - vs / 2
        tw = np.r_[rng.uniform(-0.004, 0.004, 3), rng.uniform(-0.015, 0.015, 3)]
        shape = 'ellipsoid' if i % 2 else 'box'
        out.append(MovingBox(half, T0, tw, rng.uniform(0.3, 0.9, 3), vs, res, shape=shape))
    return out


# OpenCV getStructuringElement(MORPH_ELLIPSE, (5, 5))
ELLIPSE5 = np.array([[0, 0, 1, 0, 0], [1, 1, 1, 1, 1], [1, 1, 1, 1, 1], [1, 1, 1, 1, 1], [0, 0, 1, 0, 0]], bool)


def dilate5(mask):
    m = np.asarray(mask, bool)
    h, w = m.shape
    pad = np.pad(m, 2)
    out = np.zeros_like(m)
    for dy in range(5):
        for dx in range(5):
            if ELLIPSE5[dy, dx]:
                out |= pad[dy:dy + h, dx:dx + w]
    return out


def generalize_label(instance_mask, fg_value=1.0):
    """instance_seg::generalize_label (segmentation.cpp:10-47) for a Mask R-CNN
    style mask: fg_value inside, 0 on the dilated ring, -1 outside."""
    m = np.asarray(instance_mask, bool)
    ring = dilate5(m) & ~m
    out = np.full(m.shape, -1.0, np.float32)
    out[m] = fg_value
    out[ring] = 0.0
    return out


def masks_for_labels(labels, n_objects):
    """Generalised integration masks for the background (id 0) and each object."""
    out = []
    for i in range(n_objects + 1):
        out.append(generalize_label(labels == i))
    return out
