"""Pins of the CPU oracle (CPU only).

Sources of truth, in order: known answers copied as numbers from the
reference's own unit tests; the reference's header-only octree core compiled
in place (oracle/_ref, when it was built in this container); closed forms and
invariants that a dropped term, wrong sign or transposed index would break.
"""
import numpy as np
import pytest
import scipy.linalg

from conftest import oracle

import ref as refmod

L = oracle.lib(False)
needs_ref = pytest.mark.skipif(not refmod.available(), reason='oracle/_ref not built (needs /root/reference)')


# ---------------------------------------------------------------- KATs
def checkerboard():
    # apps/kfusion/test/preprocessing/interp_unittest.cpp:5-16
    img = np.zeros((64, 64), np.float32)
    for y in range(63):
        for x in range(63):
            img[y, x] = (x & 1) ^ (y & 1)
    return img


@pytest.mark.parametrize('px,py,want', [(31.0, 32.0, 1.0), (24.0, 52.0, 0.0), (24.5, 52.5, 0.5),
                                        (24.25, 52.25, 0.375), (0.82, 0.74, 0.3464)])
def test_bilinear_kat(px, py, want):
    # interp_unittest.cpp:24-52 (ASSERT_FLOAT_EQ: within 4 ulp)
    got = L.orc_bilinear(checkerboard().reshape(-1), 64, px, py)
    assert abs(got - want) <= 4 * np.spacing(np.float32(max(abs(want), 1e-30))) or got == pytest.approx(want, abs=1e-7)


def test_encode_parent_kats():
    # lib/test/octree/octree_unittest.cpp:68-86: parent(parent(encode(112,80,160,5,8)))
    # == encode(96,64,160,3,8), and the next parent == encode(64,64,128,2,8).
    # The parent of a key is its encode at the lower level.
    e = L.orc_encode
    assert e(112, 80, 160, 3, 8) == e(96, 64, 160, 3, 8)
    assert e(96, 64, 160, 2, 8) == e(64, 64, 128, 2, 8)
    assert e(112, 80, 160, 5, 8) & 0x1FF == 5
    # descendant (octree_unittest.cpp:50-62): (110,80,159) at level 5 lies under (96,64,128) at level 3
    mask3 = e(110, 80, 159, 3, 8) & ~0x1FF
    assert mask3 == (e(96, 64, 128, 3, 8) & ~0x1FF)
    assert (e(110, 80, 159, 3, 8) & ~0x1FF) != (e(128, 64, 64, 3, 8) & ~0x1FF)


def unique_kat_keys():
    # lib/test/algorithms/unique_unittest.cpp:72-98 (256^3 tree, levels 7/3/5/6)
    blocks = [(56, 12, 12), (56, 12, 15), (128, 128, 128), (128, 128, 125), (128, 128, 127), (128, 136, 129),
              (128, 136, 127), (136, 128, 136), (128, 240, 136), (128, 241, 136)]
    lv = [7] * 10
    lv[2], lv[3], lv[4] = 3, 5, 6
    keys = np.array([L.orc_encode(x, y, z, l, 8) for (x, y, z), l in zip(blocks, lv)], np.uint64)
    return np.sort(keys)


def test_unique_multiscale_kats():
    k = unique_kat_keys()
    assert L.orc_unique_ms(k.copy(), 10, -1) == 7            # unique_unittest.cpp:113-127
    assert L.orc_unique_ms(k.copy(), 10, 6) == 6             # unique_unittest.cpp:129-143


def test_morton_roundtrip():
    # lib/test/utils/morton_unittest.cpp:38-67 (seeded here)
    rng = np.random.default_rng(0)
    out = np.zeros(3, np.uint32)
    for _ in range(1000):
        x, y, z = (int(v) for v in rng.integers(0, 1 << 21, 3))
        L.orc_unpack(L.orc_morton(x, y, z), out)
        assert tuple(out) == (x, y, z)
    for x in range(0, 64, 7):
        for y in range(64):
            for z in range(0, 64, 5):
                L.orc_unpack(L.orc_morton(x, y, z), out)
                assert tuple(out) == (x, y, z)


# ------------------------------------------------ against the reference core
@needs_ref
def test_keys_match_reference_core():
    R = refmod.lib()
    rng = np.random.default_rng(1)
    for _ in range(2000):
        md = int(rng.integers(4, 12))
        lv = int(rng.integers(0, md + 1))
        x, y, z = (int(v) for v in rng.integers(0, 1 << md, 3))
        assert L.orc_encode(x, y, z, lv, md) == R.ref_encode(x, y, z, lv, md)
        assert L.orc_morton(x, y, z) == R.ref_morton(x, y, z)
    k = unique_kat_keys()
    for lvl in (-1, 3, 5, 6, 7):
        a, b = k.copy(), k.copy()
        na, nb = L.orc_unique_ms(a, 10, lvl), R.ref_unique_ms(b, 10, lvl)
        assert na == nb and np.array_equal(a[:na], b[:nb])


def sphere_tree(res=64, dim=1.28, centre=(0.61, 0.67, 0.63), radius=0.33, mu=0.1, seed=0):
    """Allocate the blocks of a sphere's +-mu band (plus a few random ones)
    in both octrees and write tsdf = clamp((|p-c|-r)/mu), fg = 1 there."""
    R = refmod.lib()
    vs = dim / res
    o = L.orc_octree_create(res, dim)
    r = R.ref_octree_create(res, dim)
    md = int(np.log2(res))
    keys = set()
    c = np.array(centre)
    for bx in range(0, res, 8):
        for by in range(0, res, 8):
            for bz in range(0, res, 8):
                lo = np.array([bx, by, bz]) * vs
                hi = lo + 8 * vs
                q = np.clip(c, lo, hi)
                dmin = np.linalg.norm(q - c)
                far = np.max(np.abs(np.stack([lo, hi]) - c), 0)
                dmax = np.linalg.norm(far)
                if dmin <= radius + mu and dmax >= radius - mu:
                    keys.add(L.orc_encode(bx, by, bz, md - 3, md))
    rng = np.random.default_rng(seed)
    for _ in range(6):
        bx, by, bz = (int(v) * 8 for v in rng.integers(0, res // 8, 3))
        keys.add(L.orc_encode(bx, by, bz, md - 3, md))
    ka = np.array(sorted(keys), np.uint64)
    L.orc_octree_alloc(o, ka.copy(), len(ka))
    R.ref_octree_alloc(r, ka.copy(), len(ka))
    vals = {}
    for key in ka:
        out = np.zeros(3, np.uint32)
        L.orc_unpack(int(key) & ~0x1FF, out)
        bx, by, bz = (int(v) for v in out)
        for z in range(bz, bz + 8):
            for y in range(by, by + 8):
                for x in range(bx, bx + 8):
                    p = np.array([x, y, z]) * vs
                    t = np.float32(np.clip((np.linalg.norm(p - c) - radius) / mu, -1, 1))
                    vals[(x, y, z)] = t
                    L.orc_octree_set(o, x, y, z, t, 1.0)
                    R.ref_octree_set(r, x, y, z, t, 1.0)
    return o, r, ka


@needs_ref
def test_alloc_update_matches_reference_core():
    R = refmod.lib()
    rng = np.random.default_rng(2)
    for res in (64, 256):
        md = int(np.log2(res))
        o = L.orc_octree_create(res, 2.56)
        r = R.ref_octree_create(res, 2.56)
        for frame in range(3):
            n = int(rng.integers(50, 400))
            pts = rng.integers(0, res, (n, 3))
            keys = np.array([L.orc_encode(int(x), int(y), int(z), md - 3, md) for x, y, z in pts], np.uint64)
            L.orc_octree_alloc(o, keys.copy(), n)
            R.ref_octree_alloc(r, keys.copy(), n)
            a = np.zeros(100000, np.uint64)
            b = np.zeros(100000, np.uint64)
            na, nb = L.orc_octree_keys(o, a, 100000), R.ref_octree_keys(r, b, 100000)
            assert na == nb and np.array_equal(a[:na], b[:nb])
            na, nb = L.orc_octree_node_keys(o, a, 100000), R.ref_octree_node_keys(r, b, 100000)
            assert na == nb and np.array_equal(a[:na], b[:nb])
        L.orc_octree_destroy(o)
        R.ref_octree_destroy(r)


@needs_ref
def test_interp_grad_match_reference_core():
    R = refmod.lib()
    o, r, _ = sphere_tree()
    rng = np.random.default_rng(3)
    g1, g2 = np.zeros(3, np.float32), np.zeros(3, np.float32)
    for _ in range(4000):
        p = rng.uniform(-0.5, 63.0, 3).astype(np.float32)
        for f in (0, 1):
            a = L.orc_octree_interp(o, *p, f)
            b = R.ref_octree_interp(r, *p, f)
            assert a == b, (p, f, a, b)
        if (p >= 0).all():
            L.orc_octree_grad(o, *p, g1)
            R.ref_octree_grad(r, *p, g2)
            assert np.array_equal(g1, g2), (p, g1, g2)


@needs_ref
def test_ray_first_block_matches_reference_core():
    R = refmod.lib()
    o, r, _ = sphere_tree()
    rng = np.random.default_rng(4)
    hits = 0
    for _ in range(3000):
        org = rng.uniform(-0.3, 1.6, 3).astype(np.float32)
        d = rng.normal(size=3)
        if rng.uniform() < 0.1:
            d[rng.integers(0, 3)] = 0.0          # axis-parallel rays exercise the epsilon clamp
        d = (d / np.linalg.norm(d)).astype(np.float32)
        nearp, farp = np.float32(0.1), np.float32(rng.uniform(0.5, 4.0))
        a = L.orc_octree_ray_tmin(o, org, d, nearp, farp, np.float32(2.0 ** -6))
        b = R.ref_octree_ray_tmin(r, org, d, nearp, farp)
        assert a == b, (org, d, farp, a, b)
        hits += a > 0
    assert hits > 300


@needs_ref
def test_raycast_march_matches_reference():
    R = refmod.lib()
    res, dim = 64, 1.28
    o, r, ka = sphere_tree(res, dim)
    v = R.ref_volume_create(res, dim)
    R.ref_volume_alloc(v, ka.copy(), len(ka))
    out = np.zeros(3, np.uint32)
    vs = dim / res
    c, rad, mu = np.array((0.61, 0.67, 0.63)), 0.33, 0.1
    for key in ka:
        L.orc_unpack(int(key) & ~0x1FF, out)
        bx, by, bz = (int(q) for q in out)
        for z in range(bz, bz + 8):
            for y in range(by, by + 8):
                for x in range(bx, bx + 8):
                    t = np.float32(np.clip((np.linalg.norm(np.array([x, y, z]) * vs - c) - rad) / mu, -1, 1))
                    R.ref_volume_set(v, x, y, z, t, 1.0)
    rng = np.random.default_rng(5)
    h1, h2 = np.zeros(3, np.float32), np.zeros(3, np.float32)
    hits = 0
    for _ in range(2000):
        org = rng.uniform(0.05, 1.2, 3).astype(np.float32)
        tgt = (c + rng.normal(size=3) * 0.2).astype(np.float32)
        d = (tgt - org) / np.linalg.norm(tgt - org)
        d = d.astype(np.float32)
        a = L.orc_octree_raycast(o, org, d, 0.05, 3.0, mu, vs, 8 * vs, h1)
        b = R.ref_volume_raycast(v, org, d, 0.05, 3.0, mu, vs, 8 * vs, h2)
        assert a == b and np.array_equal(h1, h2), (org, d, a, b)
        hits += a > 0
    assert hits > 200
    # the hit lies on the sphere to within the march's linear-interpolation error
    R.ref_volume_destroy(v)


# ------------------------------------------------------------ closed forms
def test_se3_exp_matches_matrix_exponential():
    rng = np.random.default_rng(6)
    T = np.zeros(16)
    for scale in (1e-5, 1e-3, 0.1, 1.0):
        for _ in range(20):
            x = rng.normal(size=6) * scale
            L.orc_se3_exp(x, T)
            A = np.zeros((4, 4))
            w = x[3:]
            A[:3, :3] = [[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]]
            A[:3, 3] = x[:3]          # translation first (TooN SE3 layout)
            assert np.allclose(T.reshape(4, 4), scipy.linalg.expm(A), atol=1e-12, rtol=1e-10)


def test_solve_is_svd_pseudo_inverse():
    rng = np.random.default_rng(7)
    x, upd = np.zeros(6), np.zeros(16, np.float32)
    for cond in (1e1, 1e4, 1e8):
        for _ in range(10):
            Q, _ = np.linalg.qr(rng.normal(size=(6, 6)))
            s = np.logspace(0, -np.log10(cond), 6)
            C = Q @ np.diag(s) @ Q.T
            b = rng.normal(size=6)
            sums = np.zeros(32)
            sums[1:7] = b
            sums[7:28] = C[np.triu_indices(6)]        # upper triangle, row-major (makeJTJ)
            sums[0], sums[28] = 1.0, 1.0
            L.orc_solve(sums, x, upd)
            keep = s * 1e6 > s.max()
            want = Q @ np.diag(np.where(keep, 1 / s, 0)) @ Q.T @ b
            assert np.allclose(x, want, rtol=1e-6, atol=1e-9 * np.abs(want).max())


def test_inverse_rigid():
    rng = np.random.default_rng(8)
    A, B = np.zeros(16, np.float32), np.zeros(16, np.float32)
    for _ in range(50):
        x = rng.normal(size=6)
        T = np.zeros(16)
        L.orc_se3_exp(x, T)
        A[:] = T
        L.orc_inverse(A, B)
        assert np.allclose(B.reshape(4, 4) @ A.reshape(4, 4), np.eye(4), atol=1e-5)


@pytest.mark.parametrize('fn,k', [(1, 1.345), (1, 0.1), (3, 2.3849), (2, 4.6851)])
def test_robust_weight_is_psi_over_r(fn, k):
    # IRLS weight w(r) = rho'(r)/r for the loss rho (tracking.cpp:1474-1532)
    for r in np.linspace(-5, 5, 101):
        if abs(r) < 1e-3:
            continue
        h = 1e-3
        d = (L.orc_robust_loss(r + h, fn, k) - L.orc_robust_loss(r - h, fn, k)) / (2 * h)
        if fn == 1 and abs(abs(r) - k) < 2 * h:
            continue
        assert L.orc_robust_weight(r, fn, k) == pytest.approx(d / r, rel=2e-2, abs=2e-3)


def test_reduce_layout_is_JtWJ():
    rng = np.random.default_rng(9)
    jw, jh = 37, 23
    n = jw * jh
    res = rng.choice([1, 1, 1, -1, -2, -3, -4, -5], size=n).astype(np.int32)
    err = rng.normal(size=n).astype(np.float32) * 0.5
    J = rng.normal(size=(n, 6)).astype(np.float32)
    out = np.zeros(32)
    L.orc_reduce_rows(res, err, J.reshape(-1), jw, jh, jw, jh, 1, 3, 2.3849, out)
    ok = res == 1
    w = 1.0 / (1.0 + (err[ok].astype(np.float64) / 2.3849) ** 2)
    JW = J[ok].astype(np.float64) * w[:, None]
    JtJ = JW.T @ J[ok].astype(np.float64)
    assert np.allclose(out[1:7], JW.T @ err[ok], rtol=1e-4, atol=1e-3)
    assert np.allclose(out[7:28], JtJ[np.triu_indices(6)], rtol=1e-4, atol=1e-3)
    assert out[28] == ok.sum()
    assert out[29] == (res == -4).sum() and out[30] == (res == -5).sum()
    assert out[31] == ((res > -4) & (res < 1)).sum()


def numpy_bilateral(img, e_d=0.1):
    h, w = img.shape
    g = np.exp(-(np.arange(-2, 3) ** 2) / 32.0)
    out = np.zeros_like(img, np.float64)
    num = np.zeros((h, w))
    den = np.zeros((h, w))
    ys, xs = np.mgrid[0:h, 0:w]
    for i in range(-2, 3):
        for j in range(-2, 3):
            p = img[np.clip(ys + j, 0, h - 1), np.clip(xs + i, 0, w - 1)].astype(np.float64)
            f = g[i + 2] * g[j + 2] * np.exp(-(p - img) ** 2 / (2 * e_d * e_d))
            f = np.where(p > 0, f, 0)
            num += f * p
            den += f
    out = np.where(img == 0, 0, num / np.where(den == 0, 1, den))
    return out


def test_bilateral_against_independent_numpy():
    rng = np.random.default_rng(10)
    img = (1.0 + rng.uniform(0, 0.3, (29, 41))).astype(np.float32)
    img[rng.uniform(size=img.shape) < 0.1] = 0
    out = np.zeros_like(img)
    L.orc_bilateral(out.reshape(-1), img.reshape(-1), 41, 29)
    assert np.allclose(out, numpy_bilateral(img), rtol=1e-5, atol=1e-6)
    const = np.full((16, 16), 1.7, np.float32)
    L.orc_bilateral(out[:16, :16].copy().reshape(-1), const.reshape(-1), 16, 16)
    o2 = np.zeros_like(const)
    L.orc_bilateral(o2.reshape(-1), const.reshape(-1), 16, 16)
    assert np.allclose(o2, 1.7, rtol=1e-6)


def test_half_sample_against_independent_numpy():
    rng = np.random.default_rng(11)
    img = rng.uniform(0.5, 1.5, (30, 44)).astype(np.float32)
    out = np.zeros((15, 22), np.float32)
    L.orc_half_sample(out.reshape(-1), img.reshape(-1), 44, 30)
    blk = img.reshape(15, 2, 22, 2).transpose(0, 2, 1, 3).reshape(15, 22, 4).astype(np.float64)
    c = blk[..., :1]
    m = np.abs(blk - c) < np.float32(0.3)
    want = (blk * m).sum(-1) / m.sum(-1)
    assert np.allclose(out, want, rtol=1e-6)


def test_tsdf_update_closed_form_plane():
    """One integration of a fronto-parallel plane at depth d: tsdf =
    min(1, (d - z) * sqrt(1 + (x/z)^2 + (y/z)^2) / mu) where > -mu."""
    W, H = 64, 48
    K = np.array([60.0, 60.0, 31.5, 23.5], np.float32)
    P = np.eye(4, dtype=np.float32)
    P[:3, 3] = [0.64, 0.64, 0.0]
    o = oracle.Oracle(W, H, P, use_rgb=False)
    o.add_object(1.28, 64, np.eye(4))
    depth = np.full((H, W), 1000, np.uint16)
    o.preprocess(depth, np.zeros((H, W, 3), np.uint8))
    o.set_mask(0, np.ones((H, W), np.float32))
    o.track(K, 0)
    o.integrate(K, 0)
    keys = o.keys(0)
    assert len(keys) > 10
    vs, mu = 0.02, 0.1
    out = np.zeros(3, np.uint32)
    checked = 0
    for key in keys:
        L.orc_unpack(int(key) & ~0x1FF, out)
        bx, by, bz = (int(v) for v in out)
        blk = o.block(0, bx, by, bz)
        for idx in range(512):
            x, y, z = bx + idx % 8, by + (idx // 8) % 8, bz + idx // 64
            pc = np.array([x, y, z]) * vs - P[:3, 3]
            if pc[2] < 1e-4:
                continue
            u, v = 60 * pc[0] / pc[2] + 31.5 + 0.5, 60 * pc[1] / pc[2] + 23.5 + 0.5
            if not (0.501 <= u <= W - 1.501 and 0.501 <= v <= H - 1.501):
                continue      # boundary pixels: float rounding decides
            diff = (1.0 - pc[2]) * np.sqrt(1 + (pc[0] / pc[2]) ** 2 + (pc[1] / pc[2]) ** 2)
            if diff <= -mu:
                assert blk[idx, 1] == 0
                continue
            assert blk[idx, 0] == pytest.approx(min(1.0, diff / mu), abs=2e-6)
            assert blk[idx, 1] == 1
            checked += 1
    assert checked > 1000


# ------------------------------------------------ band march, active list
def camK4(k):
    K4 = np.eye(4)
    K4[0, 0], K4[1, 1], K4[0, 2], K4[1, 2] = k
    return K4


def band_blocks_fp64(depth_m, mask, k, T_wc, T_wo, res, dim, mu, eps=2e-3):
    """buildVolumeAllocationList (alloc_impl.hpp:121-191) in FP64 from its
    definition: for every pixel with depth and mask >= 0, numSteps =
    ceil(band / voxel) samples origin + i * step (i < numSteps) along the
    band of width 2 mu centred on the back-projected pixel centre (x + 0.5,
    y + 0.5), directed towards the camera centre, in the volume frame.
    Returns (required, allowed) block sets: blocks of samples farther than eps
    voxels from every block face, and those blocks together with the blocks
    across any face a sample lies within eps of (where FP32 rounding of the
    incremental march decides)."""
    voxel = float(np.float32(dim) / np.float32(res))
    inv_voxel = np.float32(1) / np.float32(voxel)
    band = np.float32(2) * np.float32(mu)
    n_steps = int(np.ceil(np.float32(band * inv_voxel)))
    kpose = T_wc @ np.linalg.inv(camK4(k))
    inv_wo = np.linalg.inv(T_wo)
    o_cam = (inv_wo @ np.r_[T_wc[:3, 3], 1.0])[:3]
    h, w = depth_m.shape
    ys, xs = np.nonzero((depth_m > 0) & (mask >= 0))
    d = depth_m[ys, xs].astype(np.float64)
    pw = kpose @ np.stack([(xs + 0.5) * d, (ys + 0.5) * d, d, np.ones_like(d)])
    po = (inv_wo @ pw)[:3].T
    dirn = o_cam - po
    dirn /= np.linalg.norm(dirn, axis=1, keepdims=True)
    origin = po - 0.5 * float(band) * dirn
    step = dirn * float(band) / n_steps
    pos = origin[:, None, :] + np.arange(n_steps)[None, :, None] * step[:, None, :]
    c = (pos / voxel).reshape(-1, 3)
    required, allowed = set(), set()
    nb = res // 8
    for cv in c:
        vox = np.floor(cv)
        blk = (vox // 8).astype(int)
        off = cv - 8 * np.round(cv / 8)
        near = np.abs(off) < eps
        alts = [blk]
        for a in range(3):
            if near[a]:
                other = blk.copy()
                other[a] += 1 if off[a] < 0 else -1   # the block on the other side of the face
                alts += [b2 for b2 in (np.where(np.arange(3) == a, other, b1) for b1 in alts)]
        inb = lambda b: bool(np.all(b >= 0) and np.all(b < nb))
        if not near.any() and inb(blk) and np.all(vox >= 0) and np.all(vox < res):
            required.add(tuple(blk))
        for b in alts:
            if inb(b):
                allowed.add(tuple(int(v) for v in b))
    return required, allowed


def block_key(b, res):
    level = int(np.log2(res)) - 3
    return (int(L.orc_morton(int(b[0]), int(b[1]), int(b[2]))) << 9) | level


def rot_xyz(ax, ay, az):
    import scipy.spatial.transform as st
    return st.Rotation.from_euler('xyz', [ax, ay, az]).as_matrix()


@pytest.mark.parametrize('seed', [1, 2, 3, 4])
def test_band_march_keys_brute_force(seed):
    """I1: the oracle's allocated key set of one frame equals the FP64
    brute-force band of every pixel, up to blocks whose samples lie within
    2e-3 voxels of a block face; for the background (identity pose) and an
    object volume under a rotated, translated pose."""
    rng = np.random.default_rng(seed)
    W, H = 8, 8
    k = (8.0, 8.0, 3.5, 3.5)
    T_wc = np.eye(4)
    T_wc[:3, :3] = rot_xyz(*rng.uniform(-0.2, 0.2, 3))
    T_wc[:3, 3] = [0.32, 0.32, -0.05] + rng.uniform(-0.03, 0.03, 3)
    T_wc = T_wc.astype(np.float32).astype(np.float64)
    T_wo = np.eye(4)
    T_wo[:3, :3] = rot_xyz(*rng.uniform(-0.5, 0.5, 3))
    T_wo[:3, 3] = [0.1, 0.1, 0.15] + rng.uniform(-0.02, 0.02, 3)
    T_wo = T_wo.astype(np.float32).astype(np.float64)
    depth = rng.integers(300, 520, (H, W)).astype(np.uint16)
    depth[rng.uniform(size=(H, W)) < 0.1] = 0
    mask = np.where(rng.uniform(size=(H, W)) < 0.15, -1.0, 1.0).astype(np.float32)
    o = oracle.Oracle(W, H, T_wc.astype(np.float32), use_rgb=False)
    o.add_object(0.64, 64, np.eye(4))
    o.add_object(0.40, 64, T_wo.astype(np.float32))
    for i in (0, 1):
        o.set_mask(i, mask)
    o.preprocess(depth, np.zeros((H, W, 3), np.uint8))
    o.integrate(np.array(k, np.float32), 0)
    dm = (depth.astype(np.float32) / np.float32(1000)).astype(np.float64)
    for i, (dim, pose) in enumerate(((0.64, np.eye(4)), (0.40, T_wo))):
        req, allowed = band_blocks_fp64(dm, mask, k, T_wc, pose, 64, dim, 0.1)
        got = set(int(v) for v in o.keys(i))
        want_req = {block_key(b, 64) for b in req}
        want_all = {block_key(b, 64) for b in allowed}
        assert len(want_req) > 20
        assert want_req <= got, (i, len(want_req - got))
        assert got <= want_all, (i, len(got - want_all))


def test_band_march_brute_force_catches_a_half_pixel_shift():
    """The brute force separates plausible mistakes: the band of pixel
    corners (no + 0.5) is not the oracle's key set."""
    W, H = 8, 8
    k = (8.0, 8.0, 3.5, 3.5)
    T_wc = np.eye(4)
    T_wc[:3, 3] = [0.32, 0.32, -0.05]
    depth = np.full((H, W), 410, np.uint16)
    mask = np.ones((H, W), np.float32)
    o = oracle.Oracle(W, H, T_wc.astype(np.float32), use_rgb=False)
    o.add_object(0.64, 64, np.eye(4))
    o.set_mask(0, mask)
    o.preprocess(depth, np.zeros((H, W, 3), np.uint8))
    o.integrate(np.array(k, np.float32), 0)
    got = set(int(v) for v in o.keys(0))
    dm = (depth.astype(np.float32) / np.float32(1000)).astype(np.float64)
    # shifting every pixel by half a pixel is the same as moving the
    # principal point by +0.5
    req, allowed = band_blocks_fp64(dm, mask, (8.0, 8.0, 4.0, 4.0), T_wc, np.eye(4), 64, 0.64, 0.1)
    assert not ({block_key(b, 64) for b in req} <= got and got <= {block_key(b, 64) for b in allowed})


def active_expect(blocks, old, k, T_wc, T_wo, W, H, vs, eps=1e-4):
    """The active list (projective_functor::build_active_list,
    projective_functor.hpp:54-71) and update_block's visibility
    (projective_functor.hpp:73-111) in FP64: a block is updated when it was
    active or its corner projects, by integer truncation, into the frame
    (filter.hpp:42-51); an updated block becomes active iff one of its voxels
    with camera z >= 1e-4 projects (+0.5) into [0.5, W-1.5] x [0.5, H-1.5].
    Returns {block: expected active} over the blocks that no boundary within
    eps decides."""
    Tco = np.linalg.inv(T_wc) @ T_wo
    KT = camK4(k) @ Tco
    g = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing='ij'), -1).reshape(-1, 3)
    out = {}
    for b in blocks:
        bc = np.array(b, float)
        vc = KT @ np.r_[bc * vs, 1.0]
        u, v = vc[0] / vc[2], vc[1] / vc[2]
        if min(abs(u + 1), abs(u - W), abs(v + 1), abs(v - H)) < eps:
            continue
        corner_in = -1 < u < W and -1 < v < H      # trunc(u) in [0, W)
        if not (old[b] or corner_in):
            out[b] = old[b]
            continue
        p = (Tco @ np.c_[(bc + g) * vs, np.ones(512)].T)[:3]
        cam = camK4(k)[:3, :3] @ p
        px, py = cam[0] / cam[2] + 0.5, cam[1] / cam[2] + 0.5
        front = p[2] >= 1e-4
        inside = lambda m: front & (px >= 0.5 + m) & (px <= W - 1.5 - m) & (py >= 0.5 + m) & (py <= H - 1.5 - m)
        strict, loose = inside(eps).any(), inside(-eps).any()
        if strict != loose or np.any(np.abs(p[2] - 1e-4) < 1e-7):
            continue
        out[b] = bool(strict)
    return out


def test_active_list_brute_force():
    """I4: block activity after three integrations (the second and third of
    all-zero depth from other poses, so only visibility changes), oracle
    against the FP64 brute force, with blocks that are inactive, off the
    corner test and yet visible among them (the case the active list
    decides)."""
    W, H = 16, 16
    k = (12.0, 12.0, 7.5, 7.5)
    kf = np.array(k, np.float32)
    poses = []
    for ang, t in ((0.0, [0.32, 0.32, 0.0]), (3.0, [0.32, 0.32, 0.1]), (0.45, [0.20, 0.30, 0.02])):
        T = np.eye(4)
        T[:3, :3] = rot_xyz(0.0, ang, 0.05)
        T[:3, 3] = t
        poses.append(T.astype(np.float32).astype(np.float64))
    o = oracle.Oracle(W, H, poses[0].astype(np.float32), use_rgb=False)
    o.add_object(0.64, 64, np.eye(4))
    o.set_mask(0, np.ones((H, W), np.float32))
    depth0 = np.full((H, W), 400, np.uint16)
    keys = None
    old = None
    xyz = np.zeros(3, np.uint32)
    checked = decided = 0
    for f, T in enumerate(poses):
        o.set_pose(-1, T.astype(np.float32))
        o.preprocess(depth0 if f == 0 else np.zeros((H, W), np.uint16), np.zeros((H, W, 3), np.uint8))
        o.integrate(kf, f)
        if keys is None:
            keys = o.keys(0)
            blocks = []
            for key in keys:
                L.orc_unpack(int(key) & ~0x1FF, xyz)
                blocks.append(tuple(int(v) for v in xyz))
            old = {b: True for b in blocks}     # new blocks start active (octree.hpp:809, 848)
        got = {b: o.L.orc_object_block_active(o.h_, 0, *b) == 1 for b in blocks}
        want = active_expect(blocks, old, k, T, np.eye(4), W, H, 0.01)
        for b, a in want.items():
            assert got[b] == a, (f, b, a)
            checked += 1
            decided += (not old[b]) and a
        assert len(want) > 0.9 * len(blocks)
        old = got
    assert checked > 100 and decided > 0


# ------------------------------------------------------- preprocessing
def pyramid_oracle(depth_u16, rgb, k, use_rgb=True):
    H, W = depth_u16.shape
    P = np.eye(4, dtype=np.float32)
    P[:3, 3] = [2.56, 2.56, 1.0]
    o = oracle.Oracle(W, H, P, use_rgb=use_rgb)
    o.add_object(5.12, 64, np.eye(4))
    o.set_mask(0, np.ones((H, W), np.float32))
    o.preprocess(depth_u16, rgb)
    o.track_camera(np.array(k, np.float32), 0)      # the seeding frame builds the pyramid
    return o


def test_depth2vertex_closed_form():
    """P5: v = d * ((x - cx_l) / fx_l, (y - cy_l) / fy_l, 1) with the
    intrinsics divided by 2^l at level l (preprocessing.cpp:52-71,
    tracking.cpp:803-804), on a constant depth whose pyramid stays constant."""
    W, H = 64, 48
    k = (60.0, 58.0, 31.5, 23.5)
    o = pyramid_oracle(np.full((H, W), 1234, np.uint16), np.full((H, W, 3), 100, np.uint8), k)
    d = float(np.float32(1234) / np.float32(1000))
    for lvl in range(3):
        v = o.level(lvl, 2).astype(np.float64)
        s = 2 ** lvl
        ys, xs = np.mgrid[0:H >> lvl, 0:W >> lvl]
        want = np.stack([d * (xs - k[2] / s) / (k[0] / s), d * (ys - k[3] / s) / (k[1] / s), np.full(xs.shape, d)], -1)
        assert np.allclose(v, want, rtol=0, atol=2e-6), lvl


def numpy_vertex2normal(V):
    """vertex2normalKernel (preprocessing.h:39-81, NegY = false): n =
    normalize(cross(up - down, right - left)) where, for NegY = false, `up`
    is row y+1 and `down` row y-1, left = (x-1, y), right = (x+1, y), all
    clamped to the image; x = INVALID when any of the four has z == 0."""
    H, W, _ = V.shape
    ys, xs = np.mgrid[0:H, 0:W]
    left = V[ys, np.maximum(xs - 1, 0)]
    right = V[ys, np.minimum(xs + 1, W - 1)]
    up = V[np.minimum(ys + 1, H - 1), xs]
    down = V[np.maximum(ys - 1, 0), xs]
    n = np.cross(up - down, right - left)
    n = n / np.linalg.norm(n, axis=-1, keepdims=True)
    bad = (left[..., 2] == 0) | (right[..., 2] == 0) | (up[..., 2] == 0) | (down[..., 2] == 0)
    n[bad] = [-2.0, 0.0, 0.0]
    return n


def test_vertex2normal_tilted_plane():
    """P6: normals of a plane tilted about the x axis (depth depends on the
    row only) point along the plane's normal towards (0, +-sin, +-cos), are
    INVALID next to a hole, and equal an independent numpy restatement of the
    neighbour rule on the oracle's own vertices."""
    W, H = 64, 48
    k = (60.0, 60.0, 31.5, 23.5)
    th = 0.5
    ys = np.arange(H)
    # plane n . X = c with n = (0, sin th, -cos th): depth = c / (n . K^-1 (x, y, 1))
    c = -1.0
    z = c / (np.sin(th) * (ys - k[3]) / k[1] - np.cos(th))
    depth = np.repeat(np.round(z * 1000)[:, None], W, 1).astype(np.uint16)
    depth[20, 30] = 0
    o = pyramid_oracle(depth, np.full((H, W, 3), 100, np.uint8), k)
    V = o.level(0, 2).astype(np.float64)
    N = o.level(0, 3).astype(np.float64)
    want = numpy_vertex2normal(V)
    inv = N[..., 0] == -2
    assert np.array_equal(inv, want[..., 0] == -2)
    assert np.allclose(N[~inv], want[~inv], atol=1e-5)
    # the test reads the four neighbours only: the hole's own pixel keeps a
    # normal, its neighbours are INVALID
    assert not inv[20, 30] and inv[19, 30] and inv[21, 30] and inv[20, 29] and inv[20, 31] and inv.sum() == 4
    nrm = np.array([0.0, np.sin(th), -np.cos(th)])
    ok = ~inv
    ok[16:25, 26:35] = False      # the bilateral filter skips the hole, changing depths next to it
    cosang = np.abs(N[ok] @ nrm)
    assert np.median(cosang) > 0.995 and np.all(np.abs(N[ok][:, 0]) < 1e-6)


def test_gradients_linear_ramp_and_zero_border():
    """P7: central differences x 0.5 of a grey ramp rising 2 levels per
    column: gx = 2/255 * 2^l inside at level l (the bilateral filter and
    half-sampling keep a linear ramp linear), gy = 0, and the one-pixel
    border never written, 0 (preprocessing.cpp:235-246)."""
    W, H = 64, 48
    k = (60.0, 60.0, 31.5, 23.5)
    v = (10 + 2 * np.arange(W)).astype(np.uint8)
    rgb = np.repeat(np.repeat(v[None, :, None], H, 0), 3, 2)
    o = pyramid_oracle(np.full((H, W), 1000, np.uint16), rgb, k)
    for lvl in range(3):
        gx, gy = o.level(lvl, 4), o.level(lvl, 5)
        h, w = gx.shape
        inner = (slice(3, h - 3), slice(3, w - 3))
        assert np.allclose(gx[inner], 2.0 / 255 * 2 ** lvl, rtol=2e-4), lvl
        assert np.abs(gy[inner]).max() < 1e-6
        for g in (gx, gy):
            assert np.all(g[0, :] == 0) and np.all(g[-1, :] == 0) and np.all(g[:, 0] == 0) and np.all(g[:, -1] == 0)


# ------------------------------------------------------------ tracking
def plane_pair(W=64, H=48, k=(60.0, 60.0, 31.5, 23.5), d_mm=1500, use_rgb=False):
    """A constant-depth frame (a fronto-parallel plane) with the camera at
    identity rotation: seed frame 0, then frame 1's pyramid."""
    P = np.eye(4, dtype=np.float32)
    P[:3, 3] = [2.56, 2.56, 1.0]
    o = oracle.Oracle(W, H, P, use_rgb=use_rgb)
    o.add_object(5.12, 64, np.eye(4))
    depth = np.full((H, W), d_mm, np.uint16)
    rgb = np.full((H, W, 3), 120, np.uint8)
    o.set_mask(0, np.ones((H, W), np.float32))
    kf = np.array(k, np.float32)
    o.preprocess(depth, rgb)
    o.track(kf, 0)
    o.preprocess(depth, rgb)
    return o, P, kf


def test_vertex2depth_closed_form():
    """T1: the rendered reference depth is z of inverse(T_w_r) * v for model
    vertices with world z > 0, else 0 (preprocessing.cpp:292-309), and its
    pyramid the robust half-sample of it (tracking.cpp:218-238)."""
    W, H = 64, 48
    o, P, kf = plane_pair(use_rgb=True)
    rng = np.random.default_rng(5)
    cam = np.stack(np.meshgrid(np.linspace(-0.4, 0.4, W), np.linspace(-0.3, 0.3, H)), -1)
    z = 1.2 + 0.1 * rng.uniform(size=(H, W))
    Xc = np.concatenate([cam * z[..., None], z[..., None]], -1)
    Xw = (Xc + P[:3, 3]).astype(np.float32)
    Xw[5, 7, 2] = 0.0          # world z <= 0: no depth
    Xw[9, 3, 2] = -0.5
    nrm = np.zeros((H, W, 3), np.float32)
    nrm[..., 2] = -1
    o.set_maps(Xw, nrm, np.zeros((H, W), np.int32))
    o.track_camera(kf, 1)
    rd = o.level(0, 7).astype(np.float64)
    want = Xw[..., 2].astype(np.float64) - P[2, 3]
    want[Xw[..., 2] <= 0] = 0
    assert np.allclose(rd, want, atol=2e-6)
    assert rd[5, 7] == 0 and rd[9, 3] == 0
    r1 = o.level(1, 7).astype(np.float64)
    blk = rd.reshape(H // 2, 2, W // 2, 2).transpose(0, 2, 1, 3).reshape(H // 2, W // 2, 4)
    c = blk[..., 0]
    m = np.abs(blk - c[..., None]) < 0.3
    assert np.allclose(r1, (blk * m).sum(-1) / m.sum(-1), atol=2e-6)


def test_object_icp_residual_on_a_plane():
    """T4: object ICP rows (track_Obj_ICP_Kernel, tracking.cpp:370-390,
    392-486) of a live plane at depth d against an object model plane at
    d + delta, object frame aligned with the camera: every valid row has
    e = inv_cov * dot(n, o2 - v_ref) = inv_cov * n_z * (-delta) with inv_cov
    = f b / d^2 (the depth variance (d^2 sigma / (f b))^2, b = 0.075), the
    sign opposite to the camera's ICP residual, J = inv_cov * [n, c2v x n]."""
    W, H = 64, 48
    k = (60.0, 60.0, 31.5, 23.5)
    d, delta = 1.5, 0.01
    o, P, kf = plane_pair(W, H, k, d_mm=1500)
    T_wo = np.eye(4, dtype=np.float32)
    T_wo[:3, 3] = P[:3, 3] + np.array([-0.5, -0.5, 1.0], np.float32)
    o.add_object(1.0, 64, T_wo)
    n_live = o.level(0, 3)
    nz = float(np.sign(np.median(n_live[..., 2])))
    ys, xs = np.mgrid[0:H, 0:W]
    Xc = np.stack([(xs - k[2]) / k[0] * (d + delta), (ys - k[3]) / k[1] * (d + delta), np.full(xs.shape, d + delta)], -1)
    nrm = np.zeros((H, W, 3), np.float32)
    nrm[..., 2] = nz
    o.set_object_maps(1, (Xc + P[:3, 3]).astype(np.float32), nrm)
    o.track_camera(kf, 1)
    T_c2_o2 = np.linalg.inv(o.pose().astype(np.float64)) @ T_wo.astype(np.float64)
    o.eval(1, 0, T_c2_o2)
    res, err, J = o.rows(0)
    icp = slice(0, W * H)
    ok = res[icp] == 1
    assert ok.sum() > 0.8 * (W - 4) * (H - 4)
    w = k[0] * 0.075 / d ** 2
    assert np.allclose(err[icp][ok], w * nz * (-delta), rtol=2e-3)
    Jv = J[icp][ok]
    assert np.allclose(Jv[:, :2], 0, atol=1e-3 * w) and np.allclose(Jv[:, 2], w * nz, rtol=1e-4)
    c2v = o.level(0, 2).reshape(-1, 3)[ok].astype(np.float64)
    assert np.allclose(Jv[:, 3:], w * np.cross(c2v, [0, 0, nz]), atol=1e-3 * w)


def test_converged_iteration_breaks_without_update():
    """T8: model maps holding the live plane's own points (camera rotation
    the identity): every residual is 0, so x = 0 and sqrt(s0 / s28) = 0 < 2e-3
    at the first iteration of every level, which breaks without applying
    the update (tracking.cpp:326-331, 1365); the pose is unchanged."""
    o, P, kf = plane_pair()
    V = o.level(0, 2)
    N = o.level(0, 3)
    o.set_maps((V + P[:3, 3]).astype(np.float32), N, np.zeros(V.shape[:2], np.int32))
    o.trace_clear()
    o.track_camera(kf, 1)
    t = o.trace()
    assert [tuple(int(v) for v in r[[1, 2, 5]]) for r in t] == [(2, 0, 1), (1, 0, 1), (0, 0, 1)]
    assert np.array_equal(o.pose(), P.astype(np.float64))


def test_step_back_restores_the_previous_pose():
    """T8: with the live plane 5 mm nearer than the model plane, Gauss-Newton
    reaches the solution within a step; further steps cannot lower the error
    much, and every step-back in the trace puts back the pose of two records
    earlier (previous_pose, tracking.cpp:304-323), and every executed
    iteration records its error; a step-back follows an error increase."""
    o, P, kf = plane_pair(d_mm=1500)
    V = o.level(0, 2)
    N = o.level(0, 3)
    Vm = V.copy()
    Vm[..., 2] += 0.005
    o.set_maps((Vm + P[:3, 3]).astype(np.float32), N, np.zeros(V.shape[:2], np.int32))
    o.trace_clear()
    o.track_camera(kf, 1)
    t = o.trace()
    flags = t[:, 5].astype(int)
    assert abs(o.pose()[2, 3] - (P[2, 3] + 0.005)) < 1e-4
    for i in np.nonzero(flags == 2)[0]:
        assert t[i, 3] > t[i - 1, 3]
        if i >= 2 and t[i, 1] == t[i - 2, 1]:
            assert np.array_equal(t[i, 6:22], t[i - 2, 6:22])


def test_object_without_surface_breaks_on_zero_inliers():
    """T9: an object whose model maps are empty (every pixel a raycast miss)
    has no valid row: each level breaks at its first iteration (flag 3,
    tracking.cpp:674-676) and the pose stays at the static guess."""
    W, H = 64, 48
    o, P, kf = plane_pair(W, H)
    T_wo = np.eye(4, dtype=np.float32)
    T_wo[:3, 3] = P[:3, 3] + np.array([-0.2, -0.2, 1.3], np.float32)
    o.add_object(0.4, 64, T_wo)
    o.set_moving(1, True)
    nrm = np.zeros((H, W, 3), np.float32)
    nrm[..., 0] = -2
    o.set_object_maps(1, np.zeros((H, W, 3), np.float32), nrm)
    lab = np.zeros((H, W), np.int32)
    lab[10:20, 10:20] = 1                      # the last raycast showed the object
    v, n, _ = o.maps()
    o.set_maps(v, n, lab)
    o.trace_clear()
    o.track_camera(kf, 1)
    o.track_objects(kf, 1)
    t = o.trace()
    rec = t[t[:, 0] == 1]
    assert [tuple(int(x) for x in r[[1, 2, 5]]) for r in rec] == [(2, 0, 3), (1, 0, 3), (0, 0, 3)]
    assert np.all(rec[:, 4] == 0)
    assert np.abs(o.pose(1) - T_wo).max() < 1e-6


# ------------------------------------------------------- raycast arbitration
def two_planes(z_bg, z_obj, fg):
    """The background volume holds a plane at camera depth z_bg, an object
    volume a plane at z_obj integrated with foreground value fg; both
    fronto-parallel.  Returns the label image of the raycast."""
    W, H = 32, 24
    k = np.array([30.0, 30.0, 15.5, 11.5], np.float32)
    P = np.eye(4, dtype=np.float32)
    P[:3, 3] = [0.32, 0.32, 0.0]
    o = oracle.Oracle(W, H, P, use_rgb=False)
    o.add_object(0.64, 64, np.eye(4))
    T_wo = np.eye(4, dtype=np.float32)
    T_wo[:3, 3] = [0.16, 0.16, 0.25]
    o.add_object(0.32, 64, T_wo)
    rgb = np.zeros((H, W, 3), np.uint8)
    on, off = np.ones((H, W), np.float32), np.full((H, W), -1.0, np.float32)
    for f, (z, m0, m1) in enumerate(((z_bg, on, off), (z_obj, off, on * fg))):
        o.set_mask(0, m0)
        o.set_mask(1, m1)
        o.preprocess(np.full((H, W), int(round(z * 1000)), np.uint16), rgb)
        o.integrate(k, f)
    o.raycast(k, 2)
    return o.maps()[2]


def test_raycast_arbitration_background_bias_and_fg_gate():
    """R1 (raycastObjectList, rendering.cpp:72-142): the background's hit
    distance carries +0.05 m, so an object surface up to 5 cm behind the
    background surface wins the pixel; 8 cm behind it loses; an object whose
    interpolated foreground is not above 0.5 never wins."""
    centre = (slice(8, 16), slice(10, 22))
    lab = two_planes(0.40, 0.43, 1.0)
    assert np.all(lab[centre] == 1), lab
    lab = two_planes(0.40, 0.48, 1.0)
    assert np.all(lab[centre] == 0), lab
    lab = two_planes(0.40, 0.43, 0.5)
    assert np.all(lab[centre] == 0), lab
    lab = two_planes(0.45, 0.40, 1.0)     # object in front: wins regardless
    assert np.all(lab[centre] == 1), lab
