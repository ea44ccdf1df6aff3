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
