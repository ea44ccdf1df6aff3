"""The device 6x6 tracking solve against the oracle's pseudo-inverse solve.

`solve` (commons.h:672-696) is TooN's GR_SVD backsub with condition 1e6:
singular values at or below max/1e6 are dropped.  The device first tries a
Gauss-Jordan solve certified by tr(C)*tr(C^-1) < 1e6, and otherwise runs a
wave-parallel Jacobi eigendecomposition.  Both branches are checked here
through mf_solve6 (the same device function the GN iterations call) on
systems built from chosen spectra, against the oracle's serial Jacobi
(`orc_solve`, pinned against numpy's SVD pseudo-inverse in the CPU suite)
and against the closed form Q diag(1/s, kept) Q^T b.  Spectra with a
singular value within a factor 3 of the cut are left out: there the cut
itself is decided by the last bits.
"""
import numpy as np
import pytest

from conftest import midfusion, oracle

pytestmark = pytest.mark.gpu


def make(rng, kind):
    Q, _ = np.linalg.qr(rng.normal(size=(6, 6)))
    if kind == 'well':
        s = np.logspace(0, -rng.uniform(0, 3), 6)
    elif kind == 'rank':
        s = np.logspace(0, -rng.uniform(0, 3), 6)
        s[rng.choice(6, rng.integers(1, 4), replace=False)] *= 1e-10
    elif kind == 'indefinite':
        s = np.logspace(0, -rng.uniform(0, 3), 6)
        s[rng.integers(0, 6)] *= -1
    else:   # 'cond1e8': two values below the cut, none near it
        s = np.array([1.0, 1e-2, 1e-4, 1e-5, 1e-7, 1e-8])
    scale = 10.0 ** rng.uniform(-3, 6)
    s = s * scale
    C = Q @ np.diag(s) @ Q.T
    C = 0.5 * (C + C.T)
    b = rng.normal(size=6) * scale
    return Q, s, C, b


def closed_form(Q, s, b):
    keep = np.abs(s) * 1e6 > np.abs(s).max()
    return Q @ np.diag(np.where(keep, 1 / np.where(keep, s, 1), 0)) @ Q.T @ b


@pytest.fixture(scope='module')
def ctx():
    g = midfusion.Kfusion(64, 48, np.eye(4, dtype=np.float32), bg_size=1.0, bg_res=64)
    yield g
    g.close()


@pytest.mark.parametrize('kind,want_solver', [('well', 1), ('rank', 2), ('indefinite', 2), ('cond1e8', 2)])
def test_solve6_vs_oracle(ctx, kind, want_solver):
    rng = np.random.default_rng({'well': 11, 'rank': 12, 'indefinite': 13, 'cond1e8': 14}[kind])
    L = oracle.lib(True)
    s27, ref, cf = [], [], []
    for _ in range(2000):
        if len(s27) == 400:
            break
        Q, s, C, b = make(rng, kind)
        r = np.abs(s) / np.abs(s).max()
        if np.any((r > 1e-6 / 3) & (r < 3e-6)):
            continue
        sys27 = np.r_[b, C[np.triu_indices(6)]]
        sums = np.zeros(32)
        sums[1:28] = sys27
        sums[0], sums[28] = 1.0, 1.0
        x = np.zeros(6)
        L.orc_solve(sums, x, np.zeros(16, np.float32))
        s27.append(sys27)
        ref.append(x)
        cf.append(closed_form(Q, s, b))
    assert len(s27) == 400
    x_gpu, solver = ctx.solve6(np.array(s27))
    ref, cf = np.array(ref), np.array(cf)
    assert np.all(solver == want_solver), np.bincount(solver)
    scale = np.abs(ref).max(axis=1, keepdims=True)
    err = np.abs(x_gpu - ref) / scale
    print(kind, 'max rel vs oracle', err.max(), 'vs closed form', (np.abs(x_gpu - cf) / scale).max())
    assert err.max() < 1e-8
    # the kept spectrum's condition is at most ~6e4, so rounding C costs
    # far less than this against the exact pseudo-inverse
    assert (np.abs(x_gpu - cf) / scale).max() < 1e-7


def test_solve6_degenerate(ctx):
    """All-zero and diagonal systems: Jacobi starts converged; zero
    eigenvalues are dropped, as the oracle does."""
    L = oracle.lib(True)
    cases = [np.zeros(27)]
    C = np.diag([4.0, 0.0, 1.0, 2.0, 0.0, 3.0])
    cases.append(np.r_[np.arange(1.0, 7.0), C[np.triu_indices(6)]])
    C = np.diag([1.0, 1e-12, 1.0, 1.0, 1.0, 1.0])
    cases.append(np.r_[np.ones(6), C[np.triu_indices(6)]])
    x_gpu, solver = ctx.solve6(np.array(cases))
    assert np.all(solver == 2)
    for c, xg in zip(cases, x_gpu):
        sums = np.zeros(32)
        sums[1:28] = c
        x = np.zeros(6)
        L.orc_solve(sums, x, np.zeros(16, np.float32))
        assert np.array_equal(xg, x), (xg, x)
    assert np.array_equal(x_gpu[1], [0.25, 0, 3, 2, 0, 2])
    assert np.array_equal(x_gpu[2], [1, 0, 1, 1, 1, 1])


def test_volume_capacity_limit(ctx):
    """Voxel pools hold at most 2^21 blocks per volume (the raycast's 32-bit
    TSDF byte offsets); a larger request fails with MF_E_ARG and leaves the
    context usable."""
    with pytest.raises(midfusion.MFError, match='2\\^21'):
        ctx.add_object(2.0, 2048, np.eye(4, dtype=np.float32), capacity=(1 << 21) + 1)
    assert ctx.add_object(0.5, 64, np.eye(4, dtype=np.float32), capacity=64) >= 1
