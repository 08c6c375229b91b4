"""Pins the oracle's Gaussian3f, Cloud::transformInPlace and Merger::merge restatements
(oracle/nicp_oracle.cpp) with checks that do not restate them: the closed form of the
unprojection covariance, the information-form product of Gaussians, R S R^T in FP64, and a
brute force of the merge decisions on small clouds. The reference ships no tests or fixtures
for these functions."""
import numpy as np
import pytest

import oracle as O
from nicp_amd import se3, synth
from tests.helpers import identity16


def _depth(seed, rows=48, cols=64):
    dr, dc, Tgt = synth.pair(seed, rows, cols)
    to_m = lambda d: np.where(d > 0, np.float32(0.001) * d.astype(np.float32), 0).astype(np.float32)
    return to_m(dr), to_m(dc), Tgt


def _camera(rows, cols):
    fx, fy, cx, cy = synth.camera_for(rows, cols)
    cam = O.camera(fx, fy, cx, cy, 0.5, 4.5)
    K = np.array(cam.K, np.float32)
    return cam, K


def _gauss_cloud(depth, cam, K, baseline=0.075, alpha=0.1):
    _, iKRt = O.projector_matrices(K, identity16())
    return O.unproject_gaussians(depth, K, iKRt, cam.min_distance, cam.max_distance, baseline, alpha)


def _cov(g):
    return g[:, O.G_COV].reshape(-1, 3, 3).transpose(0, 2, 1).astype(np.float64)  # column-major -> [i, j]


def _info(g):
    return g[:, O.G_INFO].reshape(-1, 3, 3).transpose(0, 2, 1).astype(np.float64)


@pytest.mark.parametrize("baseline,alpha", [(0.075, 0.1), (0.12, 0.05)])
def test_unprojection_covariance_closed_form(baseline, alpha):
    rows, cols = 48, 64
    d, _, _ = _depth(1, rows, cols)
    cam, K = _camera(rows, cols)
    pts, g, idx = _gauss_cloud(d, cam, K, baseline, alpha)
    assert np.array_equal(O.gaussian_flags(g), np.ones(len(g), np.int32))
    assert np.array_equal(g[:, O.G_MEAN], pts[:, :3])
    r, c = np.nonzero(idx >= 0)
    z = d[r, c].astype(np.float64)
    Km = K.reshape(3, 3, order="F").astype(np.float64)
    fB = baseline * Km[0, 0]
    zvar = alpha * z * z / (fB + z * alpha)
    J0 = np.zeros((len(z), 3, 3))
    J0[:, 0, 0] = z
    J0[:, 1, 1] = z
    J0[:, 0, 2] = c
    J0[:, 1, 2] = r
    J0[:, 2, 2] = 1.0
    J = np.linalg.inv(Km)[None] @ J0
    D = np.zeros((len(z), 3, 3))
    D[:, 0, 0] = 3.0
    D[:, 1, 1] = 3.0
    D[:, 2, 2] = zvar
    expect = J @ D @ J.transpose(0, 2, 1)
    got = _cov(g)[idx[r, c]]
    scale = np.abs(expect).max(axis=(1, 2), keepdims=True)
    assert np.abs(got - expect).max() / 1 <= 2e-6 * scale.max()
    assert np.all(np.abs(got - expect) <= 3e-6 * scale)


def _two_copies(seed, rows=32, cols=40, jitter=0.0005):
    """Two unprojections of one depth image, the second with depth jitter: every pixel holds
    two points, and the merger folds the farther one (or the later one on a tie) into the
    nearer."""
    d, _, _ = _depth(seed, rows, cols)
    cam, K = _camera(rows, cols)
    rng = np.random.default_rng(seed)
    d2 = np.where(d > 0, d + rng.uniform(-jitter, jitter, d.shape).astype(np.float32), 0).astype(np.float32)
    p1, g1, i1 = _gauss_cloud(d, cam, K)
    p2, g2, i2 = _gauss_cloud(d2, cam, K)
    assert np.array_equal(i1 >= 0, i2 >= 0)
    return cam, K, np.concatenate([p1, p2]), np.concatenate([g1, g2]), len(p1)


def test_merge_fuses_information_form():
    """Every kept owner's Gaussian is the product of its own and the folded point's: information
    matrices and vectors add, the point moves to info^-1 * vector (FP64 restatement of the
    textbook product of Gaussians)."""
    rows, cols = 32, 40
    cam, K, pts, g, n1 = _two_copies(3, rows, cols)
    g0 = g.copy()
    pts0 = pts.copy()
    k, col = O.merge(pts, g, identity16(), cam, rows, cols)
    n = len(pts0)
    owners = np.nonzero(col == np.arange(n))[0]
    folded = np.nonzero((col >= 0) & (col != np.arange(n)))[0]
    assert len(owners) == n1 and len(folded) == n1 and k == n1
    # the fold partner of every owner (the other copy of its pixel)
    partner = np.where(owners < n1, owners + n1, owners - n1)
    assert np.array_equal(np.sort(col[partner]), np.sort(owners))
    S1, S2 = _cov(g0)[owners], _cov(g0)[partner]
    m1, m2 = g0[owners, 0:3].astype(np.float64), g0[partner, 0:3].astype(np.float64)
    I1, I2 = np.linalg.inv(S1), np.linalg.inv(S2)
    info = I1 + I2
    vec = (I1 @ m1[..., None] + I2 @ m2[..., None])[..., 0]
    mean = np.linalg.solve(info, vec[..., None])[..., 0]
    # the kept owners in order are the first k entries
    kept = np.nonzero((col < 0) | (col == np.arange(n)))[0]
    assert np.array_equal(kept, owners)
    got_info = _info(g[:k])
    scale = np.abs(info).max(axis=(1, 2), keepdims=True)
    assert np.all(np.abs(got_info - info) <= 1e-5 * scale)
    assert np.abs(pts[:k, :3] - mean).max() < 2e-5
    assert np.array_equal(O.gaussian_flags(g[:k]), np.full(k, 3, np.int32))
    assert np.all(pts[:k, 3] == 1.0)


def _round_half_away(x):
    x = np.asarray(x, np.float64)
    return (np.sign(x) * np.floor(np.abs(x) + 0.5)).astype(np.int64)


def _brute_decisions(pts, KRt, rows, cols, min_d, max_d, thr, max_depth):
    """Merger::merge's decision per point, written from merger.cpp:32-72 with an independent
    z-buffer: the winner of a pixel is the smallest depth, ties to the lower index."""
    f = np.float32
    A = KRt.reshape(4, 4, order="F").astype(np.float32)
    n = len(pts)
    ip = np.zeros((n, 3), np.float32)
    for i in range(3):
        acc = A[i, 0] * pts[:, 0]
        acc = (A[i, 1] * pts[:, 1]).astype(f) + acc
        acc = (A[i, 2] * pts[:, 2]).astype(f) + acc
        acc = (A[i, 3] * pts[:, 3]).astype(f) + acc
        ip[:, i] = acc
    d = ip[:, 2]
    ok = ~((d < f(min_d)) | (d > f(max_d)))
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = (f(1.0) / d).astype(f)
        u = (ip[:, 0] * inv).astype(f)
        v = (ip[:, 1] * inv).astype(f)
    c = np.where(ok, _round_half_away(u), -1)
    r = np.where(ok, _round_half_away(v), -1)
    inimg = ok & (c >= 0) & (c < cols) & (r >= 0) & (r < rows)
    best = {}
    for i in range(n):
        if inimg[i]:
            key = (r[i], c[i])
            if key not in best or d[i] < d[best[key]]:
                best[key] = i
    out = np.full(n, -1, np.int32)
    for i in range(n):
        if d[i] < 0 or d[i] > f(max_depth) or not inimg[i]:
            continue
        t = best[(r[i], c[i])]
        if t == i:
            out[i] = i
        elif abs(float(d[i]) - float(d[t])) < float(f(thr)) and np.float32(abs(d[i] - d[t])) < f(thr):
            out[i] = t
    return out


@pytest.mark.parametrize("seed,thr,max_depth", [(5, 0.1, 10.0), (6, 0.02, 10.0), (7, 0.1, 2.0)])
def test_merge_decisions_bruteforce(seed, thr, max_depth):
    """Two frames of a pair added into one cloud at the ground-truth motion and merged from the
    second camera: decisions, kept count and order against a brute force."""
    rows, cols = 36, 48
    dr, dc, Tgt = _depth(seed, rows, cols)
    cam, K = _camera(rows, cols)
    p1, g1, _ = _gauss_cloud(dr, cam, K)
    p2, g2, _ = _gauss_cloud(dc, cam, K)
    T = np.asarray(Tgt, np.float32).reshape(16)
    O.cloud_transform(T, points=p2, gaussians=g2)
    pts = np.concatenate([p1, p2])
    g = np.concatenate([g1, g2])
    pts0 = pts.copy()
    KRt, _ = O.projector_matrices(K, T)
    expect = _brute_decisions(pts0, KRt, rows, cols, cam.min_distance, cam.max_distance, thr, max_depth)
    k, col = O.merge(pts, g, T, cam, rows, cols, distance_threshold=thr, max_point_depth=max_depth)
    assert np.array_equal(col, expect)
    n = len(pts0)
    kept = np.nonzero((col < 0) | (col == np.arange(n)))[0]
    assert k == len(kept)
    assert (col >= 0).sum() > n // 4 and ((col >= 0) & (col != np.arange(n))).sum() > 0
    # points kept as they are keep their bits; owners moved to their mean
    as_is = col[kept] < 0
    assert np.array_equal(pts[:k][as_is].view(np.uint32), pts0[kept][as_is].view(np.uint32))
    assert np.abs(pts[:k][~as_is, :3] - pts0[kept][~as_is, :3]).max() < 0.05


def test_transform_closed_form():
    rows, cols = 32, 40
    d, _, _ = _depth(9, rows, cols)
    cam, K = _camera(rows, cols)
    pts, g, _ = _gauss_cloud(d, cam, K)
    n = len(pts)
    rng = np.random.default_rng(9)
    nrm = rng.normal(size=(n, 4)).astype(np.float32)
    nrm[:, 3] = 0
    A = rng.normal(size=(n, 3, 3))
    om = np.zeros((n, 4, 4), np.float32)
    om[:, :3, :3] = A @ A.transpose(0, 2, 1)
    st = np.zeros((n, 4, 4), np.float32)
    st[:, :3, :3] = np.linalg.qr(rng.normal(size=(n, 3, 3)))[0]
    st[:, :3, 3] = pts[:, :3]
    st[:, 3, 3] = 1
    om16 = om.transpose(0, 2, 1).reshape(n, 16).copy()
    st16 = st.transpose(0, 2, 1).reshape(n, 16).copy()
    T16 = se3.v2t(np.array([0.3, -0.2, 0.5, 0.1, -0.2, 0.15])).astype(np.float32).reshape(16)
    Tm = T16.reshape(4, 4, order="F").astype(np.float64)
    R, t = Tm[:3, :3], Tm[:3, 3]
    p0, n0, g0 = pts.copy(), nrm.copy(), g.copy()
    O.cloud_transform(T16, points=pts, normals=nrm, stats16=st16, omega_p=om16, gaussians=g)
    assert np.allclose(pts[:, :3], p0[:, :3] @ R.T + t, atol=2e-6) and np.all(pts[:, 3] == 1)
    assert np.allclose(nrm[:, :3], n0[:, :3] @ R.T, atol=2e-6) and np.all(nrm[:, 3] == 0)
    assert np.allclose(g[:, O.G_MEAN], p0[:, :3] @ R.T + t, atol=2e-6)
    S = _cov(g0)
    assert np.allclose(_cov(g), R @ S @ R.T, rtol=1e-5, atol=1e-6 * np.abs(S).max())
    omx = om16.reshape(n, 4, 4).transpose(0, 2, 1).astype(np.float64)
    W = om.astype(np.float64)[:, :3, :3]
    assert np.allclose(omx[:, :3, :3], R @ W @ R.T, rtol=1e-5, atol=1e-5 * np.abs(W).max())
    assert np.all(omx[:, 3, :] == 0) and np.all(omx[:, :, 3] == 0)
    stx = st16.reshape(n, 4, 4).transpose(0, 2, 1).astype(np.float64)
    assert np.allclose(stx[:, :3, :3], R @ st[:, :3, :3].astype(np.float64), atol=2e-6)
    assert np.allclose(stx[:, :3, 3], p0[:, :3] @ R.T + t, atol=2e-6)
    # the identity leaves every bit alone, also the lazy state of the Gaussians
    g1 = g.copy()
    g1[:, 3] = np.array([2], np.int32).view(np.float32)[0]
    keep = g1.copy()
    O.cloud_transform(identity16(), gaussians=g1)
    assert np.array_equal(g1.view(np.uint32), keep.view(np.uint32))


def test_information_form_used_as_is():
    """A Gaussian whose information form is current (bit 1 only) is transformed from the moments
    its information form implies (gaussian.h:108-114), and a merge adds to that form directly."""
    rows, cols = 24, 32
    d, _, _ = _depth(11, rows, cols)
    cam, K = _camera(rows, cols)
    pts, g, _ = _gauss_cloud(d, cam, K)
    n = len(g)
    S = _cov(g)
    info = np.linalg.inv(S)
    vec = (info @ g[:, 0:3, None].astype(np.float64))[..., 0]
    h = g.copy()
    h[:, O.G_INFO] = info.transpose(0, 2, 1).reshape(n, 9).astype(np.float32)
    h[:, O.G_IVEC] = vec.astype(np.float32)
    h[:, O.G_MEAN] = 0
    h[:, O.G_COV] = 0
    h[:, 3] = np.array([2], np.int32).view(np.float32)[0]
    T = np.eye(4, dtype=np.float32)
    T[0, 3] = 0.25
    O.cloud_transform(T.flatten(order="F"), gaussians=h)
    assert np.array_equal(O.gaussian_flags(h), np.ones(n, np.int32))
    assert np.allclose(h[:, 0:3] - [0.25, 0, 0], g[:, 0:3], atol=2e-5)
    assert np.allclose(_cov(h), S, rtol=2e-3, atol=1e-4 * np.abs(S).max())
