"""Pins the CPU oracle (oracle/nicp_oracle.cpp) with checks that do not restate it:
closed forms, brute force on small inputs, library routines (LAPACK, numpy),
invariants and finite differences. The reference ships no tests, fixtures or
golden vectors for this path, and it cannot be compiled here (Eigen3, OpenCV
and FLANN are absent), so these checks are what pin it.
"""
import numpy as np
import pytest

import oracle as O
from nicp_amd import synth
from tests.helpers import identity16, offset16, offset_frame_arrays

CAMS = [(131.25, 131.25, 79.5, 59.5), (262.5, 262.5, 159.5, 119.5), (525.0, 525.0, 319.5, 239.5),
        (1050.0, 1050.0, 639.5, 479.5)]


def K9(fx, fy, cx, cy):
    return np.array([fx, 0, 0, 0, fy, 0, cx, cy, 1], np.float32)


# ---- A1 -------------------------------------------------------------------------------
def test_depth_convert_closed_form():
    raw = np.array([0, 1, 500, 4500, 65535, 1234], np.uint16)
    out = np.empty(len(raw), np.float32)
    O.lib().orc_depth_u16_to_f32(raw.ctypes.data_as(O.C.c_void_p), len(raw), O.C.c_float(0.001),
                                 out.ctypes.data_as(O.C.c_void_p))
    expect = np.where(raw > 0, np.float32(0.001) * raw.astype(np.float32), np.float32(0))
    assert np.array_equal(out.view(np.uint32), expect.astype(np.float32).view(np.uint32))


# ---- A2 -------------------------------------------------------------------------------
@pytest.mark.parametrize("cam", CAMS)
def test_inverse_camera(cam):
    K = K9(*cam)
    iK = np.empty(9, np.float32)
    O.lib().orc_inverse3(K.ctypes.data_as(O.C.c_void_p), iK.ctypes.data_as(O.C.c_void_p))
    Km = K.reshape(3, 3, order="F").astype(np.float64)
    iKm = iK.reshape(3, 3, order="F").astype(np.float64)
    assert np.allclose(Km @ iKm, np.eye(3), atol=2e-4)
    assert np.allclose(iKm, np.linalg.inv(Km), rtol=1e-6, atol=1e-9)
    # a fact of these four cameras: iK(2,2) is exactly 1, so unprojected z equals d
    assert iK[8] == np.float32(1.0)


# ---- A3 / A4 ----------------------------------------------------------------------------
def _frame(seed=0, rows=60, cols=80, **kw):
    return synth.pair(seed, rows, cols, **kw)


def test_unproject_index_and_points():
    dr, _, _ = _frame(1, 60, 80)
    fx, fy, cx, cy = synth.camera_for(60, 80)
    d = np.where(dr > 0, np.float32(0.001) * dr.astype(np.float32), 0).astype(np.float32)
    _, iKRt = O.projector_matrices(K9(fx, fy, cx, cy), identity16())
    pts, idx = O.unproject(d, iKRt, 0.5, 4.5)
    valid = (d != 0) & (d >= 0.5) & (d <= 4.5)
    expect = np.where(valid.ravel(), np.cumsum(valid.ravel()) - 1, -1).reshape(valid.shape)
    assert np.array_equal(idx, expect)
    rr, cc = np.nonzero(valid)
    dd = d[valid].astype(np.float64)
    assert np.array_equal(pts[:, 2], d[valid])          # z == d (iK22 == 1)
    assert np.all(pts[:, 3] == 1.0)
    assert np.allclose(pts[:, 0], (cc - cx) * dd / fx, rtol=0, atol=2e-6)
    assert np.allclose(pts[:, 1], (rr - cy) * dd / fy, rtol=0, atol=2e-6)
    # round trip: projecting at the identity returns every point to its own pixel
    KRt, _ = O.projector_matrices(K9(fx, fy, cx, cy), identity16())
    pidx, pdep = O.project(pts, KRt, 60, 80, 0.5, 4.5)
    assert np.array_equal(pidx, idx)


def test_intervals_closed_form():
    fx, fy, cx, cy = synth.camera_for(480, 640)
    d = np.linspace(0.3, 5.0, 4001).astype(np.float32).reshape(1, -1)
    iv = O.intervals(d, K9(fx, fy, cx, cy), 0.5, 4.5, 0.1)[0, :, 0]
    exact = np.trunc(fx * 0.1 / d[0].astype(np.float64))
    inside = (d[0] >= 0.5) & (d[0] <= 4.5)
    assert np.all(iv[~inside] == -1)
    frac = np.abs(fx * 0.1 / d[0].astype(np.float64) - np.round(fx * 0.1 / d[0].astype(np.float64)))
    safe = inside & (frac > 1e-4)
    assert np.array_equal(iv[safe], exact[safe].astype(np.int32))


# ---- A6 / A7 ----------------------------------------------------------------------------
def test_integral_image_order_and_region():
    rng = np.random.default_rng(0)
    rows, cols = 23, 31
    n = rows * cols
    # integer-valued coordinates: every sum is exact in FP32
    pts = np.concatenate([rng.integers(-8, 9, (n, 3)), np.ones((n, 1))], 1).astype(np.float32)
    idx = np.arange(n, dtype=np.int32).reshape(rows, cols)
    idx[rng.random((rows, cols)) < 0.2] = -1
    ii = O.integral_image(idx, pts).reshape(rows, cols, 10)
    vals = np.zeros((rows, cols, 10), np.float64)
    for r in range(rows):
        for c in range(cols):
            i = idx[r, c]
            if i < 0:
                continue
            x, y, z, w = pts[i].astype(np.float64)
            vals[r, c] = [w * w, x, y, z, x * x, x * y, x * z, y * y, y * z, z * z]
    for _ in range(300):
        c, r = rng.integers(0, cols), rng.integers(0, rows)
        R = rng.integers(1, 12)
        out = O.get_region(ii.ravel(), rows, cols, c - R, c + R, r - R, r + R)
        # documented window: columns (clamp(c-R-1), clamp(c+R-1)], rows likewise (column/row 0 drop)
        x0, x1 = np.clip(c - R - 1, 0, cols - 1), np.clip(c + R - 1, 0, cols - 1)
        y0, y1 = np.clip(r - R - 1, 0, rows - 1), np.clip(r + R - 1, 0, rows - 1)
        brute = vals[y0 + 1:y1 + 1, x0 + 1:x1 + 1].sum(axis=(0, 1))
        assert np.array_equal(out.astype(np.float64), brute)
    # FP32 order: an x-prefix per row, then a y-prefix per column (N/pointintegralimage.cpp:30-43)
    ptsf = np.concatenate([rng.normal(0, 2, (n, 3)), np.ones((n, 1))], 1).astype(np.float32)
    ii2 = O.integral_image(idx, ptsf).reshape(rows, cols, 10)
    v32 = np.zeros((rows, cols, 10), np.float32)
    ok = idx >= 0
    p = ptsf[idx[ok]]
    v32[ok] = np.stack([p[:, 3] * p[:, 3], p[:, 0], p[:, 1], p[:, 2], p[:, 0] * p[:, 0], p[:, 0] * p[:, 1],
                        p[:, 0] * p[:, 2], p[:, 1] * p[:, 1], p[:, 1] * p[:, 2], p[:, 2] * p[:, 2]], 1)
    seq = np.cumsum(np.cumsum(v32, axis=1, dtype=np.float32), axis=0, dtype=np.float32)
    assert np.array_equal(ii2.view(np.uint32), seq.view(np.uint32))
    other = np.cumsum(np.cumsum(v32, axis=0, dtype=np.float32), axis=1, dtype=np.float32)
    assert not np.array_equal(seq.view(np.uint32), other.view(np.uint32))


# ---- A8 / A9 (Eigen computeDirect restatement) -------------------------------------------
def test_eigen_direct_diagonal_and_lapack():
    ev, V = O.eigen3_direct(np.diag([3.0, 1.0, 2.0]))
    assert np.allclose(ev, [1, 2, 3], atol=1e-6)
    assert np.allclose(np.abs(V), np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]]), atol=1e-6)
    rng = np.random.default_rng(1)
    for k in range(500):
        A = rng.normal(size=(3, 3))
        # covariance-like spectra: one small, two larger
        Q, _ = np.linalg.qr(A)
        lam = np.sort(np.abs(rng.normal(size=3))) * 10 ** rng.uniform(-6, 0)
        lam[0] *= rng.uniform(0.0, 0.3)
        C = (Q * lam) @ Q.T
        C = ((C + C.T) / 2).astype(np.float32)
        ev, V = O.eigen3_direct(C)
        w, U = np.linalg.eigh(C.astype(np.float64))
        scale = np.abs(w).max()
        assert np.allclose(ev, w, atol=2e-5 * scale)
        assert np.allclose(V.T.astype(np.float64) @ V, np.eye(3), atol=1e-5)
        res = np.linalg.norm(C.astype(np.float64) @ V - V * ev, axis=0).max()
        assert res < 5e-5 * scale
        gap = (w[1] - w[0]) / scale
        if gap > 1e-2:
            assert abs(abs(U[:, 0] @ V[:, 0]) - 1) < 1e-6 / gap ** 2 + 1e-6


def test_plane_normals_and_fp64_companion():
    # a tilted wall filling the image: every normal equals the plane normal
    rows, cols = 120, 160
    fx, fy, cx, cy = synth.camera_for(rows, cols)
    n = np.array([0.3, -0.2, -1.0])
    n /= np.linalg.norm(n)
    p0 = np.array([0.0, 0.0, 2.0])
    cc, rr = np.meshgrid(np.arange(cols), np.arange(rows))
    ray = np.stack([(cc - cx) / fx, (rr - cy) / fy, np.ones_like(cc, dtype=float)], -1)
    t = (p0 @ n) / (ray @ n)
    depth = np.round(t * 1000 * 4).astype(np.uint16)  # quarter-millimetre units keep quantisation small
    cam = O.camera(fx, fy, cx, cy, 0.5, 12.0)
    sp = O.stats_params(min_image_radius=3)
    f = O.Frame(depth, 0.00025, cam, sp)
    has = np.abs(f.normals[:, :3]).sum(1) > 0
    assert has.mean() > 0.9
    dots = f.normals[has, :3] @ n
    # oriented like n (towards the camera, n . p <= 0). The FP32 integral image cancels
    # catastrophically on thin border windows (the reference's own arithmetic), so the FP32
    # normals match in bulk while the FP64 companion matches everywhere.
    assert np.all(dots > 0.9) and np.quantile(dots, 0.99) > 0.9999 and np.median(dots) > 0.99999
    assert np.median(f.curvature[has]) < 1e-3
    n64, k64, ok = O.stats_f64(depth, 0.00025, cam, sp, f)
    assert np.array_equal(ok, has)
    assert np.all(n64[has] @ n > 0.99999)
    err32 = np.abs(n64[has] - f.normals[has, :3]).max(1)
    assert np.median(err32) < 5e-3 and np.quantile(err32, 0.99) < 0.1


# ---- A10 / A11 ----------------------------------------------------------------------------
def test_curvature_and_information_matrices():
    dr, _, _ = _frame(2, 120, 160)
    fx, fy, cx, cy = synth.camera_for(120, 160)
    cam = O.camera(fx, fy, cx, cy)
    sp = O.stats_params(min_image_radius=3)
    f = O.Frame(dr, 0.001, cam, sp)
    has = np.abs(f.normals[:, :3]).sum(1) > 0
    e = f.eigvals.astype(np.float64)
    k = (e[:, 0] / (((f.eigvals[:, 0] + f.eigvals[:, 1]) + f.eigvals[:, 2]).astype(np.float64) + 1e-9)).astype(np.float32)
    assert np.array_equal(f.curvature, k)
    assert np.all(f.eigvals[:, 0] >= 0)
    for which, Om, flat, nonflat in ((0, f.omega_p, [1000, .001, .001], [1, .1, .1]),
                                     (1, f.omega_n, [1000, .001, .001], [.1, 1, 1])):
        O4 = Om.reshape(-1, 4, 4).transpose(0, 2, 1)
        assert np.all(O4[~has] == 0)
        assert np.all(O4[:, 3, :] == 0) and np.all(O4[:, :, 3] == 0)
        Uh = f.eigvecs[has].reshape(-1, 3, 3).transpose(0, 2, 1).astype(np.float64)
        D = np.where((f.curvature[has] < 0.02)[:, None], np.array(flat), np.array(nonflat))
        expect = np.einsum("nij,nj,nkj->nik", Uh, D, Uh)
        got = O4[has][:, :3, :3].astype(np.float64)
        assert np.allclose(got, expect, rtol=1e-5, atol=1e-5 * np.abs(expect).max())


# ---- A14 -----------------------------------------------------------------------------------
def test_zbuffer_bruteforce_and_ties():
    rng = np.random.default_rng(3)
    rows, cols = 30, 40
    fx, fy, cx, cy = 40.0, 40.0, 19.5, 14.5
    n = 3000
    pts = np.concatenate([rng.uniform(-1, 1, (n, 2)), rng.uniform(0.6, 4.0, (n, 1)), np.ones((n, 1))], 1).astype(np.float32)
    pts[n - 100:] = pts[:100]  # exact duplicates: the lower index must win the tie
    KRt, _ = O.projector_matrices(K9(fx, fy, cx, cy), identity16())
    idx, dep = O.project(pts, KRt, rows, cols, 0.5, 4.5)
    p = pts.astype(np.float64)
    u = fx * p[:, 0] / p[:, 2] + cx
    v = fy * p[:, 1] / p[:, 2] + cy
    x = np.floor(u + 0.5).astype(int)
    y = np.floor(v + 0.5).astype(int)
    inb = (x >= 0) & (x < cols) & (y >= 0) & (y < rows) & (p[:, 2] >= 0.5) & (p[:, 2] <= 4.5)
    near_half = (np.abs(u - np.floor(u) - 0.5) < 1e-4) | (np.abs(v - np.floor(v) - 0.5) < 1e-4)
    best = -np.ones((rows, cols), int)
    bestd = np.full((rows, cols), np.inf)
    for i in range(n):
        if inb[i] and pts[i, 2] < bestd[y[i], x[i]]:
            bestd[y[i], x[i]] = pts[i, 2]
            best[y[i], x[i]] = i
    amb = np.zeros((rows, cols), bool)
    amb[y[inb & near_half], x[inb & near_half]] = True
    assert np.array_equal(idx[~amb], best[~amb])
    assert np.all(idx[idx >= 0] < n - 100)


# ---- A15 / A16 ---------------------------------------------------------------------------
def _pair_frames(seed, rows, cols, **kw):
    dr, dc, T = synth.pair(seed, rows, cols, **kw)
    fx, fy, cx, cy = synth.camera_for(rows, cols)
    cam = O.camera(fx, fy, cx, cy)
    sp = O.stats_params()
    return O.Frame(dr, 0.001, cam, sp), O.Frame(dc, 0.001, cam, sp), T, cam


def test_correspondence_gates_bruteforce_and_partition():
    rows, cols = 240, 320
    fr, fc, T, cam = _pair_frames(4, rows, cols)
    ap = O.align_params()
    KRt, _ = O.projector_matrices(np.array(cam.K, np.float32), T)
    ri, _ = O.project(fr.points, KRt, rows, cols, 0.5, 4.5)
    KRt0, _ = O.projector_matrices(np.array(cam.K, np.float32), identity16())
    ci, _ = O.project(fc.points, KRt0, rows, cols, 0.5, 4.5)
    Tinv = np.linalg.inv(T.reshape(4, 4, order="F").astype(np.float64)).astype(np.float32).flatten(order="F")
    pairs = O.correspondences(fr.cloud(), fc.cloud(), ri, ci, Tinv, ap, 1)
    # float64 brute force of the documented gates, with a margin around each threshold
    Ti = Tinv.reshape(4, 4, order="F").astype(np.float64)
    acc, amb = set(), set()
    for pix in np.flatnonzero((ri.ravel() >= 0) & (ci.ravel() >= 0)):
        a, b = ri.ravel()[pix], ci.ravel()[pix]
        nr, nc = fr.normals[a, :3].astype(np.float64), fc.normals[b, :3].astype(np.float64)
        if not nr.any() or not nc.any():
            continue
        rp = Ti[:3, :3] @ fr.points[a, :3] + Ti[:3, 3]
        rn = Ti[:3, :3] @ nr
        m = [1.0 - np.sum((fc.points[b, :3] - rp) ** 2), nc @ rn - 0.95]
        kr, kc = max(fr.curvature[a], 0.02), max(fc.curvature[b], 0.02)
        ratio = (kr + 1e-5) / (kc + 1e-5)
        m += [ratio - 1 / 1.3, 1.3 - ratio]
        key = (a, b)
        if min(abs(x) for x in m) < 1e-5:
            amb.add(key)
        elif min(m) > 0:
            acc.add(key)
    got = set(map(tuple, pairs.tolist()))
    assert len(acc) > 1000
    assert acc <= got and (got - acc) <= amb
    # row-major order
    pix = [np.flatnonzero((ri.ravel() == a) & (ci.ravel() == b))[0] for a, b in pairs[:200]]
    assert pix == sorted(pix)
    # OpenMP partition quirk (N/correspondencefinderprojective.cpp:41-53): 7 threads over 240 rows
    # only visit 7 * 34 = 238 rows
    p7 = O.correspondences(fr.cloud(), fc.cloud(), ri, ci, Tinv, ap, 7)
    rows_of = {(a, b): np.flatnonzero((ri.ravel() == a) & (ci.ravel() == b))[0] // cols for a, b in pairs}
    expect7 = [tuple(p) for p in pairs.tolist() if rows_of[tuple(p)] < 238]
    assert [tuple(p) for p in p7.tolist()] == expect7


# ---- A17 / A18 -----------------------------------------------------------------------------
def test_linearizer_finite_differences():
    rows, cols = 120, 160
    fr, fc, T, cam = _pair_frames(5, rows, cols)
    ap = O.align_params(robust_kernel=1, inlier_max_chi2=1e30)
    KRt, _ = O.projector_matrices(np.array(cam.K, np.float32), T)
    sp = O.stats_params(min_image_radius=3)
    dr, dc, _ = synth.pair(5, rows, cols)
    fr = O.Frame(dr, 0.001, cam, sp)
    fc = O.Frame(dc, 0.001, cam, sp)
    ri, _ = O.project(fr.points, KRt, rows, cols, 0.5, 4.5)
    ci, _ = O.project(fc.points, O.projector_matrices(np.array(cam.K, np.float32), identity16())[0], rows, cols, 0.5, 4.5)
    Tinv = np.linalg.inv(T.reshape(4, 4, order="F").astype(np.float64)).astype(np.float32).flatten(order="F")
    pairs = O.correspondences(fr.cloud(), fc.cloud(), ri, ci, Tinv, ap, 1)[:400]
    H, b, e, inl = O.linearize_f64(fr.cloud(), fc.cloud(), pairs, Tinv, ap)
    assert inl == len(pairs)

    def energy(dx):
        # chi2 at v2t(dx) * invT, evaluated in float64 from the same float inputs
        Td = O.v2t(dx.astype(np.float32)).reshape(4, 4, order="F").astype(np.float64)
        Tn = Td @ Tinv.reshape(4, 4, order="F").astype(np.float64)
        E = 0.0
        for a, c in pairs:
            rp = Tn[:3, :3] @ fr.points[a, :3] + Tn[:3, 3]
            rn = Tn[:3, :3] @ fr.normals[a, :3]
            ep = rp - fc.points[c, :3]
            en = rn - fc.normals[c, :3]
            Op = fc.omega_p[c].reshape(4, 4, order="F")[:3, :3]
            On = fc.omega_n[c].reshape(4, 4, order="F")[:3, :3]
            E += ep @ Op @ ep + en @ On @ en
        return E

    assert abs(energy(np.zeros(6)) - e) / e < 1e-9
    h = 1e-4
    for k in range(6):
        d = np.zeros(6)
        d[k] = h
        g = (energy(d) - energy(-d)) / (2 * h)
        assert abs(g - 2 * b[k]) < 2e-3 * (abs(2 * b[k]) + np.abs(H).max() * h)
    # the FP32 restatement agrees with the FP64 twin
    H32, b32, e32, i32 = O.linearize(fr.cloud(), fc.cloud(), pairs, Tinv, ap, 1)
    assert np.abs(H32 - H).max() < 1e-4 * np.abs(H).max()
    # OpenMP partition quirk (N/linearizer.cpp:36-43): 3 threads drop C mod 3 correspondences
    pairs7 = pairs[:400 - 2]
    _, _, _, i3 = O.linearize(fr.cloud(), fc.cloud(), pairs7, Tinv, ap, 3)
    assert i3 == len(pairs7) - len(pairs7) % 3


# ---- A19 -----------------------------------------------------------------------------------
def test_ldlt_and_se3():
    rng = np.random.default_rng(7)
    for _ in range(50):
        A = rng.normal(size=(6, 6))
        H = A @ A.T + 6 * np.eye(6)
        rhs = rng.normal(size=6)
        x64, ok = O.ldlt_solve(H, rhs, np.float64)
        assert ok and np.allclose(x64, np.linalg.solve(H, rhs), rtol=1e-10, atol=1e-12)
        x32, ok = O.ldlt_solve(H, rhs, np.float32)
        assert np.allclose(x32, np.linalg.solve(H, rhs), rtol=1e-3, atol=1e-4)
    # semi-definite: the pseudo-inverse of D zeroes the null direction
    H = np.diag([4.0, 2.0, 0.0, 1.0, 3.0, 5.0])
    x, _ = O.ldlt_solve(H, np.ones(6), np.float64)
    assert np.allclose(x, [0.25, 0.5, 0.0, 1.0, 1 / 3, 0.2])
    from scipy.spatial.transform import Rotation
    for _ in range(200):
        q = rng.normal(size=3)
        q *= rng.uniform(0, 0.99) / np.linalg.norm(q)
        x = np.concatenate([rng.normal(size=3), q]).astype(np.float32)
        T = O.v2t(x).reshape(4, 4, order="F").astype(np.float64)
        w = np.sqrt(1 - q @ q)
        assert np.allclose(T[:3, :3], Rotation.from_quat([q[0], q[1], q[2], w]).as_matrix(), atol=1e-6)
        assert np.allclose(T[:3, 3], x[:3])
        assert np.allclose(O.t2v(O.v2t(x)), x, atol=2e-6)


# ---- A13 / A20 ------------------------------------------------------------------------------
def test_align_recovers_motion_and_statistics():
    rows, cols = 480, 640
    ok = 0
    for seed in (2, 3, 4):
        fr, fc, T, cam = _pair_frames(seed, rows, cols)
        ap = O.align_params()
        Tout, res = O.align(fr.cloud(), fc.cloud(), rows, cols, cam, identity16(), ap, 1)
        rot, tr = synth.pose_error(Tout, T)
        ok += rot < 0.1 and tr < 0.002
        Om = np.array(res.omega).reshape(6, 6, order="F")
        assert np.allclose(Om, Om.T)
        assert res.translational_ratio >= 1 and res.rotational_ratio >= 1
        assert res.solution_valid == 1
    assert ok >= 2
    # degenerate: zero information -> invalid; a single plane -> large eigen ratio -> invalid
    r0 = O.compute_statistics(np.zeros((6, 6)), identity16(), O.align_params())
    assert r0.solution_valid == 0
    Hp = np.zeros((6, 6))
    Hp[2, 2] = 1e6
    Hp[3, 3] = Hp[4, 4] = 1e5
    Hp[5, 5] = 1e-3
    Hp[0, 0] = Hp[1, 1] = 1e-3
    rp = O.compute_statistics(Hp, identity16(), O.align_params())
    assert rp.solution_valid == 0 and max(rp.translational_ratio, rp.rotational_ratio) > 50


# A20 against the mathematics: for a well-conditioned H (small Sigma = H^-1) the unscented
# transform through f(s) = t2v(T v2t(s)^-1) (N/aligner.cpp:111-120) is exact to first order, so
# Omega must equal J^-T H J^-1 with J the Jacobian of f at s = 0, and the mean f(0) = t2v(T).
# f is written here with scipy's rotations (v2t: quaternion (x, y, z, w >= 0) -> matrix), not with
# the oracle's own v2t/t2v.
def _f64_v2t(x):
    from scipy.spatial.transform import Rotation
    x = np.atleast_2d(np.asarray(x, np.float64))
    q = x[:, 3:6]
    w = np.sqrt(1.0 - (q * q).sum(1))
    T = np.tile(np.eye(4), (len(x), 1, 1))
    T[:, :3, :3] = Rotation.from_quat(np.column_stack([q, w])).as_matrix()
    T[:, :3, 3] = x[:, :3]
    return T


def _f64_t2v(T):
    from scipy.spatial.transform import Rotation
    q = Rotation.from_matrix(T[:, :3, :3]).as_quat()
    q = np.where(q[:, 3:4] < 0, -q, q)
    return np.column_stack([T[:, :3, 3], q[:, :3]])


def _stat_map(T44, s):
    """f(s) = t2v(T v2t(s)^-1) for rows s."""
    return _f64_t2v(T44[None] @ np.linalg.inv(_f64_v2t(s)))


def _stat_cases():
    rng = np.random.default_rng(11)
    cases = []
    for deg in (0.0, 20.0):
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        q = np.sin(np.radians(deg) / 2) * ax
        x = np.concatenate([rng.uniform(-0.3, 0.3, 3), q])
        A = rng.normal(size=(6, 6))
        U, _ = np.linalg.qr(A)
        H = U @ np.diag(10.0 ** rng.uniform(6, 8, 6)) @ U.T
        cases.append((deg, H, _f64_v2t(x)[0]))
    return cases


@pytest.mark.parametrize("k", [0, 1])
def test_statistics_against_linearised_covariance(k):
    deg, H, T44 = _stat_cases()[k]
    T16 = T44.astype(np.float32).flatten(order="F")
    T44 = T16.reshape(4, 4, order="F").astype(np.float64)   # the transform the oracle sees
    res = O.compute_statistics(H, T16, O.align_params())
    Om = np.array(res.omega).reshape(6, 6, order="F")
    h = 1e-6
    J = np.column_stack([(_stat_map(T44, h * e[None])[0] - _stat_map(T44, -h * e[None])[0]) / (2 * h)
                         for e in np.eye(6)])
    Ji = np.linalg.inv(J)
    expect = Ji.T @ H @ Ji
    rel = np.abs(Om - expect).max() / np.abs(expect).max()
    assert rel < 1e-4, (deg, rel)
    # the mean is f(0) = t2v(T) up to O(Sigma)
    assert np.abs(np.array(res.mean) - _stat_map(T44, np.zeros((1, 6)))[0]).max() < 1e-6
    # ratios: largest over smallest singular value of each 3x3 block of Omega
    for blk, ratio in ((slice(0, 3), res.translational_ratio), (slice(3, 6), res.rotational_ratio)):
        sv = np.abs(np.linalg.eigvalsh(expect[blk, blk]))
        assert abs(ratio - sv.max() / sv.min()) <= 1e-3 * sv.max() / sv.min()
    # sigma points from the rows of the Cholesky factor (a transposed factor) would carry the
    # covariance L^T L instead of L L^T = Sigma: for this H that moves Omega far beyond the bar
    L = np.linalg.cholesky(np.linalg.inv(H))
    assert np.abs(L.T @ L - L @ L.T).max() > 1e-2 * np.abs(L @ L.T).max()


def test_statistics_monte_carlo():
    """Omega against the inverse of the empirical covariance of f(s), s ~ N(0, H^-1)."""
    deg, H, T44 = _stat_cases()[1]
    T16 = T44.astype(np.float32).flatten(order="F")
    T44 = T16.reshape(4, 4, order="F").astype(np.float64)
    res = O.compute_statistics(H, T16, O.align_params())
    Om = np.array(res.omega).reshape(6, 6, order="F")
    rng = np.random.default_rng(5)
    s = rng.multivariate_normal(np.zeros(6), np.linalg.inv(H), size=400_000)
    y = _stat_map(T44, s)
    Cy = np.cov(y.T)
    # whitened: Cy Omega must be the identity up to sampling error (~ sqrt(2/n) = 0.002 per entry)
    M = Cy @ Om
    assert np.abs(M - np.eye(6)).max() < 0.03
    assert np.abs(np.array(res.mean) - y.mean(0)).max() < 1e-5


# ---- DepthImage_scale (N/imageutils.cpp:6-37) --------------------------------------------
def test_depth_scale_closed_form():
    dr, _, _ = _frame(4, 120, 160)
    d = np.where(dr > 0, np.float32(0.001) * dr.astype(np.float32), 0).astype(np.float32)
    assert np.array_equal(O.depth_scale(d, 1), d)              # step 1 is the identity
    for step in (2, 3, 4):
        out = O.depth_scale(d, step)
        rows, cols = d.shape[0] // step, d.shape[1] // step
        blk = d[:rows * step, :cols * step].reshape(rows, step, cols, step).transpose(0, 2, 1, 3).reshape(rows, cols, -1)
        blk = blk.astype(np.float64)
        npos = (blk > 0).sum(-1)
        mu = np.where(npos > 0, blk.sum(-1) / np.maximum(npos, 1), 0)
        var = np.where(npos > 0, (blk ** 2).sum(-1) / np.maximum(npos, 1) - mu ** 2, 0)
        keep = (npos > 0) & (var <= 0.01)
        safe = np.abs(var - 0.01) > 1e-4
        assert np.allclose(out[keep & safe], mu[keep & safe], rtol=1e-6)
        assert np.all(out[~keep & safe] == 0)


# ---- sensor offsets ---------------------------------------------------------------------
def test_align_sensor_offsets_conjugate():
    """AlignerProjective::align with sensor offsets (N/alignerprojective.cpp:36,51): clouds
    converted with the offset S on both sides register to S T S^-1, where T registers the
    clouds without offsets; identity or absent offsets give the plain align bit for bit."""
    rows, cols = 240, 320
    fr, fc, _, cam = _pair_frames(5, rows, cols, max_rot_deg=6.0, max_trans_m=0.06)
    ap = O.align_params()
    T0, r0 = O.align(fr.cloud(), fc.cloud(), rows, cols, cam, identity16(), ap, 1, fp64_solve=True)
    for ro, co in ((None, None), (identity16(), identity16()), (None, identity16())):
        T1, r1 = O.align_offsets(fr.cloud(), fc.cloud(), rows, cols, cam, identity16(), ro, co, ap, 1, fp64_solve=True)
        np.testing.assert_array_equal(T1.view(np.uint32), T0.view(np.uint32))
        assert r1.inliers == r0.inliers
    for S16 in (offset16(12.0, (1, -2, 0.5), (0.0, 0.0, 0.0)), offset16(25.0, (0.3, 1, 0.2), (0.15, -0.4, 0.9))):
        ar = offset_frame_arrays(fr, S16)
        ac = offset_frame_arrays(fc, S16)
        cr, cc = O.cloud_from_arrays(*ar), O.cloud_from_arrays(*ac)
        Ts, rs = O.align_offsets(cr, cc, rows, cols, cam, identity16(), S16, S16, ap, 1, fp64_solve=True)
        S = S16.reshape(4, 4, order="F").astype(np.float64)
        expect = (S @ T0.reshape(4, 4, order="F").astype(np.float64) @ np.linalg.inv(S)).flatten(order="F")
        rot, tr = synth.pose_error(Ts, expect)
        # the same clouds aligned with the offsets ignored
        Tw, _ = O.align(cr, cc, rows, cols, cam, identity16(), ap, 1, fp64_solve=True)
        rw, tw = synth.pose_error(Tw, expect)
        print("offset: %.2e deg %.2e m from S T S^-1 (inliers %d vs %d); without offsets %.3f deg %.4f m" % (
            rot, tr, rs.inliers, r0.inliers, rw, tw))
        # the bar: rotation within 2e-3 deg, translation within 1e-4 m (the rotation error times
        # the offset's 1 m lever arm); ignoring the offsets misses it by an order of magnitude
        assert rot < 2e-3 and tr < 1e-4
        assert rw > 1e-2 or tw > 5e-4
        assert abs(rs.inliers - r0.inliers) <= max(5, r0.inliers // 1000)


# ---- the accuracy ceiling of the reference's algorithm -----------------------------------
def _ceiling_bucket(rot, trans, noise, holes, fp64_normals=False, seeds=range(1000, 1012)):
    """The oracle's AlignerProjective (8 threads) on 640x480 scene pairs rendered as the learned
    workload renders them (scene = seed, motion of the seed), started at the ground truth and at
    the identity. Returns (rotation drift from truth in deg, aligned from truth, aligned from I)."""
    rows, cols = 480, 640
    cam = O.camera(*synth.camera_for(rows, cols))
    sp, ap = O.stats_params(), O.align_params()
    I = identity16()
    drift, a_gt, a_id = [], 0, 0

    def cloud(raw, f):
        if not fp64_normals:
            return f.cloud()
        # normals and curvature from the FP64 companion (no FP32 integral-image cancellation)
        n64, c64, v = O.stats_f64(raw, 0.001, cam, sp, f)
        nrm, curv = f.normals.copy(), f.curvature.copy()
        sel = v & (np.abs(nrm[:, :3]).sum(1) > 0)
        s = np.sign((n64[sel] * nrm[sel, :3]).sum(1))
        nrm[sel, :3] = (n64[sel] * s[:, None]).astype(np.float32)
        curv[sel] = c64[sel].astype(np.float32)
        return O.cloud_from_arrays(f.points, nrm, curv, f.omega_p, f.omega_n)

    for s in seeds:
        T = synth.motion(s, rot, trans)
        dr = synth.render(s, I, rows, cols, stream=10, noise=noise, holes=holes)
        dc = synth.render(s, T, rows, cols, stream=11, noise=noise, holes=holes)
        fr, fc = O.Frame(dr, 0.001, cam, sp), O.Frame(dc, 0.001, cam, sp)
        cr, cc = cloud(dr, fr), cloud(dc, fc)
        for start, is_gt in ((T, True), (I, False)):
            To, _ = O.align(cr, cc, rows, cols, cam, start, ap, 8)
            e = synth.pose_error(To, T) if np.all(np.isfinite(To)) else (np.inf, np.inf)
            ok = e[0] < 0.1 and e[1] < 0.002
            if is_gt:
                drift.append(e[0])
                a_gt += ok
            else:
                a_id += ok
    return np.array(drift), a_gt, a_id


def test_accuracy_ceiling_from_ground_truth():
    """Started at the ground truth, the reference's loop drifts to its own fixed point, and that
    drift grows with the motion: at <=5 deg / 5 cm nearly every pair stays under 0.1 deg / 2 mm, at
    <=30 deg / 0.5 m few do -- with or without sensor noise and holes, and with normals from the
    FP64 companion. The identity start never does better. So at large motions no initial guess can
    raise the aligned fraction much: the accuracy ceiling of the learned workload."""
    rows = []
    med = {}
    for rot, trans in ((5, 0.05), (10, 0.1), (20, 0.3), (30, 0.5)):
        d, a_gt, a_id = _ceiling_bucket(rot, trans, 1.5e-3, 0.1)
        med[rot] = float(np.median(d))
        rows.append(("noisy", rot, a_gt, a_id, med[rot]))
        assert a_gt >= a_id - 1
    d0, a0, ai0 = _ceiling_bucket(30, 0.5, 0.0, 0.0)
    rows.append(("noise-free", 30, a0, ai0, float(np.median(d0))))
    d64, a64, _ = _ceiling_bucket(30, 0.5, 1.5e-3, 0.1, fp64_normals=True)
    rows.append(("FP64 normals", 30, a64, None, float(np.median(d64))))
    for kind, rot, a_gt, a_id, m in rows:
        print("%-12s <=%2d deg: aligned from truth %2d/12, from identity %s/12, median drift %.3f deg" % (
            kind, rot, a_gt, "-" if a_id is None else "%2d" % a_id, m))
    assert rows[0][2] >= 11 and med[5] < 0.05
    assert med[30] > 2 * med[5]
    assert rows[3][2] <= 4 and med[30] > 0.1
    assert a0 <= 6 and float(np.median(d0)) > 0.1
    assert abs(a64 - rows[3][2]) <= 1 and abs(float(np.median(d64)) - med[30]) < 0.1 * med[30]


# ---- the learned aligner's start choice -------------------------------------------------
def _keep_prediction(pred, fb):
    """The choice of nicp_batch_align_learned (include/nicp_amd.h): the prediction is kept unless
    the fallback has a lower chi2 per correspondence with at least half the prediction's
    correspondences, or the prediction has none; pred / fb = (chi2, inliers)."""
    if pred[1] <= 0:
        return fb[1] <= 0
    return not (fb[1] >= 0.5 * pred[1] and fb[0] / fb[1] < pred[0] / pred[1])


def test_start_choice_on_the_reference_linearizer():
    """One linearisation of the reference's aligner (the oracle, one outer iteration) at each
    candidate start judges them. A guess 15 deg off the truth loses to the identity at small
    motions (<= 5 deg / 5 cm); the ground truth wins over the identity at 20 deg although the identity
    collects more gated correspondences there (the robust kernel counts every gated pair as an
    inlier, linearizer.cpp:82-90), which is why the rule compares chi2 per correspondence."""
    rows, cols = 240, 320
    cam = O.camera(*synth.camera_for(rows, cols))
    sp, ap = O.stats_params(), O.align_params(outer_iterations=1)
    I = identity16()
    off = offset16(15.0, (1, 0.5, -0.2), (0.1, 0.0, 0.05)).reshape(4, 4, order="F").astype(np.float64)

    def probe(seed, rot, trans):
        dr, dc, T = synth.pair(seed, rows, cols, max_rot_deg=rot, max_trans_m=trans)
        fr, fc = O.Frame(dr, 0.001, cam, sp), O.Frame(dc, 0.001, cam, sp)
        bad = (T.reshape(4, 4, order="F").astype(np.float64) @ off).astype(np.float32).flatten(order="F")
        out = {}
        for name, start in (("truth", T), ("identity", I), ("bad", bad)):
            _, r = O.align(fr.cloud(), fc.cloud(), rows, cols, cam, start, ap, 1)
            out[name] = (float(r.error), int(r.inliers))
        return out

    for seed, rot, trans in ((15, 5.0, 0.05), (16, 3.0, 0.03)):
        p = probe(seed, rot, trans)
        print(seed, p)
        assert not _keep_prediction(p["bad"], p["identity"])      # the bad guess falls back
        assert _keep_prediction(p["truth"], p["identity"])        # the truth is kept
    p = probe(14, 20.0, 0.3)
    print(14, p)
    assert p["identity"][1] > p["truth"][1]                       # more inliers at the identity
    assert _keep_prediction(p["truth"], p["identity"])            # and still the truth is kept
