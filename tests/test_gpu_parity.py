"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bars: index images, points, z-buffers and correspondence sets bit-exact on
identical float inputs; normals within 1e-5 (the eigensolver's atan2/cos/sin
come from different math libraries); H and b within 1e-5 relative of the FP64
twin of the linearizer on identical inputs; transforms within 1e-5 rad and
1e-5 m of the oracle's AlignerProjective on identical clouds.
"""
import sys

import numpy as np
import pytest

import oracle as O
from nicp_amd import capi, synth
from tests.helpers import aparams, cams, identity16, offset16, offset_frame_arrays, params

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _frames(seed, rows, cols, **kw):
    dr, dc, T = synth.pair(seed, rows, cols, **kw)
    return dr, dc, T


@pytest.mark.parametrize("seed,res", [(0, (240, 320)), (1, (480, 640)), (2, (120, 160)), (3, (960, 1280))])
def test_frame_conversion(ctx, seed, res):
    rows, cols = res
    say("case frame_conversion seed=%d %dx%d" % (seed, rows, cols))
    dr, _, _ = _frames(seed, rows, cols)
    ocam, dcam = cams(rows, cols)
    osp, dsp = params()
    fo = O.Frame(dr, 0.001, ocam, osp)
    cl = ctx.cloud_from_depth(dr, dcam, dsp, 0.001, keep_stats=True)
    assert cl.n == fo.n
    np.testing.assert_array_equal(cl.index_image(), fo.index_image)
    pts, nrm, curv, op, on = cl.download()
    assert pts.view(np.uint32).tolist() == fo.points.view(np.uint32).tolist() or np.array_equal(
        pts.view(np.uint32), fo.points.view(np.uint32))
    # which points carry a normal
    has_o = np.abs(fo.normals[:, :3]).sum(1) > 0
    has_d = np.abs(nrm[:, :3]).sum(1) > 0
    flips = int((has_o != has_d).sum())
    both = has_o & has_d
    err = np.abs(nrm[both, :3] - fo.normals[both, :3]).max(axis=1)
    bad = int((err > 1e-5).sum())
    say("  points %d normals %d flips %d |dn|max %.3g >1e-5: %d" % (fo.n, int(both.sum()), flips, err.max(), bad))
    # gate flips (curvature at the stats threshold) and tolerance misses must be rare
    assert flips <= max(2, fo.n // 20000)
    assert bad <= max(2, int(both.sum()) // 10000)
    cerr = np.abs(curv - fo.curvature)
    assert np.quantile(cerr[both], 0.999) < 1e-5
    ev, el, mn, npt = cl.download_stats()
    np.testing.assert_array_equal(npt, fo.npts)
    np.testing.assert_array_equal(mn[:, :3].view(np.uint32), fo.means[:, :3].view(np.uint32))
    # information matrices: upper triangle within a relative 1e-4 of the largest entry
    up = [0, 4, 8, 5, 9, 10]
    for dev, ref in ((op, fo.omega_p), (on, fo.omega_n)):
        d = np.abs(dev[both][:, up] - ref[both][:, up]).max(axis=1)
        scale = np.abs(ref[both][:, up]).max(axis=1) + 1e-30
        assert np.quantile(d / scale, 0.999) < 1e-4


PROJ_CASES = [(5, (240, 320), 5.0, 0.05), (11, (480, 640), 30.0, 0.5), (12, (480, 640), 15.0, 0.3),
              (13, (960, 1280), 30.0, 0.5), (14, (960, 1280), 8.0, 0.1)]


@pytest.mark.parametrize("seed,res,rot,trans", PROJ_CASES)
def test_project_bitexact(ctx, seed, res, rot, trans):
    rows, cols = res
    say("case project_bitexact seed=%d %dx%d rot<=%g" % (seed, rows, cols, rot))
    dr, _, Tgt = _frames(seed, rows, cols, max_rot_deg=rot, max_trans_m=trans)
    ocam, dcam = cams(rows, cols)
    osp, _ = params()
    fo = O.Frame(dr, 0.001, ocam, osp)
    cl = ctx.cloud_from_host(fo.points, fo.normals, fo.curvature, fo.omega_p, fo.omega_n)
    for T in (identity16(), Tgt):
        KRt, _ = O.projector_matrices(np.array(ocam.K, np.float32), T)
        oi, od = O.project(fo.points, KRt, rows, cols, 0.5, 4.5)
        di, dd = ctx.project(cl, T, dcam, rows, cols)
        np.testing.assert_array_equal(di, oi)
        np.testing.assert_array_equal(dd.view(np.uint32), od.view(np.uint32))


def _oracle_clouds(seed, rows, cols, **kw):
    dr, dc, Tgt = _frames(seed, rows, cols, **kw)
    ocam, dcam = cams(rows, cols)
    osp, _ = params()
    fr = O.Frame(dr, 0.001, ocam, osp)
    fc = O.Frame(dc, 0.001, ocam, osp)
    return fr, fc, Tgt, ocam, dcam


def _upload(ctx, f):
    return ctx.cloud_from_host(f.points, f.normals, f.curvature, f.omega_p, f.omega_n)


@pytest.mark.parametrize("multi", [False, True])
@pytest.mark.parametrize("seed,res,rot,trans", [(6, (240, 320), 5.0, 0.05)] + PROJ_CASES[1:])
def test_correspondences_bitexact(ctx, multi, seed, res, rot, trans):
    rows, cols = res
    say("case correspondences multi=%s seed=%d %dx%d rot<=%g" % (multi, seed, rows, cols, rot))
    fr, fc, Tgt, ocam, dcam = _oracle_clouds(seed, rows, cols, max_rot_deg=rot, max_trans_m=trans)
    kw = dict(row_search_region=1, col_search_region=2) if multi else {}
    oap, dap = aparams(**kw)
    # reference projected with T, current with the identity, finder transform T^-1
    T = Tgt
    KRt, _ = O.projector_matrices(np.array(ocam.K, np.float32), T)
    ri, _ = O.project(fr.points, KRt, rows, cols, 0.5, 4.5)
    KRt0, _ = O.projector_matrices(np.array(ocam.K, np.float32), identity16())
    ci, _ = O.project(fc.points, KRt0, rows, cols, 0.5, 4.5)
    Tinv = np.linalg.inv(T.reshape(4, 4, order="F").astype(np.float64)).astype(np.float32).flatten(order="F")
    op = O.correspondences(fr.cloud(), fc.cloud(), ri, ci, Tinv, oap, 1)
    dp = ctx.correspondences(_upload(ctx, fr), _upload(ctx, fc), ri, ci, Tinv, dap)
    say("  pairs oracle %d device %d" % (len(op), len(dp)))
    assert len(op) > 500
    np.testing.assert_array_equal(dp, op)


def test_linearize_vs_fp64(ctx):
    rows, cols = 480, 640
    say("case linearize_vs_fp64")
    fr, fc, Tgt, ocam, dcam = _oracle_clouds(7, rows, cols)
    oap, dap = aparams()
    KRt, _ = O.projector_matrices(np.array(ocam.K, np.float32), Tgt)
    ri, _ = O.project(fr.points, KRt, rows, cols, 0.5, 4.5)
    KRt0, _ = O.projector_matrices(np.array(ocam.K, np.float32), identity16())
    ci, _ = O.project(fc.points, KRt0, rows, cols, 0.5, 4.5)
    Tinv = np.linalg.inv(Tgt.reshape(4, 4, order="F").astype(np.float64)).astype(np.float32).flatten(order="F")
    pairs = O.correspondences(fr.cloud(), fc.cloud(), ri, ci, Tinv, oap, 1)
    H64, b64, e64, i64 = O.linearize_f64(fr.cloud(), fc.cloud(), pairs, Tinv, oap)
    H32, b32, e32, i32 = O.linearize(fr.cloud(), fc.cloud(), pairs, Tinv, oap, 1)
    Hd, bd, ed, idd = ctx.linearize(_upload(ctx, fr), _upload(ctx, fc), pairs, Tinv, dap)
    relH = np.abs(Hd - H64).max() / np.abs(H64).max()
    relb = np.abs(bd - b64).max() / np.abs(b64).max()
    relH32 = np.abs(H32 - H64).max() / np.abs(H64).max()
    relb32 = np.abs(b32 - b64).max() / np.abs(b64).max()
    say("  pairs %d  device vs fp64: H %.2e b %.2e | fp32 oracle vs fp64: H %.2e b %.2e" % (
        len(pairs), relH, relb, relH32, relb32))
    assert idd == i64
    # H sums mostly positive terms; b sums signed residual terms whose FP32 evaluation
    # (inputs and transforms are FP32, as in the reference) bounds its agreement with FP64
    assert relH < 1e-5 and relb < 1e-4
    assert abs(ed - e64) / max(e64, 1e-30) < 1e-5


def test_linearize_closed_omega_vs_fp64(ctx):
    """Clouds converted on the device carry their information matrices in closed form (both
    calculators' diagonals have equal second and third entries: b I + (a - b) n n^T), and the
    linearisers compute them from the normals instead of loading them. Pinned two ways: the
    closed form against the stored matrices (the reference's U diag(d) U^T, FP32) of every point
    within 1e-4 of the largest entry, and the device's H and b on such clouds against the FP64
    twin run on the stored matrices."""
    rows, cols = 480, 640
    say("case linearize_closed_omega")
    dr, dc, Tgt = _frames(7, rows, cols)
    ocam, dcam = cams(rows, cols)
    oap, dap = aparams()
    _, dsp = params()
    cr = ctx.cloud_from_depth(dr, dcam, dsp)
    cc = ctx.cloud_from_depth(dc, dcam, dsp)
    pr_, nr_, kr_, opr, onr = cr.download()
    pc_, nc_, kc_, opc, onc = cc.download()
    # the closed form of every current point's matrices against the stored ones
    n = np.asarray(nc_, np.float64)[:, :3]
    curv = np.asarray(kc_, np.float64)
    sp = dsp
    pa = np.where(curv < sp.point_curvature_threshold, sp.point_flat[0], sp.point_nonflat[0])
    pb = np.where(curv < sp.point_curvature_threshold, sp.point_flat[1], sp.point_nonflat[1])
    na = np.where(curv < sp.normal_curvature_threshold, sp.normal_flat[0], sp.normal_nonflat[0])
    nb = np.where(curv < sp.normal_curvature_threshold, sp.normal_flat[1], sp.normal_nonflat[1])
    iu = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    has = np.any(n != 0.0, axis=1)
    for a, b, stored in ((pa, pb, opc), (na, nb, onc)):
        closed = np.stack([(a - b) * n[:, i] * n[:, j] + (b if i == j else 0.0) for i, j in iu], axis=1)
        closed[~has] = 0.0
        m = np.asarray(stored, np.float64).reshape(-1, 4, 4)  # 4x4 per point, the 3x3 block used
        st = np.stack([m[:, i, j] for i, j in iu], axis=1)
        err = np.abs(closed - st).max(axis=1) / np.maximum(np.abs(st).max(axis=1), 1e-30)
        say("  closed form vs stored: max rel %.2e (99.9%% %.2e) over %d points" % (
            err[has].max(), np.quantile(err[has], 0.999), int(has.sum())))
        # the stored U diag(d) U^T carries the FP32 eigenvectors' departure from orthonormality
        # (up to ~3e-5 at near-degenerate windows), which b I + (a - b) n n^T does not; the bar is
        # that of the information matrices against the oracle (test_frame_conversion)
        assert err[has].max() < 1e-4
        assert np.all(st[~has] == 0.0)
    # linearisation on the device's clouds against the FP64 twin on the stored matrices
    fr = O.cloud_from_arrays(pr_, _no_w(nr_), kr_, opr, onr)
    fc = O.cloud_from_arrays(pc_, _no_w(nc_), kc_, opc, onc)
    KRt, _ = O.projector_matrices(np.array(ocam.K, np.float32), Tgt)
    ri, _ = O.project(np.asarray(pr_, np.float32), KRt, rows, cols, 0.5, 4.5)
    KRt0, _ = O.projector_matrices(np.array(ocam.K, np.float32), identity16())
    ci, _ = O.project(np.asarray(pc_, np.float32), KRt0, rows, cols, 0.5, 4.5)
    Tinv = np.linalg.inv(Tgt.reshape(4, 4, order="F").astype(np.float64)).astype(np.float32).flatten(order="F")
    pairs = O.correspondences(fr, fc, ri, ci, Tinv, oap, 1)
    H64, b64, e64, i64 = O.linearize_f64(fr, fc, pairs, Tinv, oap)
    Hd, bd, ed, idd = ctx.linearize(cr, cc, pairs, Tinv, dap)
    relH = np.abs(Hd - H64).max() / np.abs(H64).max()
    relb = np.abs(bd - b64).max() / np.abs(b64).max()
    say("  pairs %d  device (closed form) vs fp64 (stored): H %.2e b %.2e" % (len(pairs), relH, relb))
    assert len(pairs) > 500
    assert idd == i64
    assert relH < 1e-5 and relb < 1e-4
    assert abs(ed - e64) / max(e64, 1e-30) < 1e-5
    cr.close()
    cc.close()


def _no_w(nrm):
    n = np.array(nrm, np.float32, copy=True)
    n[:, 3] = 0.0
    return n


@pytest.mark.parametrize("seed", [1, 4])
def test_align_vs_oracle(ctx, seed):
    """AlignerProjective::align on identical clouds. The device computes H and b in
    FP32 per thread and FP64 beyond, and solves in FP64: it must match the oracle loop run with the same
    design within 1e-5 rad / 1e-5 m, and the reference's FP32 loop within the larger of 1e-5 and
    twice the spread of that FP32 loop itself across OpenMP thread counts (1 vs 8 threads
    group its partial sums differently)."""
    rows, cols = 480, 640
    say("case align_vs_oracle seed=%d" % seed)
    fr, fc, Tgt, ocam, dcam = _oracle_clouds(seed, rows, cols)
    oap, dap = aparams()
    To, ro = O.align(fr.cloud(), fc.cloud(), rows, cols, ocam, identity16(), oap, 1)
    To8, _ = O.align(fr.cloud(), fc.cloud(), rows, cols, ocam, identity16(), oap, 8)
    T64, r64 = O.align(fr.cloud(), fc.cloud(), rows, cols, ocam, identity16(), oap, 1, fp64_solve=True)
    Td, rd = ctx.align(_upload(ctx, fr), _upload(ctx, fc), dcam, rows, cols, identity16(), dap)
    r64d, t64d = synth.pose_error(Td, T64)
    r32d, t32d = synth.pose_error(Td, To)
    rs, ts = synth.pose_error(To8, To)
    gr, gt = synth.pose_error(Td, Tgt)
    say("  device vs fp64-solve oracle: %.2e rad %.2e m | vs FP32 reference %.2e rad %.2e m "
        "(reference 1 vs 8 threads: %.2e rad %.2e m) | vs truth %.4f deg %.5f m | inliers %d/%d/%d valid %d/%d" % (
            np.radians(r64d), t64d, np.radians(r32d), t32d, np.radians(rs), ts, gr, gt, rd.inliers, r64.inliers,
            ro.inliers, rd.solution_valid, ro.solution_valid))
    assert np.radians(r64d) < 1e-5 and t64d < 1e-5
    assert np.radians(r32d) < max(1e-5, 2 * np.radians(rs)) and t32d < max(1e-5, 2 * ts)
    assert abs(rd.inliers - r64.inliers) <= max(5, r64.inliers // 5000)
    assert rd.solution_valid == ro.solution_valid
    # solution statistics (computed on the device in every align call) against the FP64-solve
    # oracle's: H agrees to ~1e-5 and T to ~1e-7, which bounds how far Omega and the ratios move
    Od = np.array(rd.omega).reshape(6, 6)
    Oo = np.array(r64.omega).reshape(6, 6)
    relO = np.abs(Od - Oo).max() / np.abs(Oo).max()
    say("  statistics: |dOmega|/|Omega| %.2e ratios %.4g/%.4g vs %.4g/%.4g" % (
        relO, rd.translational_ratio, rd.rotational_ratio, r64.translational_ratio, r64.rotational_ratio))
    assert relO < 1e-3
    assert abs(rd.translational_ratio - r64.translational_ratio) <= 1e-3 * r64.translational_ratio
    assert abs(rd.rotational_ratio - r64.rotational_ratio) <= 1e-3 * r64.rotational_ratio
    assert np.abs(np.array(rd.mean) - np.array(r64.mean)).max() < 1e-5
    # the device's last correspondences agree with the oracle's up to gate-boundary flips
    dp = ctx.last_correspondences()
    assert abs(len(dp) - r64.num_correspondences) <= max(5, r64.num_correspondences // 5000)


def test_align_sensor_offsets_vs_oracle(ctx):
    """AlignerProjective::align with reference and current sensor offsets (alignerprojective.cpp:
    36, 51) on clouds moved by them, against the FP64-solve oracle. Equal offsets: within 1e-5 rad
    / 1e-5 m. Offsets one degree and 2 cm apart make the projection and the error frames disagree:
    residuals grow, ~0.02% of the inliers sit on a gate and rounding flips them (the oracle's own
    FP32 and FP64 solves differ by 6e-6..4e-5 rad and 1e-5 m there), so the bar is 1e-4 rad / 1e-4 m,
    still 300 times below what swapping the two offsets changes (2 deg, 4 cm). Identity offsets give
    nicp_align bit for bit."""
    rows, cols = 480, 640
    say("case align_sensor_offsets 640x480")
    fr, fc, _, ocam, dcam = _oracle_clouds(7, rows, cols, max_rot_deg=6.0, max_trans_m=0.06)
    oap, dap = aparams()
    Tp, rp = ctx.align(_upload(ctx, fr), _upload(ctx, fc), dcam, rows, cols, identity16(), dap)
    Ti, ri = ctx.align_offsets(_upload(ctx, fr), _upload(ctx, fc), dcam, rows, cols, identity16(), identity16(),
                               identity16(), dap)
    np.testing.assert_array_equal(Ti.view(np.uint32), Tp.view(np.uint32))
    assert ri.inliers == rp.inliers
    S = offset16(25.0, (0.3, 1, 0.2), (0.15, -0.4, 0.9))
    S2 = offset16(26.0, (0.3, 1, 0.2), (0.17, -0.4, 0.9))
    got = {}
    for ro, co, tag in ((S, S, "equal"), (S, S2, "S,S2"), (S2, S, "S2,S")):
        ar, ac = offset_frame_arrays(fr, ro), offset_frame_arrays(fc, co)
        oref, ocur = O.cloud_from_arrays(*ar), O.cloud_from_arrays(*ac)
        To, r_o = O.align_offsets(oref, ocur, rows, cols, ocam, identity16(), ro, co, oap, 1, fp64_solve=True)
        To32, _ = O.align_offsets(oref, ocur, rows, cols, ocam, identity16(), ro, co, oap, 1)
        Td, rd = ctx.align_offsets(ctx.cloud_from_host(*ar), ctx.cloud_from_host(*ac), dcam, rows, cols,
                                   identity16(), ro, co, dap)
        rot, tr = synth.pose_error(Td, To)
        rs, ts = synth.pose_error(To32, To)
        bar_r = bar_t = 1e-5 if tag == "equal" else 1e-4
        say("  offsets %s: device vs oracle %.2e rad %.2e m (bar %.0e), oracle FP32 vs FP64 %.2e rad %.2e m, "
            "inliers %d/%d" % (tag, np.radians(rot), tr, bar_r, np.radians(rs), ts, rd.inliers, r_o.inliers))
        assert np.radians(rot) < bar_r and tr < bar_t
        assert abs(rd.inliers - r_o.inliers) <= max(5, r_o.inliers // 2000)
        got[tag] = (Td, bar_r)
    swap_r, _ = synth.pose_error(got["S,S2"][0], got["S2,S"][0])
    assert np.radians(swap_r) > 100 * max(got["S,S2"][1], got["S2,S"][1])


def test_solution_statistics_vs_oracle(ctx):
    """Aligner::_computeStatistics on the device (nicp_solution_statistics, the kernel every align
    call runs) against the oracle's FP64 restatement on the same H and T: both are FP64 with
    different factorisations (Jacobi eigen-decompositions here, SVD/LLT/inverse there)."""
    say("case solution_statistics_vs_oracle")
    rng = np.random.default_rng(3)
    oap, dap = aparams()
    cases = []
    for k in range(12):
        U, _ = np.linalg.qr(rng.normal(size=(6, 6)))
        H = U @ np.diag(10.0 ** rng.uniform(4, 4.0 + k / 4.0, 6)) @ U.T
        q = rng.normal(size=3)
        q *= np.sin(np.radians(rng.uniform(0, 30)) / 2) / np.linalg.norm(q)
        T = O.v2t(np.concatenate([rng.uniform(-0.5, 0.5, 3), q]).astype(np.float32))
        cases.append((H, T))
    # degenerate: zero information, and a single plane (large eigen ratio)
    cases.append((np.zeros((6, 6)), identity16()))
    Hp = np.diag([1e-3, 1e-3, 1e6, 1e5, 1e5, 1e-3])
    cases.append((Hp, identity16()))
    n_valid = n_cmp = 0
    for H, T in cases:
        ro = O.compute_statistics(H, T, oap)
        rd = ctx.solution_statistics(H, T, dap)
        assert rd.solution_valid == ro.solution_valid
        n_valid += ro.solution_valid
        if not np.isfinite(ro.translational_ratio):
            continue
        n_cmp += 1
        Oo = np.array(ro.omega).reshape(6, 6)
        Od = np.array(rd.omega).reshape(6, 6)
        assert np.abs(Od - Oo).max() <= 1e-8 * np.abs(Oo).max()
        assert np.abs(np.array(rd.mean) - np.array(ro.mean)).max() < 1e-12
        assert abs(rd.translational_ratio - ro.translational_ratio) <= 1e-8 * ro.translational_ratio
        assert abs(rd.rotational_ratio - ro.rotational_ratio) <= 1e-8 * ro.rotational_ratio
    say("  %d cases, %d valid, %d compared" % (len(cases), n_valid, n_cmp))
    assert n_valid >= 3 and n_cmp >= 12


def test_pipeline_accuracy(ctx):
    """Depth -> clouds -> align entirely on the device, against the sampled ground truth."""
    rows, cols = 480, 640
    ocam, dcam = cams(rows, cols)
    _, dsp = params()
    _, dap = aparams()
    ok = 0
    for seed in range(2, 6):
        say("case pipeline seed=%d" % seed)
        dr, dc, Tgt = _frames(seed, rows, cols)
        cr = ctx.cloud_from_depth(dr, dcam, dsp)
        cc = ctx.cloud_from_depth(dc, dcam, dsp)
        T, res = ctx.align(cr, cc, dcam, rows, cols, identity16(), dap)
        rot, tr = synth.pose_error(T, Tgt)
        say("  err %.4f deg %.5f m valid %d time %.2f ms" % (rot, tr, res.solution_valid, res.total_time_ms))
        ok += rot < 0.1 and tr < 0.002
    assert ok >= 3


@pytest.mark.parametrize("iters", [10, 15])
def test_batch_matches_single(ctx, iters):
    """Uneven batches (the four resolutions of the batched configuration) equal single-pair runs
    bit for bit, eagerly and replayed from the captured graph."""
    shapes = [(120, 160), (240, 320), (480, 640), (960, 1280), (240, 320)]
    say("case batch_matches_single iters=%d" % iters)
    _, dsp = params()
    _, dap = aparams(outer_iterations=iters)
    frames, cam_list, singles = [], [], []
    for i, (rows, cols) in enumerate(shapes):
        dr, dc, Tgt = _frames(20 + i, rows, cols, max_rot_deg=8.0, max_trans_m=0.1)
        _, dcam = cams(rows, cols)
        frames += [dr, dc]
        cam_list += [dcam, dcam]
    b = capi.Batch(ctx, [f.shape for f in frames], cam_list, dsp)
    for k, f in enumerate(frames):
        b.upload(k, f)
    b.convert()
    ref = np.arange(0, len(frames), 2)
    cur = ref + 1
    Tb, rb = b.align(ref, cur, None, dap, use_graph=False)
    Tg, rg = b.align(ref, cur, None, dap, use_graph=True)
    Tg2, _ = b.align(ref, cur, None, dap, use_graph=True)
    for i, (rows, cols) in enumerate(shapes):
        _, dcam = cams(rows, cols)
        cr = ctx.cloud_from_depth(frames[2 * i], dcam, dsp)
        cc = ctx.cloud_from_depth(frames[2 * i + 1], dcam, dsp)
        Ts, rs = ctx.align(cr, cc, dcam, rows, cols, identity16(), dap)
        np.testing.assert_array_equal(Tb[i].view(np.uint32), Ts.view(np.uint32))
        assert rb[i].inliers == rs.inliers
        bc = b.frame_cloud(2 * i, (rows, cols))
        assert bc.n == cr.n
    np.testing.assert_array_equal(Tg.view(np.uint32), Tb.view(np.uint32))
    np.testing.assert_array_equal(Tg2.view(np.uint32), Tb.view(np.uint32))


def test_depth_pyramid_and_float_input(ctx):
    """DepthImage_scale bit-exact; the float-depth conversion equals the uint16 one."""
    say("case depth_pyramid_and_float_input")
    rows, cols = 480, 640
    dr, _, _ = _frames(9, rows, cols)
    dm = np.where(dr > 0, np.float32(0.001) * dr.astype(np.float32), 0).astype(np.float32)
    for step in (1, 2, 3):
        o = O.depth_scale(dm, step)
        d16 = ctx.depth_scale(dr, step)
        d32 = ctx.depth_scale(dm, step)
        np.testing.assert_array_equal(d16.view(np.uint32), o.view(np.uint32))
        np.testing.assert_array_equal(d32.view(np.uint32), o.view(np.uint32))
    _, dcam = cams(rows, cols)
    _, dsp = params()
    a = ctx.cloud_from_depth(dr, dcam, dsp)
    b = ctx.cloud_from_depth_f32(dm, dcam, dsp)
    assert a.n == b.n
    for x, y in zip(a.download(), b.download()):
        np.testing.assert_array_equal(x.view(np.uint32), y.view(np.uint32))


def test_batch_conversion_equals_single(ctx):
    """A batch of many frames (>= 16: both prefix passes then run as one workgroup per frame)
    converts every frame to the clouds of single-frame conversion (the per-row and per-column
    prefix kernels), bit for bit: index images, points, normals and curvature. A batch does not
    store the information matrices (its linearisers form them from the normals in closed form,
    b I + (a - b) n n^T) and fills a frame's arrays that way when the frame is handed out; they
    are within 1e-4 of the largest entry of the single conversion's U diag(d) U^T, the bar of
    the information matrices against the oracle. Odd sizes exercise partial bands and chunks."""
    say("case batch_conversion_equals_single")
    shapes = [(120, 160), (480, 640), (37, 53), (1, 1), (240, 320), (41, 49), (81, 96), (960, 1280), (130, 20)] * 2
    _, dsp = params(min_image_radius=1, min_points=3)
    frames, cam_list = [], []
    for i, (rows, cols) in enumerate(shapes):
        if rows * cols >= 60 * 80:
            d = _frames(60 + i, rows, cols)[0]
        else:
            rng = np.random.default_rng(60 + i)
            d = rng.uniform(800, 3000, (rows, cols))
            d[rng.uniform(size=d.shape) < 0.1] = 0
            d = d.astype(np.uint16)
        frames.append(d)
        cam_list.append(cams(rows, cols)[1])
    b = capi.Batch(ctx, shapes, cam_list, dsp)
    for k, f in enumerate(frames):
        b.upload(k, f)
    b.convert()
    for k, (rows, cols) in enumerate(shapes):
        s = ctx.cloud_from_depth(frames[k], cam_list[k], dsp)
        bc = b.frame_cloud(k, (rows, cols))
        assert bc.n == s.n, k
        np.testing.assert_array_equal(bc.index_image(), s.index_image())
        xb, xs = bc.download(), s.download()
        for x, y in zip(xb[:3], xs[:3]):
            np.testing.assert_array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32))
        for x, y in zip(xb[3:], xs[3:]):
            x = np.asarray(x, np.float64).reshape(-1, 16)
            y = np.asarray(y, np.float64).reshape(-1, 16)
            scale = np.maximum(np.abs(y).max(axis=1), 1e-30)
            assert (np.abs(x - y).max(axis=1) <= 1e-4 * scale).all(), k
            np.testing.assert_array_equal(x[scale <= 1e-30], 0.0)
