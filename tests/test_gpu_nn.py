"""kNN correspondence variant (CorrespondenceFinderNN / AlignerNN) on the device
against the oracle's exact brute force: raw neighbour indices and gated pairs
bit-exact, transforms within 1e-5 rad / 1e-5 m, and the 200k-point
configuration against its ground truth."""
import sys

import numpy as np
import pytest

import oracle as O
from nicp_amd import capi, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _clouds(ctx, seed, n, **kw):
    d = synth.cloud_pair(seed, n=n, **kw)
    oref = O.cloud_from_arrays(d["src_points"], d["src_normals"], d["curvature"], d["omega_p"], d["omega_n"])
    ocur = O.cloud_from_arrays(d["tgt_points"], d["tgt_normals"], d["curvature"], d["omega_p"], d["omega_n"])
    dref = ctx.cloud_from_host(d["src_points"], d["src_normals"], d["curvature"], d["omega_p"], d["omega_n"])
    dcur = ctx.cloud_from_host(d["tgt_points"], d["tgt_normals"], d["curvature"], d["omega_p"], d["omega_n"])
    return d, oref, ocur, dref, dcur


@pytest.mark.parametrize("cell", [0.0, 0.011, 0.2])
def test_nn_correspondences_bitexact(ctx, cell):
    say("case nn_correspondences cell=%g" % cell)
    d, oref, ocur, dref, dcur = _clouds(ctx, 3, 6000)
    oap = O.align_params(inlier_distance_threshold=0.5)
    dap = capi.align_params(inlier_distance_threshold=0.5)
    # a transform between identity and the truth: many neighbours, some beyond the radius
    x = O.t2v(np.linalg.inv(d["T_gt"].reshape(4, 4, order="F").astype(np.float64)).astype(np.float32).flatten(order="F"))
    T = O.v2t((x * 0.5).astype(np.float32))
    op, onn = O.nn_correspondences(oref, ocur, T, oap)
    dp, dnn = ctx.nn_correspondences(dref, dcur, T, dap, cell_size=cell)
    say("  neighbours %d pairs oracle %d device %d" % (int((onn >= 0).sum()), len(op), len(dp)))
    assert len(op) > 1000
    np.testing.assert_array_equal(dnn, onn)
    np.testing.assert_array_equal(dp, op)


def test_align_nn_vs_oracle(ctx):
    say("case align_nn_vs_oracle")
    d, oref, ocur, dref, dcur = _clouds(ctx, 4, 5000)
    oap, dap = O.align_params(outer_iterations=20), capi.align_params(outer_iterations=20)
    I = np.eye(4, dtype=np.float32).flatten(order="F")
    To, ro = O.align_nn(oref, ocur, I, oap)
    T64, r64 = O.align_nn(oref, ocur, I, oap, fp64_solve=True)
    Td, rd = ctx.align_nn(dref, dcur, I, dap)
    r, t = synth.pose_error(Td, T64)
    r32, t32 = synth.pose_error(Td, To)
    gr, gt = synth.pose_error(Td, d["T_gt"])
    say("  device vs fp64-solve oracle %.2e rad %.2e m | vs FP32 reference %.2e rad %.2e m | vs truth %.4f deg %.5f m"
        " | inliers %d/%d" % (np.radians(r), t, np.radians(r32), t32, gr, gt, rd.inliers, ro.inliers))
    assert np.radians(r) < 1e-5 and t < 1e-5
    assert np.radians(r32) < 1e-5 and t32 < 1e-5
    assert rd.inliers == r64.inliers


def test_align_nn_200k(ctx):
    say("case align_nn_200k")
    ok = 0
    for seed in (10, 11):
        d, oref, ocur, dref, dcur = _clouds(ctx, seed, 200000)
        dap = capi.align_params(outer_iterations=20)
        Td, rd = ctx.align_nn(dref, dcur, np.eye(4, dtype=np.float32).flatten(order="F"), dap)
        gr, gt = synth.pose_error(Td, d["T_gt"])
        ir, it = synth.pose_error(np.eye(4).flatten(), d["T_gt"])
        say("  seed %d: init %.2f deg %.3f m -> %.4f deg %.5f m, inliers %d, %.1f ms" % (
            seed, ir, it, gr, gt, rd.inliers, rd.total_time_ms))
        ok += gr < 0.1 and gt < 0.002
    assert ok >= 1


def _empty(ctx):
    z = np.zeros((0, 4), np.float32)
    return ctx.cloud_from_host(z, z, np.zeros(0, np.float32), np.zeros((0, 16), np.float32),
                               np.zeros((0, 16), np.float32))


@pytest.mark.parametrize("cell", [0.0, 0.011])
def test_align_nn_batch_matches_single(ctx, cell):
    """nicp_align_nn_batch over ragged pairs (different sizes, an empty reference cloud, an
    empty current cloud, per-pair initial guesses) gives each pair's nicp_align_nn result bit
    for bit. The batch splits the grid arena between its pairs, so its grids are coarser than
    the single pair's: the neighbours must not depend on the cell size."""
    say("case align_nn_batch cell=%g" % cell)
    dap = capi.align_params(outer_iterations=8)
    refs, curs, inits, truth = [], [], [], []
    for k, (n, empty) in enumerate([(4000, None), (1500, None), (2500, "ref"), (800, "cur"), (6000, None)]):
        d, _, _, dref, dcur = _clouds(ctx, 30 + k, n)
        refs.append(_empty(ctx) if empty == "ref" else dref)
        curs.append(_empty(ctx) if empty == "cur" else dcur)
        x = O.t2v(d["T_gt"])
        inits.append(O.v2t((x * (0.25 * (k % 3))).astype(np.float32)))
        truth.append(d["T_gt"])
    Tb, rb = ctx.align_nn_batch(refs, curs, np.stack(inits), dap, cell_size=cell)
    for i in range(len(refs)):
        Ts, rs = ctx.align_nn(refs[i], curs[i], inits[i], dap, cell_size=cell)
        assert np.array_equal(Tb[i].view(np.uint32), Ts.view(np.uint32)), i
        # every field but the time, bit for bit
        rb[i].total_time_ms = rs.total_time_ms = 0.0
        assert bytes(rb[i]) == bytes(rs), i
    for i in (0, 1, 4):
        gr, gt = synth.pose_error(Tb[i], truth[i])
        say("  pair %d: %.4f deg %.5f m, inliers %d" % (i, gr, gt, rb[i].inliers))
        assert rb[i].inliers > 0
    # empty clouds: no correspondences, T stays at the guess (H = lambda I, b = 0)
    for i in (2, 3):
        rot, tr = synth.pose_error(Tb[i], inits[i])
        assert rb[i].inliers == 0 and rot < 1e-4 and tr < 1e-6
    # identities when no guesses are given
    T0, _ = ctx.align_nn_batch(refs[:2], curs[:2], None, dap, cell_size=cell)
    T1, _ = ctx.align_nn(refs[1], curs[1], np.eye(4, dtype=np.float32).flatten(order="F"), dap, cell_size=cell)
    assert np.array_equal(T0[1], T1)


def test_align_nn_batch_rejects_bad_arguments(ctx):
    d, _, _, dref, dcur = _clouds(ctx, 40, 500)
    ap = capi.align_params()
    with pytest.raises(capi.NicpError):
        ctx.align_nn_batch([], [], None, ap)
    other = capi.Context(0)
    try:
        foreign = other.cloud_from_host(d["src_points"], d["src_normals"], d["curvature"], d["omega_p"], d["omega_n"])
        with pytest.raises(capi.NicpError):
            ctx.align_nn_batch([dref, foreign], [dcur, dcur], None, ap)
    finally:
        other.close()
    with pytest.raises(capi.NicpError):
        ctx.align_nn_batch([dref], [dcur], None, capi.align_params(inner_iterations=2))


def _fnv(pairs):
    h = 1469598103934665603
    for r, c in pairs:
        h = ((h ^ int(r)) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
        h = ((h ^ int(c)) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def test_front_end_aligner_nn(ctx, tmp_path):
    """The C++ front end's CorrespondenceFinderNN and AlignerNN give the C-ABI's correspondences
    and transform bit for bit."""
    import os
    import subprocess
    say("case nn front end")
    d, _, _, dref, dcur = _clouds(ctx, 5, 4000)
    for name, p, n in (("ref", "src_points", "src_normals"), ("cur", "tgt_points", "tgt_normals")):
        with open(str(tmp_path / (name + ".bin")), "wb") as f:
            f.write(np.int32(len(d[p])).tobytes())
            for a in (d[p], d[n], d["curvature"], d["omega_p"], d["omega_n"]):
                f.write(np.ascontiguousarray(a, np.float32).tobytes())
    x = O.t2v(np.linalg.inv(d["T_gt"].reshape(4, 4, order="F").astype(np.float64)).astype(np.float32).flatten(order="F"))
    T = O.v2t((x * 0.5).astype(np.float32))
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "nn")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(root, "include"), "-o", exe,
                           os.path.join(root, "tests", "cpp", "nn_front.cpp"), "-L", os.path.join(root, "nicp_amd"),
                           "-lnicp_amd", "-Wl,-rpath," + os.path.join(root, "nicp_amd")])
    out = subprocess.run([exe, str(tmp_path / "ref.bin"), str(tmp_path / "cur.bin")] + ["%.9g" % v for v in T],
                         capture_output=True, text=True, timeout=120)
    say(out.stdout, out.stderr)
    assert out.returncode == 0
    lines = dict((l.split()[0], l.split()[1:]) for l in out.stdout.splitlines() if l.strip())
    dp, _ = ctx.nn_correspondences(dref, dcur, T, capi.align_params(inlier_distance_threshold=0.5))
    assert int(lines["pairs"][0]) == len(dp) and int(lines["pairs"][1]) == _fnv(dp)
    Td, rd = ctx.align_nn(dref, dcur, np.eye(4, dtype=np.float32).flatten(order="F"),
                          capi.align_params(outer_iterations=20))
    assert np.array_equal(np.array([float(v) for v in lines["T"]], np.float32), Td)
    assert int(lines["inliers"][0]) == rd.inliers
    # "inliers <n> correspondences <m>"
    assert lines["inliers"][1] == "correspondences" and int(lines["inliers"][2]) == len(ctx.last_correspondences())
