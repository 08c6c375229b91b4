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
