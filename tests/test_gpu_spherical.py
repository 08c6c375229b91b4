"""GPU parity of the SphericalPointProjector path (nicp/nicp/sphericalpointprojector.h:252-303,
.cpp:19-141, depthimageconverterintegralimage.cpp:48-52) against the oracle. Validity, index
images and every integer decision without trigonometry are exact; the trigonometric values
differ from the host's float functions by at most an ulp or two (the device evaluates them in
double and rounds, glibc's float functions are not correctly rounded), so points are compared
within 2e-6 relative and truncated pixel coordinates may move where a value sits within an ulp
of a pixel border."""
import numpy as np
import pytest

import oracle as O
from nicp_amd import capi, se3, synth
from tests.helpers import identity16, spherical_cams, spherical_room

pytestmark = pytest.mark.gpu

SP = dict(world_radius=0.25, min_image_radius=10, max_image_radius=100, min_points=30, curvature_threshold=0.075)


def _pair(rows, cols, seed, x6=(0.05, 0.0, -0.04, 0.0, 0.02, 0.0)):
    oc, dc = spherical_cams(rows, cols)
    Tgt = se3.v2t(np.array(x6)).astype(np.float32)
    ref = spherical_room(rows, cols, oc, seed=seed)
    cur = spherical_room(rows, cols, oc, pose=Tgt, seed=seed + 1)
    return oc, dc, ref, cur, Tgt


@pytest.mark.parametrize("rows,cols", [(180, 360), (1000, 1500)])
def test_conversion_matches_oracle(ctx, rows, cols):
    print("case spherical conversion", rows, cols, flush=True)
    oc, dc, raw, _, _ = _pair(rows, cols, 11)
    cl = ctx.cloud_from_depth(raw, dc, capi.stats_params(**SP))
    f = O.Frame(raw, 0.001, oc, O.stats_params(**SP))
    assert np.array_equal(cl.index_image(), f.index_image)
    p, nrm, curv, op, on = cl.download()
    assert np.all(np.abs(p[:, :3] - f.points[:, :3]) <= 2e-6 * (1 + np.abs(f.points[:, :3])))
    has_d = np.linalg.norm(nrm[:, :3], axis=1) > 0.5
    has_o = np.linalg.norm(f.normals[:, :3], axis=1) > 0.5
    assert (has_d != has_o).mean() < 1e-3
    both = has_d & has_o
    cosang = np.sum(nrm[both, :3] * f.normals[both, :3], axis=1)
    # an ulp in the points moves FP32 box sums over wide windows (cancellation): most normals
    # agree to 1e-7, a tail of a few per mille to a few degrees
    print("normal cosines: median %.9f q01 %.6f q001 %.6f" % (np.median(cosang), np.quantile(cosang, 0.01),
                                                                 np.quantile(cosang, 0.001)), flush=True)
    assert np.median(cosang) > 0.9999999 and np.quantile(cosang, 0.01) > 0.9999 and np.quantile(cosang, 0.001) > 0.99
    assert both.mean() > 0.5


def test_projection_matches_oracle(ctx):
    rows, cols = 360, 720
    print("case spherical projection", rows, cols, flush=True)
    oc, dc, raw, _, _ = _pair(rows, cols, 13)
    cl = ctx.cloud_from_depth(raw, dc, capi.stats_params(**SP))
    pts = cl.download()[0]
    T = se3.v2t(np.array([0.03, 0.01, -0.02, 0.01, 0.015, -0.005])).astype(np.float32)
    idx, dep = ctx.project(cl, T, dc, rows, cols)
    oidx = np.empty(rows * cols, np.int32)
    odep = np.empty(rows * cols, np.float32)
    O.lib().orc_spherical_project(O._p(pts), len(pts), O._p(T), O.C.byref(oc), rows, cols, O._p(oidx), O._p(odep))
    oidx = oidx.reshape(rows, cols)
    differ = idx != oidx
    assert differ.mean() < 1e-3, differ.mean()
    hit = (idx >= 0) & ~differ
    assert np.array_equal(dep[hit].view(np.uint32), odep.reshape(rows, cols)[hit].view(np.uint32))
    assert hit.mean() > 0.5


@pytest.mark.parametrize("rows,cols", [(180, 360), (500, 750)])
def test_align_matches_oracle(ctx, rows, cols):
    print("case spherical align", rows, cols, flush=True)
    oc, dc, ref, cur, Tgt = _pair(rows, cols, 15)
    dsp, osp = capi.stats_params(**SP), O.stats_params(**SP)
    cr, cc = ctx.cloud_from_depth(ref, dc, dsp), ctx.cloud_from_depth(cur, dc, dsp)
    ap, oap = capi.align_params(outer_iterations=20), O.align_params(outer_iterations=20)
    T, res = ctx.align(cr, cc, dc, rows, cols, identity16(), ap)
    To, _ = O.align(O.cloud_from_arrays(*cr.download()), O.cloud_from_arrays(*cc.download()), rows, cols, oc,
                    identity16(), oap, 1, fp64_solve=True)
    # the projections differ where a truncated azimuth sits within an ulp of a pixel border,
    # which changes a few correspondences per iteration
    rot, tr = synth.pose_error(T, To)
    print("device vs oracle: %.2e rad %.2e m" % (np.radians(rot), tr), flush=True)
    assert np.radians(rot) < 1e-3 and tr < 1e-3, (rot, tr)
    rot, tr = synth.pose_error(T, Tgt)
    assert rot < 0.2 and tr < 0.01, (rot, tr)
    assert res.inliers > 1000


def test_batch_matches_single(ctx):
    """Spherical pairs in a batch project their current clouds (no own-index shortcut) and give
    the single-pair results bit for bit."""
    rows, cols = 180, 360
    print("case spherical batch", rows, cols, flush=True)
    oc, dc = spherical_cams(rows, cols)
    dsp = capi.stats_params(**SP)
    poses = [se3.v2t(np.array(x)).astype(np.float32) for x in
             ([0.05, 0.0, -0.04, 0.0, 0.02, 0.0], [-0.03, 0.02, 0.05, 0.01, -0.015, 0.0])]
    frames = []
    for k, P in enumerate(poses):
        frames += [spherical_room(rows, cols, oc, seed=20 + 2 * k), spherical_room(rows, cols, oc, pose=P, seed=21 + 2 * k)]
    b = capi.Batch(ctx, [(rows, cols)] * 4, [dc] * 4, dsp)
    for i, d in enumerate(frames):
        b.upload(i, d)
    b.convert()
    ap = capi.align_params(outer_iterations=12)
    Tb, resb = b.align([0, 2], [1, 3], np.tile(identity16(), (2, 1)), ap, use_graph=True)
    for k in range(2):
        r = ctx.cloud_from_depth(frames[2 * k], dc, dsp)
        c = ctx.cloud_from_depth(frames[2 * k + 1], dc, dsp)
        T, res = ctx.align(r, c, dc, rows, cols, identity16(), ap)
        assert np.array_equal(T.view(np.uint32), Tb[k].view(np.uint32))
        assert res.inliers == resb[k].inliers
        rot, tr = synth.pose_error(T, poses[k])
        assert rot < 0.3 and tr < 0.01, (rot, tr)
    # one projector model per batch
    with pytest.raises(capi.NicpError):
        capi.Batch(ctx, [(rows, cols)] * 2, [dc, capi.camera(100, 100, 80, 60)], dsp)


def test_gaussians_and_merge(ctx):
    """The spherical unprojection gives identity covariances (sphericalpointprojector.cpp:112-113);
    a merge seen through the spherical projector matches the oracle's decisions."""
    rows, cols = 180, 360
    print("case spherical merge", rows, cols, flush=True)
    oc, dc, ref, cur, Tgt = _pair(rows, cols, 17)
    dsp = capi.stats_params(**SP)
    world = ctx.cloud_from_depth(ref, dc, dsp)
    world.compute_gaussians()
    g = world.download_gaussians()
    assert np.array_equal(O.gaussian_flags(g), np.ones(len(g), np.int32))
    assert np.array_equal(g[:, 4:13], np.tile(np.eye(3, dtype=np.float32).ravel(), (len(g), 1)))
    c1 = ctx.cloud_from_depth(cur, dc, dsp)
    c1.compute_gaussians()
    world.add(c1, Tgt)
    p, nrm, curv, op, on = world.download()
    go = world.download_gaussians()
    # seen from a frame's own pose, every point of that frame re-projects onto a pixel border
    # (the projector truncates), where an ulp of the azimuth picks the pixel: merge from a
    # nearby pose so that decisions do not hang on the last bit of atan2
    Tm = (Tgt.reshape(4, 4, order="F").astype(np.float64) @
          se3.v2t(np.array([0.011, -0.007, 0.013, 0.004, -0.003, 0.006])).reshape(4, 4, order="F"))
    Tm = Tm.astype(np.float32).flatten(order="F")
    ns, col = world.merge(Tm, dc, rows, cols, want_collapsed=True)
    k, col_o = O.merge(p, go, Tm, oc, rows, cols, normals=nrm, curvature=curv, omega_p=op, omega_n=on)
    print("decisions differing: %.2e" % (col != col_o).mean(), flush=True)
    assert (col != col_o).mean() < 2e-3
    assert abs(world.n - k) <= 2e-3 * len(col)
    assert ns > 0.3 * c1.n


def test_front_end_spherical_projector(ctx, tmp_path):
    """The C++ front end with nicp::SphericalPointProjector gives the C-ABI's result bit for bit."""
    import os
    import subprocess
    from tests.test_gpu_app import write_pgm16
    print("case spherical front end", flush=True)
    rows, cols = 180, 360
    oc, dc, ref, cur, Tgt = _pair(rows, cols, 19)
    write_pgm16(str(tmp_path / "a.pgm"), ref)
    write_pgm16(str(tmp_path / "b.pgm"), cur)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "sph")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(root, "include"), "-o", exe,
                           os.path.join(root, "tests", "cpp", "spherical_front.cpp"), "-L",
                           os.path.join(root, "nicp_amd"), "-lnicp_amd", "-Wl,-rpath," + os.path.join(root, "nicp_amd")])
    out = subprocess.run([exe, str(tmp_path / "a.pgm"), str(tmp_path / "b.pgm")], capture_output=True, text=True,
                         timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    lines = dict((l.split()[0], l.split()[1:]) for l in out.stdout.splitlines() if l.strip())
    Tf = np.array([float(v) for v in lines["T"]], np.float32)
    dsp = capi.stats_params(**SP)
    cr, cc = ctx.cloud_from_depth(ref, dc, dsp), ctx.cloud_from_depth(cur, dc, dsp)
    T, res = ctx.align(cr, cc, dc, rows, cols, identity16(), capi.align_params(outer_iterations=20))
    assert np.array_equal(Tf, T)
    assert int(lines["inliers"][0]) == res.inliers


def test_batch_conversion_equals_single_spherical(ctx):
    """16 panoramic frames in one batch (both prefix passes as one workgroup per frame, the
    points read through the index image) convert bit for bit as single frames do: points,
    normals and curvature. The batch fills the information matrices from the normals in closed
    form when a frame is handed out (b I + (a - b) n n^T): within 1e-4 of the largest entry of
    the single conversion's U diag(d) U^T."""
    rows, cols = 90, 180
    print("case spherical batch conversion", flush=True)
    oc, dc = spherical_cams(rows, cols)
    dsp = capi.stats_params(**SP)
    frames = [spherical_room(rows, cols, oc, seed=40 + k) for k in range(16)]
    b = capi.Batch(ctx, [(rows, cols)] * 16, [dc] * 16, dsp)
    for i, d in enumerate(frames):
        b.upload(i, d)
    b.convert()
    for k in range(16):
        s = ctx.cloud_from_depth(frames[k], dc, dsp)
        bc = b.frame_cloud(k, (rows, cols))
        assert bc.n == s.n
        xb, xs = bc.download(), s.download()
        for x, y in zip(xb[:3], xs[:3]):
            assert np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32)), k
        for x, y in zip(xb[3:], xs[3:]):
            x = np.asarray(x, np.float64).reshape(-1, 16)
            y = np.asarray(y, np.float64).reshape(-1, 16)
            scale = np.maximum(np.abs(y).max(axis=1), 1e-30)
            assert (np.abs(x - y).max(axis=1) <= 1e-4 * scale).all(), k
