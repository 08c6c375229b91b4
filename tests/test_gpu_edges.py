"""Edge cases of the registration path on the device against the oracle: empty frames (no
valid depth), tiny and odd-sized images (1x1, 3x5, 37x53: a single partial tile, windows
clamped at every border), an empty reference or current cloud, zero iterations, and a ragged
batch mixing them with ordinary pairs (batch == single pair, bit for bit)."""
import numpy as np
import pytest

import oracle as O
from nicp_amd import capi, synth
from tests.helpers import aparams, cams, identity16, params

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (3, 5), (37, 53), (60, 80)]


def _depth(rows, cols, seed, empty=False):
    if empty:
        return np.zeros((rows, cols), np.uint16)
    if rows * cols >= 60 * 80:
        return synth.pair(seed, rows, cols)[0]
    rng = np.random.default_rng(seed)
    d = rng.uniform(800, 3000, (rows, cols))
    d[rng.uniform(size=d.shape) < 0.1] = 0
    return d.astype(np.uint16)


@pytest.mark.parametrize("rows,cols", SHAPES)
@pytest.mark.parametrize("empty", [False, True])
def test_conversion_tiny_and_empty(ctx, rows, cols, empty):
    print("case edges conversion", rows, cols, empty, flush=True)
    d = _depth(rows, cols, 3, empty)
    ocam, dcam = cams(rows, cols)
    osp, dsp = params(min_image_radius=1, min_points=3)
    fo = O.Frame(d, 0.001, ocam, osp)
    cl = ctx.cloud_from_depth(d, dcam, dsp, 0.001, keep_stats=True)
    assert cl.n == fo.n
    assert np.array_equal(cl.index_image(), fo.index_image)
    if empty:
        assert cl.n == 0
        return
    pts, nrm, curv, op, on = cl.download()
    assert np.array_equal(pts.view(np.uint32), fo.points.view(np.uint32))
    has_o = np.abs(fo.normals[:, :3]).sum(1) > 0
    has_d = np.abs(nrm[:, :3]).sum(1) > 0
    assert (has_o != has_d).sum() <= 1
    both = has_o & has_d
    if both.any():
        assert np.abs(nrm[both, :3] - fo.normals[both, :3]).max() < 1e-4
    ev, el, mn, npt = cl.download_stats()
    assert np.array_equal(npt, fo.npts)


@pytest.mark.parametrize("which", ["ref", "cur", "both"])
def test_align_with_empty_cloud(ctx, which):
    """No correspondences: the solve sees H = lambda I and b = 0, so T stays at the initial guess,
    as in the oracle; the solution is reported invalid."""
    print("case edges empty align", which, flush=True)
    rows, cols = 60, 80
    ocam, dcam = cams(rows, cols)
    osp, dsp = params()
    oap, dap = aparams()
    dr, dc, _ = synth.pair(5, rows, cols)
    z = np.zeros_like(dr)
    dr, dc = (z if which in ("ref", "both") else dr), (z if which in ("cur", "both") else dc)
    init = synth.motion(9, 3.0, 0.02)
    cr, cc = ctx.cloud_from_depth(dr, dcam, dsp), ctx.cloud_from_depth(dc, dcam, dsp)
    T, res = ctx.align(cr, cc, dcam, rows, cols, init, dap)
    fr, fc = O.Frame(dr, 0.001, ocam, osp), O.Frame(dc, 0.001, ocam, osp)
    To, ro = O.align(fr.cloud(), fc.cloud(), rows, cols, ocam, init, oap, 1, fp64_solve=True)
    assert np.array_equal(T, To)
    rot, tr = synth.pose_error(T, init)
    assert rot < 1e-4 and tr < 1e-6
    assert res.inliers == 0 and res.num_correspondences == 0 and res.solution_valid == 0


def test_zero_iterations(ctx):
    rows, cols = 60, 80
    _, dcam = cams(rows, cols)
    _, dsp = params()
    _, dap = aparams(outer_iterations=0)
    dr, dc, _ = synth.pair(6, rows, cols)
    init = synth.motion(11, 5.0, 0.05)
    T, res = ctx.align(ctx.cloud_from_depth(dr, dcam, dsp), ctx.cloud_from_depth(dc, dcam, dsp), dcam, rows, cols,
                       init, dap)
    assert np.array_equal(T, init)
    assert res.inliers == 0


def test_ragged_batch_with_empty_and_tiny_frames(ctx):
    print("case edges ragged batch", flush=True)
    _, dsp = params(min_image_radius=1, min_points=3)
    _, dap = aparams(outer_iterations=6)
    pairs = [((37, 53), False, False), ((1, 1), False, False), ((60, 80), False, True), ((3, 5), True, True),
             ((120, 160), False, False), ((60, 80), True, False)]
    frames, shapes, camlist = [], [], []
    for i, ((rows, cols), e0, e1) in enumerate(pairs):
        _, dcam = cams(rows, cols)
        if rows * cols >= 60 * 80:
            dr, dc, _ = synth.pair(40 + i, rows, cols)
            dr = np.zeros_like(dr) if e0 else dr
            dc = np.zeros_like(dc) if e1 else dc
        else:
            dr, dc = _depth(rows, cols, 40 + i, e0), _depth(rows, cols, 50 + i, e1)
        frames += [dr, dc]
        shapes += [(rows, cols)] * 2
        camlist += [dcam, dcam]
    b = capi.Batch(ctx, shapes, camlist, dsp)
    for k, f in enumerate(frames):
        b.upload(k, f)
    b.convert()
    ref = np.arange(0, len(frames), 2, dtype=np.int32)
    cur = ref + 1
    Tb, rb = b.align(ref, cur, None, dap, use_graph=False)
    Tg, _ = b.align(ref, cur, None, dap, use_graph=True)
    assert np.array_equal(Tg.view(np.uint32), Tb.view(np.uint32))
    for i, ((rows, cols), _, _) in enumerate(pairs):
        _, dcam = cams(rows, cols)
        cr = ctx.cloud_from_depth(frames[2 * i], dcam, dsp)
        cc = ctx.cloud_from_depth(frames[2 * i + 1], dcam, dsp)
        Ts, rs = ctx.align(cr, cc, dcam, rows, cols, identity16(), dap)
        assert np.array_equal(Tb[i].view(np.uint32), Ts.view(np.uint32)), i
        assert rb[i].inliers == rs.inliers


def test_batch_refuses_work_before_conversion(ctx):
    """A new batch's index images, points and point counts are uninitialised until the first
    nicp_batch_convert: aligning its frames or taking a frame's cloud before it is an error."""
    print("case edges batch before conversion", flush=True)
    _, dsp = params()
    _, dap = aparams()
    _, dcam = cams(240, 320)
    b = capi.Batch(ctx, [(240, 320)] * 2, [dcam, dcam], dsp)
    dr, dc, _ = synth.pair(3, 240, 320)
    b.upload(0, dr)
    b.upload(1, dc)
    with pytest.raises(capi.NicpError):
        b.align([0], [1], None, dap, use_graph=False)
    with pytest.raises(capi.NicpError):
        b.frame_cloud(0, (240, 320))
    b.convert()
    T, r = b.align([0], [1], None, dap, use_graph=False)
    assert np.all(np.isfinite(T)) and r[0].inliers > 0
    b.close()
