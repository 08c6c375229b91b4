"""nicp_align_noalloc (the fixed-workspace variant of nicp_align) against nicp_align through the
C ABI: bit-identical transforms and results on consecutive calls of decreasing size (the
workspace keeps the previous call's contents), and a clear error beyond its limits."""
import sys

import numpy as np
import pytest

from nicp_amd import capi, synth
from tests.helpers import aparams, cams, identity16, params

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _clouds(ctx, seed, rows, cols):
    dr, dc, Tgt = synth.pair(seed, rows, cols)
    _, dcam = cams(rows, cols)
    _, dsp = params()
    return ctx.cloud_from_depth(dr, dcam, dsp), ctx.cloud_from_depth(dc, dcam, dsp), dcam


def _same(a, b):
    Ta, ra = a
    Tb, rb = b
    np.testing.assert_array_equal(Ta.view(np.uint32), Tb.view(np.uint32))
    assert ra.inliers == rb.inliers and ra.num_correspondences == rb.num_correspondences
    assert np.float32(ra.error).view(np.uint32) == np.float32(rb.error).view(np.uint32)
    for f in ("H", "b", "omega", "mean"):
        np.testing.assert_array_equal(np.array(getattr(ra, f)).view(np.uint64), np.array(getattr(rb, f)).view(np.uint64))
    assert ra.solution_valid == rb.solution_valid


def test_noalloc_matches_align(ctx):
    _, dap = aparams()
    cases = [(31, 480, 640), (32, 240, 320), (33, 480, 640)]
    for seed, rows, cols in cases:
        say("case align_noalloc seed=%d %dx%d" % (seed, rows, cols))
        cr, cc, dcam = _clouds(ctx, seed, rows, cols)
        a = ctx.align(cr, cc, dcam, rows, cols, identity16(), dap)
        b = ctx.align_noalloc(cr, cc, dcam, rows, cols, identity16(), dap)
        _same(a, b)
        # a guess other than the identity, and the same call twice in a row
        c = ctx.align_noalloc(cr, cc, dcam, rows, cols, a[0], dap)
        d = ctx.align(cr, cc, dcam, rows, cols, a[0], dap)
        _same(c, d)
        cr.close()
        cc.close()


def test_noalloc_rejects_large_inputs(ctx):
    say("case align_noalloc limits")
    _, dap = aparams()
    cr, cc, dcam = _clouds(ctx, 34, 240, 320)
    with pytest.raises(capi.NicpError, match="at most 307200"):
        ctx.align_noalloc(cr, cc, dcam, 960, 1280, identity16(), dap)
    cr.close()
    cc.close()
