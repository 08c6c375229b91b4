"""The projection strategies of batches and their autotuner (nicp_batch_set_projection,
nicp_batch_set_autotune): the global z-buffer with merged atomics (0), the same reading each
cell first (1), and the tile-binned projection fused with the correspondences (2) give the same
z-buffers, so every transform and result is bit-identical whichever strategy runs -- also when
the fused tiles' source lists overflow (capacity 0 or 2: those tiles scan every source tile),
at uneven resolutions with partial tiles, eagerly and from a captured graph. The tuner's choice
is the fastest measured strategy."""
import numpy as np
import pytest

from nicp_amd import capi, initnet, synth
from tests.helpers import aparams, cams, identity16, params

pytestmark = pytest.mark.gpu


def _batch16(ctx, max_rot, max_trans, init):
    B = 16
    batch = initnet.make_batch(ctx, B)
    initnet.render_pairs(batch, np.arange(2000, 2000 + B, dtype=np.uint64), max_rot, max_trans)
    batch.convert()
    ref = np.arange(0, 2 * B, 2, dtype=np.int32)
    cur = ref + 1
    init16 = None
    if init == "far":  # starts 0.8 m off: projections pile onto few pixels
        I = np.eye(4, dtype=np.float32)
        I[2, 3] = 0.8
        init16 = np.tile(I.flatten(order="F"), (B, 1))
    return batch, ref, cur, init16


def _same(a, b):
    Ta, ra = a
    Tb, rb = b
    np.testing.assert_array_equal(Ta.view(np.uint32), Tb.view(np.uint32))
    for x, y in zip(ra, rb):
        assert x.inliers == y.inliers and x.num_correspondences == y.num_correspondences
        assert x.overlap_pixels == y.overlap_pixels
        np.testing.assert_array_equal(np.array(x.H).view(np.uint64), np.array(y.H).view(np.uint64))
        np.testing.assert_array_equal(np.array(x.omega).view(np.uint64), np.array(y.omega).view(np.uint64))


@pytest.mark.parametrize("max_rot,max_trans,init", [(30.0, 0.5, "identity"), (8.0, 0.1, "far")])
def test_autotune_results_identical(ctx, max_rot, max_trans, init):
    print("case autotune", max_rot, max_trans, init, flush=True)
    batch, ref, cur, init16 = _batch16(ctx, max_rot, max_trans, init)
    ap = capi.align_params()
    r0 = batch.align(ref, cur, init16, ap, use_graph=False)
    batch.set_autotune(True)
    r1 = batch.align(ref, cur, init16, ap, use_graph=True)   # tuning call (eager)
    r2 = batch.align(ref, cur, init16, ap, use_graph=True)   # tuned, captured
    r3 = batch.align(ref, cur, init16, ap, use_graph=False)  # tuned, eager
    for r in (r1, r2, r3):
        _same(r, r0)
    state = batch.autotune_state()
    print("autotune state", state, flush=True)
    assert len(state) == 1
    choice, m0, m1, m2 = state[0]
    assert m0 > 0 and m1 > 0 and m2 > 0 and choice == int(np.argmin([m0, m1, m2]))
    batch.set_autotune(False)
    assert batch.autotune_state() == []
    batch.close()


@pytest.mark.parametrize("max_rot,max_trans,init", [(30.0, 0.5, "identity"), (8.0, 0.1, "far")])
def test_projection_strategies_identical(ctx, max_rot, max_trans, init):
    print("case strategies", max_rot, max_trans, init, flush=True)
    batch, ref, cur, init16 = _batch16(ctx, max_rot, max_trans, init)
    ap = capi.align_params()
    batch.set_projection(0)
    base = batch.align(ref, cur, init16, ap, use_graph=False)
    for strategy in (1, 2):
        batch.set_projection(strategy)
        for graph in (False, True):
            _same(batch.align(ref, cur, init16, ap, use_graph=graph), base)
    # overflowing source lists: every fused tile (capacity 0) or most (capacity 2) scans all
    batch.set_projection(2)
    for cap in (0, 2):
        batch.set_bin_capacity(cap)
        _same(batch.align(ref, cur, init16, ap, use_graph=False), base)
    batch.set_bin_capacity(256)
    # inner iterations and the stored correspondences of the last iteration
    ap2 = capi.align_params(outer_iterations=4, inner_iterations=3)
    batch.set_projection(0)
    b2 = batch.align(ref, cur, init16, ap2, use_graph=False)
    batch.set_projection(2)
    _same(batch.align(ref, cur, init16, ap2, use_graph=True), b2)
    with pytest.raises(capi.NicpError):
        batch.set_projection(3)
    with pytest.raises(capi.NicpError):
        batch.set_bin_capacity(257)
    batch.close()


def test_projection_strategies_uneven(ctx):
    """Partial fused and correspondence tiles (160x120, 320x240 with 64x64 tiles), 1280x960, an
    empty frame and a ragged pair list in one batch."""
    print("case strategies uneven", flush=True)
    shapes = [(120, 160), (240, 320), (480, 640), (960, 1280), (120, 160), (37, 53)]
    _, dsp = params()
    frames, cam_list = [], []
    for i, (rows, cols) in enumerate(shapes):
        if rows * cols >= 60 * 80:
            dr, dc, _ = synth.pair(70 + i, rows, cols, max_rot_deg=8.0, max_trans_m=0.1)
        else:
            dr = dc = np.zeros((rows, cols), np.uint16)
        frames += [dr, dc]
        cam_list += [cams(rows, cols)[1]] * 2
    b = capi.Batch(ctx, [f.shape for f in frames], cam_list, dsp)
    for k, f in enumerate(frames):
        b.upload(k, f)
    b.convert()
    ref = np.arange(0, len(frames), 2)
    cur = ref + 1
    _, ap = aparams(outer_iterations=6)
    b.set_projection(0)
    base = b.align(ref, cur, None, ap, use_graph=False)
    b.set_projection(2)
    _same(b.align(ref, cur, None, ap, use_graph=False), base)
    _same(b.align(ref, cur, None, ap, use_graph=True), base)
    b.set_bin_capacity(1)
    _same(b.align(ref, cur, None, ap, use_graph=False), base)
    # single-pair calls on the same clouds run strategy 0's kernels
    for i, (rows, cols) in enumerate(shapes[:4]):
        cr = b.frame_cloud(2 * i, (rows, cols))
        cc = b.frame_cloud(2 * i + 1, (rows, cols))
        Ts, _ = ctx.align(cr, cc, cams(rows, cols)[1], rows, cols, identity16(), ap)
        np.testing.assert_array_equal(Ts.view(np.uint32), base[0][i].view(np.uint32))
    b.close()
