"""The projection autotuner of batches (nicp_batch_set_autotune): both z-buffer strategies give
the same z-buffers, so every result is bit-identical with the tuner on or off, whichever
strategy it picks; the choice is the faster measured one."""
import numpy as np
import pytest

from nicp_amd import capi, initnet

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("max_rot,max_trans,init", [(30.0, 0.5, "identity"), (8.0, 0.1, "far")])
def test_autotune_results_identical(ctx, max_rot, max_trans, init):
    print("case autotune", max_rot, max_trans, init, flush=True)
    B = 16
    batch = initnet.make_batch(ctx, B)
    initnet.render_pairs(batch, np.arange(2000, 2000 + B, dtype=np.uint64), max_rot, max_trans)
    batch.convert()
    ref = np.arange(0, 2 * B, 2, dtype=np.int32)
    cur = ref + 1
    ap = capi.align_params()
    init16 = None
    if init == "far":  # starts 0.8 m off: projections pile onto few pixels
        I = np.eye(4, dtype=np.float32)
        I[2, 3] = 0.8
        init16 = np.tile(I.flatten(order="F"), (B, 1))
    T0, r0 = batch.align(ref, cur, init16, ap, use_graph=False)
    batch.set_autotune(True)
    T1, r1 = batch.align(ref, cur, init16, ap, use_graph=True)   # tuning call (eager)
    T2, r2 = batch.align(ref, cur, init16, ap, use_graph=True)   # tuned, captured
    T3, r3 = batch.align(ref, cur, init16, ap, use_graph=False)  # tuned, eager
    for T in (T1, T2, T3):
        assert np.array_equal(T.view(np.uint32), T0.view(np.uint32))
    for r in (r1, r2, r3):
        assert [x.inliers for x in r] == [x.inliers for x in r0]
    state = batch.autotune_state()
    print("autotune state", state, flush=True)
    assert len(state) == 1
    choice, m0, m1 = state[0]
    assert m0 > 0 and m1 > 0 and choice == (1 if m1 < m0 else 0)
    batch.set_autotune(False)
    assert batch.autotune_state() == []
