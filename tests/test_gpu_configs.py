"""The batched and learned configurations at their own sizes, through the C ABI.

- Learned initial guess (640x480, motions up to 30 deg / 0.5 m): the trained network
  (nicp_amd/data/initnet_640x480.npy) against the identity start on 256 held-out seeded
  pairs; a sample of the refined transforms against the CPU oracle's aligner run from the
  same start on the same clouds.
- Uneven batch: 512 pairs uniform over 160x120, 320x240, 640x480 and 1280x960 (intrinsics
  scaled with the width), motions up to 8 deg / 10 cm, 15 iterations, autotuned and replayed
  from a hipGraph; pairs of every resolution against single-pair calls (bit for bit) and
  against the oracle (1e-5 rad / 1e-5 m).
- A captured graph replayed after the batch's block maps were reallocated by a larger call.
- A batch with a frame beyond the packed source pixels' size limit (2^23 pixels).

Reference: AlignerProjective::align, nicp/nicp/alignerprojective.cpp:21-136, and the
initial guess it starts from, nicp/nicp/aligner.h:123-141.
"""
import sys

import numpy as np
import pytest

import oracle as O
from nicp_amd import capi, initnet, synth
from tests.helpers import aparams, cams, identity16, params

pytestmark = pytest.mark.gpu

RES = [(120, 160), (240, 320), (480, 640), (960, 1280)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _errors(T, Tgt):
    return np.array([synth.pose_error(T[i], Tgt[i]) if np.all(np.isfinite(T[i])) else (np.inf, np.inf)
                     for i in range(len(T))])


def _aligned(e):
    return (e[:, 0] < 0.1) & (e[:, 1] < 0.002)


def _oracle_cloud(cl):
    pts, nrm, curv, op, on = cl.download()
    nrm = nrm.copy()
    nrm[:, 3] = 0.0
    return O.cloud_from_arrays(pts, nrm, curv, op, on)


def _oracle_align_on_device_clouds(b, f_ref, f_cur, shape, init16, ap_oracle, cam):
    rows, cols = shape
    cr = b.frame_cloud(f_ref, shape)
    cc = b.frame_cloud(f_cur, shape)
    # one thread: the reference's OpenMP partition drops rows and correspondences with more
    # (correspondencefinderprojective.cpp:41-53, linearizer.cpp:36-43); the device drops none
    T, res = O.align(_oracle_cloud(cr), _oracle_cloud(cc), rows, cols, cam, init16, ap_oracle, 1, fp64_solve=True)
    cr.close()
    cc.close()
    return T, res


def _reference_spread(b, f_ref, f_cur, shape, init16, ap_oracle, cam):
    """How far the reference's own result moves with the rounding of its sums on these clouds: the
    largest distance between its FP32 loop at 1, 2, 4 and 8 OpenMP threads (the thread count
    regroups the partial sums) and the FP64-solve design: (rad, m)."""
    rows, cols = shape
    cr = b.frame_cloud(f_ref, shape)
    cc = b.frame_cloud(f_cur, shape)
    ocr, occ = _oracle_cloud(cr), _oracle_cloud(cc)
    Ts = [O.align(ocr, occ, rows, cols, cam, init16, ap_oracle, n)[0] for n in (1, 2, 4, 8)]
    Ts.append(O.align(ocr, occ, rows, cols, cam, init16, ap_oracle, 1, fp64_solve=True)[0])
    cr.close()
    cc.close()
    sr = st = 0.0
    for i in range(len(Ts)):
        for j in range(i):
            rot, tr = synth.pose_error(Ts[i], Ts[j])
            sr, st = max(sr, np.radians(rot)), max(st, tr)
    return sr, st


def _one_iteration_vs_oracle(ctx, b, f_ref, f_cur, shape, cam_dev, cam_oracle):
    """One outer iteration from the identity on the batch's clouds, device against the FP64-solve
    oracle: the correspondences are those of the same start, so only the arithmetic differs."""
    rows, cols = shape
    oap1, dap1 = aparams(outer_iterations=1)
    cr = b.frame_cloud(f_ref, shape)
    cc = b.frame_cloud(f_cur, shape)
    Td, _ = ctx.align(cr, cc, cam_dev, rows, cols, identity16(), dap1)
    To, _ = O.align(_oracle_cloud(cr), _oracle_cloud(cc), rows, cols, cam_oracle, identity16(), oap1, 1,
                    fp64_solve=True)
    cr.close()
    cc.close()
    rot, tr = synth.pose_error(Td, To)
    return np.radians(rot), tr


def test_learned_gain_640x480(ctx):
    """What the trained network gains over the identity start on held-out pairs (seeds from 1000;
    training used seeds from 10,000,000)."""
    w = initnet.load_weights()
    assert w is not None, "trained weights missing"
    n_pairs, B = 256, 128
    say("case learned_gain 640x480 pairs=%d" % n_pairs)
    net = capi.InitNet(ctx, seed=1)
    net.set_params(w)
    _, ap = aparams()
    oap, _ = aparams()
    ocam, _ = cams(480, 640)
    b = initnet.make_batch(ctx, B)
    ref = np.arange(0, 2 * B, 2, dtype=np.int32)
    cur = ref + 1
    e_net, e_id, used_all = [], [], []
    sample = None
    for k in range(n_pairs // B):
        seeds = np.arange(1000 + k * B, 1000 + (k + 1) * B, dtype=np.uint64)
        Tgt = initnet.render_pairs(b, seeds, 30.0, 0.5)
        b.convert()
        Tn, rn, starts, used = b.align_learned(net, ref, cur, ap, use_graph=True)
        Ti, _ = b.align(ref, cur, None, ap, use_graph=True)
        en, ei = _errors(Tn, Tgt), _errors(Ti, Tgt)
        e_net.append(en)
        e_id.append(ei)
        used_all.append(used)
        if k == 0:
            # oracle check: four pairs the learned path aligned, from the start it chose
            idx = np.nonzero(_aligned(en))[0][:4]
            sample = []
            for i in idx:
                To, _ = _oracle_align_on_device_clouds(b, int(ref[i]), int(cur[i]), (480, 640), starts[i], oap, ocam)
                sample.append(synth.pose_error(Tn[i], To))
    e_net, e_id, used = np.concatenate(e_net), np.concatenate(e_id), np.concatenate(used_all)
    a_net, a_id = int(_aligned(e_net).sum()), int(_aligned(e_id).sum())
    m_net, m_id = float(np.median(e_net[:, 0])), float(np.median(e_id[:, 0]))
    say("  aligned: learned %d, identity %d of %d; median rotation error learned %.3f deg, identity %.3f deg; "
        "network start kept for %d pairs (fallback rate %.3f)" % (a_net, a_id, n_pairs, m_net, m_id, int(used.sum()),
                                                                   1.0 - used.mean()))
    say("  oracle from the same start (aligned sample): " + ", ".join("%.2e rad %.2e m" % (np.radians(r), t)
                                                                       for r, t in sample))
    assert a_net >= a_id
    assert m_net < m_id
    assert len(sample) >= 2
    for r, t in sample:
        assert np.radians(r) < 1e-5 and t < 1e-5
    b.close()
    net.close()


def test_uneven_batch_512(ctx):
    """512 pairs of the four resolutions in one batch, 15 iterations, autotuned, graph replay.
    Against the FP64-solve oracle after 15 iterations the bar is 1e-5 rad / 1e-5 m, or twice the
    reference's own spread (its FP32 loop across OpenMP thread counts and its FP64 design) where
    that is larger: a different rounding of H flips a few gate-boundary correspondences, and over
    15 iterations that moves a 1280x960 pair's fixed point by up to ~2e-5 m. After one iteration
    from the same start the correspondences are the same, and the bar is 1e-6."""
    n_pairs = 512
    say("case uneven_batch pairs=%d" % n_pairs)
    rng = np.random.default_rng(7)
    pick = rng.integers(0, len(RES), n_pairs)
    shapes, cam_list = [], []
    for k in pick:
        r, c = RES[k]
        cam = cams(r, c)[1]
        shapes += [(r, c), (r, c)]
        cam_list += [cam, cam]
    _, dsp = params()
    oap, dap = aparams(outer_iterations=15)
    b = capi.Batch(ctx, shapes, cam_list, dsp)
    seeds = np.arange(2_000_000, 2_000_000 + n_pairs, dtype=np.uint64)
    Tgt = np.stack([synth.motion(int(s), 8.0, 0.1) for s in seeds])
    poses = np.empty((2 * n_pairs, 16), np.float32)
    poses[0::2] = identity16()
    poses[1::2] = Tgt
    b.render_synthetic(np.arange(2 * n_pairs, dtype=np.int32), np.repeat(seeds, 2),
                       np.tile(np.array([10, 11], np.uint64), n_pairs), poses)
    ref = np.arange(0, 2 * n_pairs, 2, dtype=np.int32)
    cur = ref + 1
    b.convert()
    b.set_autotune(True)
    Te, re = b.align(ref, cur, None, dap, use_graph=True)   # tuning call (eager)
    Tg, rg = b.align(ref, cur, None, dap, use_graph=True)   # captured
    Tg2, _ = b.align(ref, cur, None, dap, use_graph=True)   # replayed
    np.testing.assert_array_equal(Tg.view(np.uint32), Te.view(np.uint32))
    np.testing.assert_array_equal(Tg2.view(np.uint32), Te.view(np.uint32))
    e = _errors(Tg, Tgt)
    ok = _aligned(e)
    for k, (r, c) in enumerate(RES):
        sel = pick == k
        say("  %dx%d: %d pairs, aligned %.3f, median rotation error %.4f deg" % (
            c, r, int(sel.sum()), float(ok[sel].mean()), float(np.median(e[sel, 0]))))
    # per resolution: two pairs against single-pair calls, one against the oracle
    for k, (r, c) in enumerate(RES):
        idx = np.nonzero(pick == k)[0]
        ocam, dcam = cams(r, c)
        for i in idx[:2]:
            cr = b.frame_cloud(int(ref[i]), (r, c))
            cc = b.frame_cloud(int(cur[i]), (r, c))
            Ts, rs = ctx.align(cr, cc, dcam, r, c, identity16(), dap)
            np.testing.assert_array_equal(Tg[i].view(np.uint32), Ts.view(np.uint32))
            assert rg[i].inliers == rs.inliers
            cr.close()
            cc.close()
        # the oracle on the same clouds: an aligned pair of this resolution when there is one
        al = idx[ok[idx]]
        i = int(al[0]) if len(al) else int(idx[0])
        To, ro = _oracle_align_on_device_clouds(b, int(ref[i]), int(cur[i]), (r, c), identity16(), oap, ocam)
        rot, tr = synth.pose_error(Tg[i], To)
        sr, st = _reference_spread(b, int(ref[i]), int(cur[i]), (r, c), identity16(), oap, ocam)
        r1, t1 = _one_iteration_vs_oracle(ctx, b, int(ref[i]), int(cur[i]), (r, c), dcam, ocam)
        say("  %dx%d pair %d vs oracle: %.2e rad %.2e m (aligned %s; reference spread %.2e rad %.2e m; "
            "one iteration %.2e rad %.2e m)" % (c, r, i, np.radians(rot), tr, bool(ok[i]), sr, st, r1, t1))
        assert r1 < 1e-6 and t1 < 1e-6
        if ok[i]:
            assert np.radians(rot) < max(1e-5, 2 * sr) and tr < max(1e-5, 2 * st)
    for k in (1, 2, 3):
        assert ok[pick == k].sum() >= 1
    b.close()


def test_graph_replay_after_map_growth(ctx):
    """A graph captured for one plan, then an eager call whose larger reference frame grows (and so
    reallocates) the batch's block maps, then the first plan again: the replay must use the new
    maps (it is captured again) and give the eager result."""
    say("case graph_replay_after_map_growth")
    _, dsp = params()
    _, dap = aparams()
    shapes = [(240, 320), (240, 320), (960, 1280)]
    cam_list = [cams(r, c)[1] for r, c in shapes]
    b = capi.Batch(ctx, shapes, cam_list, dsp)
    dr, dc, _ = synth.pair(41, 240, 320)
    big = synth.pair(42, 960, 1280)[0]
    for f, d in enumerate((dr, dc, big)):
        b.upload(f, d)
    b.convert()
    Te, _ = b.align([0], [1], None, dap, use_graph=False)
    Tg1, _ = b.align([0], [1], None, dap, use_graph=True)
    b.align([2], [1], None, dap, use_graph=False)           # 1200 reference tiles: maps grow
    Tg2, _ = b.align([0], [1], None, dap, use_graph=True)
    np.testing.assert_array_equal(Tg1.view(np.uint32), Te.view(np.uint32))
    np.testing.assert_array_equal(Tg2.view(np.uint32), Te.view(np.uint32))
    b.close()


def test_frame_beyond_packed_source_limit(ctx):
    """A frame of more than 2^23 pixels has no packed source pixels (their cloud index shares a
    32-bit word with the pixel's offset in its source tile), so a batch call with it takes the
    global z-buffer even when the tile-binned projection is forced: its pair equals the
    single-pair call bit for bit, and a small pair of the same call equals the same pair aligned
    with the tile-binned projection in a batch of small frames."""
    shapes = [(2112, 4096), (2112, 4096), (240, 320), (240, 320)]
    say("case frame_beyond_packed_source_limit pixels=%d" % (2112 * 4096))
    assert 2112 * 4096 > (1 << 23)
    _, dsp = params()
    _, dap = aparams()
    cam_list = [cams(r, c)[1] for r, c in shapes]
    b = capi.Batch(ctx, shapes, cam_list, dsp)
    big = synth.pair(51, 2112, 4096)
    small = synth.pair(52, 240, 320)
    for f, d in enumerate((big[0], big[1], small[0], small[1])):
        b.upload(f, d)
    b.convert()
    b.set_projection(2)
    T, _ = b.align([0, 2], [1, 3], None, dap, use_graph=False)
    cr, cc = b.frame_cloud(0, shapes[0]), b.frame_cloud(1, shapes[1])
    Ts, _ = ctx.align(cr, cc, cam_list[0], 2112, 4096, identity16(), dap)
    cr.close()
    cc.close()
    np.testing.assert_array_equal(T[0].view(np.uint32), Ts.view(np.uint32))
    rot, tr = synth.pose_error(T[0], big[2])
    say("  big pair: rot err %.4f deg, trans err %.5f m" % (rot, tr))
    assert rot < 1.0 and tr < 0.02
    b.close()
    # the small pair in a batch of small frames, where the tile-binned projection runs
    b2 = capi.Batch(ctx, shapes[2:], cam_list[2:], dsp)
    b2.upload(0, small[0])
    b2.upload(1, small[1])
    b2.convert()
    b2.set_projection(2)
    T2, _ = b2.align([0], [1], None, dap, use_graph=False)
    np.testing.assert_array_equal(T[1].view(np.uint32), T2[0].view(np.uint32))
    b2.close()


def test_ground_truth_ceiling_640x480(ctx):
    """The accuracy ceiling of the learned workload at its size: 1024 held-out 640x480 pairs
    (<=30 deg / 0.5 m) aligned from the ground truth, from the identity and from the learned
    start. No start can do better than the ground truth by more than chance, and at these motions
    NICP's own fixed point leaves most pairs beyond 0.1 deg even from there."""
    w = initnet.load_weights()
    assert w is not None, "trained weights missing"
    B, n_batches = 128, 8
    say("case ground_truth_ceiling 640x480 pairs=%d" % (B * n_batches))
    net = capi.InitNet(ctx, seed=1)
    net.set_params(w)
    _, ap = aparams()
    b = initnet.make_batch(ctx, B)
    ref = np.arange(0, 2 * B, 2, dtype=np.int32)
    cur = ref + 1
    a_gt = a_id = a_net = 0
    drift = []
    for k in range(n_batches):
        seeds = np.arange(1000 + k * B, 1000 + (k + 1) * B, dtype=np.uint64)
        Tgt = initnet.render_pairs(b, seeds, 30.0, 0.5)
        b.convert()
        Tg, _ = b.align(ref, cur, np.ascontiguousarray(Tgt, np.float32), ap, use_graph=True)
        Ti, _ = b.align(ref, cur, None, ap, use_graph=True)
        Tn, _, _, _ = b.align_learned(net, ref, cur, ap, use_graph=True)
        eg = _errors(Tg, Tgt)
        drift.append(eg[:, 0])
        a_gt += int(_aligned(eg).sum())
        a_id += int(_aligned(_errors(Ti, Tgt)).sum())
        a_net += int(_aligned(_errors(Tn, Tgt)).sum())
    drift = np.concatenate(drift)
    n = B * n_batches
    say("  aligned (<0.1 deg, <2 mm) of %d: from the ground truth %d, learned %d, identity %d; drift from the ground "
        "truth: median %.3f deg, 90th percentile %.3f deg" % (n, a_gt, a_net, a_id, float(np.median(drift)),
                                                             float(np.percentile(drift, 90))))
    assert a_net <= a_gt + n // 100 and a_id <= a_gt + n // 100
    assert a_gt < n // 4
    b.close()
    net.close()


def test_profiling_with_graphs(ctx):
    """Per-launch timing (nicp_ctx_profile) on a batch whose conversion, learned step and
    refinement otherwise replay hipGraphs: the launches then run eagerly, every stage is timed,
    and the transforms equal the graph replays' bit for bit."""
    say("case profiling with graphs")
    w = initnet.load_weights()
    assert w is not None, "trained weights missing"
    net = capi.InitNet(ctx, seed=1)
    net.set_params(w)
    _, ap = aparams()
    B = 16
    b = initnet.make_batch(ctx, B)
    ref = np.arange(0, 2 * B, 2, dtype=np.int32)
    cur = ref + 1
    initnet.render_pairs(b, np.arange(4000, 4000 + B, dtype=np.uint64), 30.0, 0.5)
    b.convert()
    Tl, _, sl, ul = b.align_learned(net, ref, cur, ap, use_graph=True)
    Tp, _ = b.align(ref, cur, None, ap, use_graph=True)
    ctx.profile(True)
    b.convert()
    Tl2, _, sl2, ul2 = b.align_learned(net, ref, cur, ap, use_graph=True)
    Tp2, _ = b.align(ref, cur, None, ap, use_graph=True)
    prof = ctx.profile_read()
    ctx.profile(False)
    np.testing.assert_array_equal(Tl2.view(np.uint32), Tl.view(np.uint32))
    np.testing.assert_array_equal(sl2.view(np.uint32), sl.view(np.uint32))
    np.testing.assert_array_equal(np.asarray(ul2), np.asarray(ul))
    np.testing.assert_array_equal(Tp2.view(np.uint32), Tp.view(np.uint32))
    timed = {k: v for k, v in prof.items() if v[1] > 0}
    say("  timed stages: " + ", ".join("%s %d x %.3f ms" % (k, v[1], v[0] / v[1]) for k, v in sorted(timed.items())))
    for k in ("k_stats", "k_integral_frame", "initnet_forward", "start_probe", "k_solve", "k_solution_stats"):
        assert timed.get(k, (0, 0))[1] > 0, k
    # the refinement ran twice with 10 iterations: 20 projections whichever strategy
    assert sum(timed.get(k, (0, 0))[1] for k in ("k_proj_corr_lin", "k_corr_lin")) == 20
    b.close()
    net.close()


def test_learned_iterations_to_align(ctx):
    """The learned start's ablation in iterations: on 256 held-out 640x480 pairs (<=30 deg /
    0.5 m), the pairs under 0.1 deg / 2 mm and the median rotation error after 2, 4, 6 and 10
    outer iterations, from the learned start and from the identity."""
    w = initnet.load_weights()
    assert w is not None, "trained weights missing"
    say("case learned iterations to align 640x480 pairs=256")
    net = capi.InitNet(ctx, seed=1)
    net.set_params(w)
    B = 128
    b = initnet.make_batch(ctx, B)
    ref = np.arange(0, 2 * B, 2, dtype=np.int32)
    cur = ref + 1
    iters = (2, 4, 6, 10)
    counts = {("net", k): 0 for k in iters}
    counts.update({("id", k): 0 for k in iters})
    med = {}
    errs = {key: [] for key in counts}
    for k in range(2):
        Tgt = initnet.render_pairs(b, np.arange(1000 + k * B, 1000 + (k + 1) * B, dtype=np.uint64), 30.0, 0.5)
        b.convert()
        for it in iters:
            _, ap = aparams(outer_iterations=it)
            Tn, _, _, _ = b.align_learned(net, ref, cur, ap, use_graph=True)
            Ti, _ = b.align(ref, cur, None, ap, use_graph=True)
            for name, T in (("net", Tn), ("id", Ti)):
                e = _errors(T, Tgt)
                counts[(name, it)] += int(_aligned(e).sum())
                errs[(name, it)].append(e[:, 0])
    for key in errs:
        med[key] = float(np.median(np.concatenate(errs[key])))
    for it in iters:
        say("  %2d iterations: aligned learned %3d identity %3d of 256; median rotation error learned %.3f deg "
            "identity %.3f deg" % (it, counts[("net", it)], counts[("id", it)], med[("net", it)], med[("id", it)]))
    for it in iters:
        assert counts[("net", it)] >= counts[("id", it)] - 2
        assert med[("net", it)] < med[("id", it)]
    b.close()
    net.close()
