"""Benchmark: aligned 640x480 depth pairs per second on one MI355X, batched NICP
from a learned initial guess against plain single-pair NICP.

Workload: held-out seeded synthetic pairs (20-plane/box scenes, second camera moved
by up to 30 deg / 0.5 m, depth noise sigma = 1.5e-3 z^2, 10 % holes), rendered on
the device before the timed region (input generation, not timed) into one packed device
store that holds every step's frames.
One step = the registration path over one batch of pairs resident in HBM: the copy of
its depth images from the store into a batch, both
frames of every pair converted to clouds (unprojection, integral-image normals,
information matrices), the network input and forward pass of the learned initial
guess, one linearisation at the predicted and at the identity start to choose the
start, 10 Gauss-Newton iterations of projective NICP from the chosen start, and
Aligner::_computeStatistics (the statistics linearisation and the FP64 unscented
solution statistics, on the device) -- one launch per stage for the whole batch, with
every pair's transform and result record copied back. The timed region keeps --streams
batches in flight (default 2: one context, HIP stream and host thread each, step k on lane
k mod 2), so one batch's conversion overlaps another's refinement; ms_per_step is the wall
time of all timed steps divided by their number. A pair counts as aligned when its
rotation error is below 0.1 deg and its translation error below 2 mm against the
sampled ground truth.

Prints one JSON line. Besides the contract keys: the plain baselines (single pair
per call from the identity, and batched from the identity = the same pairs without
the network), per-kernel device times from HIP events recorded on the library's
stream in a separate single-stream eager pass over the first timed batches (its
transforms checked equal to the timed ones), the HBM roofline of the dominant kernel, the
matrix-core rate of the network's forward pass, and the CPU oracle on a bounded
sample on the host cores.
"""
import argparse
import json
import os
import re
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
ROWS, COLS = 480, 640
MAX_ROT_DEG, MAX_TRANS_M = 30.0, 0.5
EVAL_SEED_BASE = 1000          # training pairs use seeds from 10,000,000 on
WARMUP_SEED_BASE = 900000
# multiply-adds of the network's forward pass per pair (conv 1..4, fc 1, fc 2), x2 flops
NET_FLOPS_PER_PAIR = 2 * (1200 * 19 * 32 + 300 * 289 * 32 + 80 * 289 * 64 + 20 * 577 * 64 + 1281 * 256 + 257 * 6)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from its committed rocprofv3 counter summary
    (profiles/<kernel>_pmc.md: FETCH_SIZE and WRITE_SIZE scaled by a copy of known size),
    measured on the default workload; None when there is no such summary."""
    path = os.path.join(ROOT, "profiles", "%s_pmc.md" % kernel)
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"^%s_traffic_bytes_per_launch = (\d+)$" % re.escape(kernel), text, re.M)
    return int(m.group(1)) if m else None


def kernel_bytes(name, pairs_info, frames_info):
    """Algorithmic HBM bytes of one launch of a kernel (covering the whole batch)."""
    P = sum(f[0] for f in frames_info)
    N = sum(f[1] for f in frames_info)
    if name == "k_corr_lin":
        # z-buffer key read + reset (16 B) and current index (4 B) per pixel; per overlap pixel the
        # reference point + normal/curvature (28 B), the current point + normal/curvature (28 B)
        # and both information matrices (48 B)
        return sum(20 * p + 104 * c2 for p, _, c2, _ in pairs_info)
    if name == "k_proj_corr_lin":
        # one outer iteration's projection, correspondences and linearisation: the per-iteration
        # figure B_iter = 24 P + 12 N + 104 C2 of the algorithm (per projector pixel the z-buffer's
        # initialisation, atomic and read, 8 B each, with both index images; 12 B per reference
        # point; per overlap pixel both points and normals with curvature, 2 x 28 B, and the
        # current point's two information matrices, 48 B)
        return sum(24 * p + 12 * n + 104 * c2 for p, n, c2, _ in pairs_info)
    if name == "k_proj_corr_lin(reads)":
        # what the fused kernel itself reads for the same work: the reference frame's packed
        # source pixels (depth and index, 8 B per point) for the projection, the current index
        # (4 B per pixel) and per overlap pixel both points and normals (2 x 32 B); its z-buffer
        # stays in LDS and the information matrices of converted frames are formed from the
        # normals
        return sum(8 * n + 4 * p + 64 * c2 for p, n, c2, _ in pairs_info)
    if name == "k_stats":
        # depth (2 B) + index (4 B) + one pass over the 40 B integral image per pixel; normal +
        # curvature (16 B) per point (a batch conversion leaves the information matrices to the
        # closed form its linearisers use)
        return 46 * P + 16 * N
    if name == "k_integral_x":
        return 42 * P
    if name == "k_integral_y":
        return 80 * P
    if name == "k_integral_frame":
        # both prefix passes at once: depth in (2 B), the 40 B integral image out, per pixel
        return 42 * P
    if name == "k_unproject":
        return 6 * P + 16 * N
    if name == "k_count":
        return 2 * P
    if name == "k_src_tiles":
        # depth (2 B) and index (4 B) per pixel in, the packed source pixels (8 B per point) out
        return 6 * P + 8 * N
    if name == "k_zproject(reference)":
        # reference points (16 B) and one 8 B atomic per point
        return sum(24 * n for _, n, _, _ in pairs_info)
    return None


def aligned(T, Tgt):
    """Aligned mask and (rotation deg, translation m) errors; a diverged pair (non-finite T, as
    the reference's quat2mat gives for an update beyond the unit quaternion) has infinite error."""
    from nicp_amd import synth
    e = np.array([synth.pose_error(T[i], Tgt[i]) if np.all(np.isfinite(T[i])) else (np.inf, np.inf)
                  for i in range(len(T))])
    return (e[:, 0] < 0.1) & (e[:, 1] < 0.002), e


def host_cpu():
    """(model name, CPUs this process may run on, CPUs of the machine)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return model, usable, os.cpu_count() or usable


def cpu_leg(rows, cols, threads, seeds, warmup):
    """Median seconds per pair of the CPU oracle: both DepthImageConverterIntegralImage::compute
    calls and a 10-iteration AlignerProjective::align from the identity (scene rendering excluded)."""
    import oracle as O
    from nicp_amd import synth
    fx, fy, cx, cy = synth.camera_for(rows, cols)
    cam = O.camera(fx, fy, cx, cy)
    sp = O.stats_params()
    ap = O.align_params()
    I = np.eye(4, dtype=np.float32).flatten(order="F")
    O.set_num_threads(threads)
    times = []
    for k, s in enumerate(seeds):
        dr, dc, _ = synth.pair(int(s), rows, cols, max_rot_deg=5.0, max_trans_m=0.05)
        t0 = time.perf_counter()
        fr = O.Frame(dr, 0.001, cam, sp)
        fc = O.Frame(dc, 0.001, cam, sp)
        O.align(fr.cloud(), fc.cloud(), rows, cols, cam, I, ap, threads)
        dt = time.perf_counter() - t0
        if k >= warmup:
            times.append(dt)
    return float(np.median(times))


def cpu_baseline(n_pairs=32, warmup=3):
    """The CPU oracle (the reference's FP32 arithmetic and OpenMP structure, built with the
    reference's Release flags) on the host: 320x240 (fx=fy=262.5) and 640x480 (fx=fy=525) pairs
    with motions <= 5 deg / 5 cm, the median of n_pairs seeded pairs after `warmup` pairs, with all
    of this process's CPUs (OMP_NUM_THREADS when set) and with one thread. `value` is the 640x480
    leg on all threads; the iteration count is fixed (10, no early exit), so the work per pair does
    not depend on the motion range."""
    model, usable, machine = host_cpu()
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(env) if env.isdigit() and int(env) > 0 else usable
    seeds = np.arange(EVAL_SEED_BASE, EVAL_SEED_BASE + warmup + n_pairs)
    legs = []
    for rows, cols in ((240, 320), (480, 640)):
        for th in (threads, 1):
            sec = cpu_leg(rows, cols, th, seeds, warmup)
            legs.append({"resolution": "%dx%d" % (cols, rows), "threads": th, "median_ms_per_pair": round(1000 * sec, 2),
                         "pairs_per_s": round(1.0 / sec, 3)})
            log("cpu oracle %dx%d, %d threads: %.1f ms per pair" % (cols, rows, th, 1000 * sec))
    main = next(l for l in legs if l["resolution"] == "640x480" and l["threads"] == threads)
    return {"value": main["pairs_per_s"], "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": "median of %d seeded pairs after %d warm-up pairs per leg (rot<=5deg, trans<=5cm): 2 "
                      "conversions + 10-iteration align from the identity each; legs 320x240 and 640x480, "
                      "%d threads and 1 thread" % (n_pairs, warmup, threads),
            "legs": legs, "cpu_model": model, "cpus_usable": usable, "cpus_machine": machine}


def parity_sample(ctx, seeds, max_rot, max_trans):
    """Per-run parity counts on pairs of the headline workload (640x480): normals of the device
    conversion beyond 1e-5 of the CPU oracle's, points whose normal is present on one side only
    (curvature at the stats threshold), and correspondences that differ from the oracle's on
    identical clouds at the ground-truth transform (gate-boundary mismatches)."""
    import oracle as O
    from nicp_amd import capi, synth
    fx, fy, cx, cy = synth.camera_for(ROWS, COLS)
    ocam, dcam = O.camera(fx, fy, cx, cy), capi.camera(fx, fy, cx, cy)
    osp, dsp = O.stats_params(), capi.stats_params()
    oap, dap = O.align_params(), capi.align_params()
    I = np.eye(4, dtype=np.float32).flatten(order="F")
    out = {"pairs": len(seeds), "points": 0, "normals_beyond_1e-5": 0, "normal_presence_flips": 0,
           "max_normal_err": 0.0, "correspondences": 0, "correspondence_mismatches": 0}
    for s in seeds:
        dr, dc, Tgt = synth.pair(int(s), ROWS, COLS, max_rot_deg=max_rot, max_trans_m=max_trans)
        fo = []
        for d in (dr, dc):
            f = O.Frame(d, 0.001, ocam, osp)
            cl = ctx.cloud_from_depth(d, dcam, dsp)
            _, nrm, _, _, _ = cl.download()
            cl.close()
            has_o = np.abs(f.normals[:, :3]).sum(1) > 0
            has_d = np.abs(nrm[:, :3]).sum(1) > 0
            both = has_o & has_d
            err = np.abs(nrm[both, :3] - f.normals[both, :3]).max(axis=1) if both.any() else np.zeros(1)
            out["points"] += int(f.n)
            out["normals_beyond_1e-5"] += int((err > 1e-5).sum())
            out["normal_presence_flips"] += int((has_o != has_d).sum())
            out["max_normal_err"] = max(out["max_normal_err"], float(err.max()))
            fo.append(f)
        fr, fc = fo
        KRt, _ = O.projector_matrices(np.array(ocam.K, np.float32), Tgt)
        ri, _ = O.project(fr.points, KRt, ROWS, COLS, ocam.min_distance, ocam.max_distance)
        KRt0, _ = O.projector_matrices(np.array(ocam.K, np.float32), I)
        ci, _ = O.project(fc.points, KRt0, ROWS, COLS, ocam.min_distance, ocam.max_distance)
        Tinv = np.linalg.inv(Tgt.reshape(4, 4, order="F").astype(np.float64)).astype(np.float32).flatten(order="F")
        op = O.correspondences(fr.cloud(), fc.cloud(), ri, ci, Tinv, oap, 1)
        ur = ctx.cloud_from_host(fr.points, fr.normals, fr.curvature, fr.omega_p, fr.omega_n)
        uc = ctx.cloud_from_host(fc.points, fc.normals, fc.curvature, fc.omega_p, fc.omega_n)
        dp = ctx.correspondences(ur, uc, ri, ci, Tinv, dap)
        ur.close()
        uc.close()
        so = set(map(tuple, op.tolist()))
        sd = set(map(tuple, dp.tolist()))
        out["correspondences"] += len(so)
        out["correspondence_mismatches"] += len(so ^ sd)
    out["max_normal_err"] = float("%.3g" % out["max_normal_err"])
    return out


def uneven_batch(ctx, n_pairs, iters, steps):
    """Batched plain NICP over pairs of uneven resolution in one launch per stage (graph replay)."""
    from nicp_amd import capi, synth
    res = [(120, 160), (240, 320), (480, 640), (960, 1280)]
    rng = np.random.default_rng(7)
    pick = rng.integers(0, len(res), n_pairs)
    shapes, cams = [], []
    for k in pick:
        r, c = res[k]
        cam = capi.camera(*synth.camera_for(r, c))
        shapes += [(r, c), (r, c)]
        cams += [cam, cam]
    batch = capi.Batch(ctx, shapes, cams, capi.stats_params())
    ap = capi.align_params(outer_iterations=iters)
    seeds = np.arange(2_000_000, 2_000_000 + n_pairs, dtype=np.uint64)
    T = np.stack([synth.motion(int(s), 8.0, 0.1) for s in seeds])
    I = np.eye(4, dtype=np.float32).flatten(order="F")
    poses = np.empty((2 * n_pairs, 16), np.float32)
    poses[0::2] = I
    poses[1::2] = T
    batch.render_synthetic(np.arange(2 * n_pairs, dtype=np.int32), np.repeat(seeds, 2),
                           np.tile(np.array([10, 11], np.uint64), n_pairs), poses)
    ref = np.arange(0, 2 * n_pairs, 2, dtype=np.int32)
    cur = ref + 1
    batch.convert()
    batch.set_autotune(True)
    batch.align(ref, cur, None, ap, use_graph=True, want_results=False)  # tuning call (eager)
    batch.align(ref, cur, None, ap, use_graph=True, want_results=False)  # graph capture
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.convert()
        Tout, _ = batch.align(ref, cur, None, ap, use_graph=True, want_results=True)
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / steps
    ok, _ = aligned(Tout, T)
    px = sum(r * c for r, c in shapes)
    tuned = batch.autotune_state()
    batch.close()
    return {"pairs": n_pairs, "iterations": iters, "ms_per_step": round(1000 * dt, 3),
            "projection_strategy": [t[0] for t in tuned],
            "pairs_per_s": round(n_pairs / dt, 1), "aligned_fraction": round(float(ok.mean()), 4),
            "depth_pixels_per_step": px,
            "pairs_per_resolution": {"%dx%d" % (c, r): int((pick == k).sum()) for k, (r, c) in enumerate(res)},
            "motion": "rot<=8deg, trans<=10cm", "graph": True}


def nn_batch(ctx, n_pairs, n_points, iters, steps):
    """kNN configuration: unstructured clouds of n_points (seeded plane/box scenes, rot <= 10 deg,
    trans <= 0.2 m, sigma 2 mm) aligned by AlignerNN from the identity, all pairs in one
    nicp_align_nn_batch call per step (clouds resident), next to the same pairs one call each."""
    from nicp_amd import capi, synth
    refs, curs, Tgt = [], [], []
    for s in range(n_pairs):
        d = synth.cloud_pair(3_000_000 + s, n=n_points)
        refs.append(ctx.cloud_from_host(d["src_points"], d["src_normals"], d["curvature"], d["omega_p"], d["omega_n"]))
        curs.append(ctx.cloud_from_host(d["tgt_points"], d["tgt_normals"], d["curvature"], d["omega_p"], d["omega_n"]))
        Tgt.append(d["T_gt"])
    Tgt = np.stack(Tgt)
    ap = capi.align_params(outer_iterations=iters)
    ctx.align_nn_batch(refs, curs, None, ap)  # warm-up: buffers grow to size
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        T, _ = ctx.align_nn_batch(refs, curs, None, ap)
    dt = (time.perf_counter() - t0) / steps
    ctx.profile(True)
    T1, _ = ctx.align_nn_batch(refs, curs, None, ap)
    prof = ctx.profile_read()
    ctx.profile(False)
    I = np.eye(4, dtype=np.float32).flatten(order="F")
    t0 = time.perf_counter()
    Ts = np.stack([ctx.align_nn(refs[i], curs[i], I, ap)[0] for i in range(n_pairs)])
    dt_single = time.perf_counter() - t0
    ok, err = aligned(T, Tgt)
    for c in refs + curs:
        c.close()
    q_ms, q_n = prof.get("k_nn_query", (0.0, 0))
    return {"pairs": n_pairs, "points_per_cloud": n_points, "iterations": iters, "ms_per_step": round(1000 * dt, 3),
            "pairs_per_s": round(n_pairs / dt, 2), "aligned_fraction": round(float(ok.mean()), 4),
            "rot_err_median_deg": round(float(np.median(err[:, 0])), 5),
            "single_pair_calls_ms": round(1000 * dt_single, 3), "vs_single_pair_calls": round(dt_single / dt, 2),
            "batch_equals_single": bool(np.array_equal(T.view(np.uint32), Ts.view(np.uint32))
                                        and np.array_equal(T1.view(np.uint32), T.view(np.uint32))),
            "k_nn_query_ms": round(q_ms / max(q_n, 1), 4),
            "data": "seeded plane/box scenes, rot<=10deg, trans<=0.2m, sigma=2mm"}


def calibrate_copy(mib):
    """A device-to-device copy of known size, for scaling rocprof's FETCH_SIZE / WRITE_SIZE."""
    from nicp_amd import capi
    ctx = capi.Context(0)
    n = mib * (1 << 20)
    a = capi.DeviceBuffer(ctx, (n,), np.uint8)
    b = capi.DeviceBuffer(ctx, (n,), np.uint8)
    a.upload(np.ones(n, np.uint8))
    t0 = time.perf_counter()
    for _ in range(5):
        capi.lib().nicp_device_copy(ctx.h, capi.C.c_void_p(b.ptr), capi.C.c_void_p(a.ptr), n, 2)
    dt = (time.perf_counter() - t0) / 5
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "copy_GB/s": round(2 * n / dt / 1e9, 1)}))


def main():
    ap_ = argparse.ArgumentParser()
    ap_.add_argument("--gpus", type=int, default=1)
    ap_.add_argument("--steps", type=int, default=32, help="timed steps (each a new batch of pairs)")
    ap_.add_argument("--warmup", type=int, default=1)
    ap_.add_argument("--pairs", type=int, default=128, help="pairs per step")
    ap_.add_argument("--iters", type=int, default=10)
    ap_.add_argument("--weights", default=None, help="network weights (.npy of bf16 bits)")
    ap_.add_argument("--no-cpu", action="store_true")
    ap_.add_argument("--no-profile", action="store_true")
    ap_.add_argument("--plain-pairs", type=int, default=32)
    ap_.add_argument("--uneven-pairs", type=int, default=512, help="pairs of the uneven-resolution batch (0: skip)")
    ap_.add_argument("--uneven-steps", type=int, default=3)
    ap_.add_argument("--nn-pairs", type=int, default=16, help="cloud pairs of the kNN batch (0: skip)")
    ap_.add_argument("--nn-points", type=int, default=200000)
    ap_.add_argument("--nn-steps", type=int, default=3)
    ap_.add_argument("--max-rot", type=float, default=MAX_ROT_DEG, help="motion range of the pairs (deg)")
    ap_.add_argument("--max-trans", type=float, default=MAX_TRANS_M, help="motion range of the pairs (m)")
    ap_.add_argument("--projection", default="auto", choices=["auto", "0", "1", "2"],
                     help="projection strategy of the batched refinement: autotuned, or forced (0: global "
                          "z-buffer, 1: global z-buffer read first, 2: tile-binned and fused)")
    ap_.add_argument("--streams", type=int, default=2,
                     help="batches in flight in the timed region: each on its own context (HIP stream), driven "
                          "by its own host thread, so one batch's conversion overlaps another's refinement")
    ap_.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                     help="only run a device-to-device copy of MIB MiB (to scale rocprof's FETCH_SIZE/WRITE_SIZE)")
    args = ap_.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (pairs shard "
              "independently across GPUs). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from nicp_amd import capi, initnet
    max_rot, max_trans = args.max_rot, args.max_trans
    weights_path = args.weights or initnet.DEFAULT_WEIGHTS
    weights = initnet.load_weights(weights_path)
    if weights is None:
        print("no trained weights at %s: train them with python -m nicp_amd.initnet" % weights_path,
              file=sys.stderr)
        sys.exit(1)
    record_path = os.path.splitext(weights_path)[0] + ".json"
    input_mode = json.load(open(record_path)).get("input_mode", 0) if os.path.exists(record_path) else 0

    B = args.pairs
    ap = capi.align_params(outer_iterations=args.iters)
    ref = np.arange(0, 2 * B, 2, dtype=np.int32)
    cur = ref + 1
    # lanes: a context (its own HIP stream), network and batch each; lane 0 also runs the eager
    # per-kernel pass and the byte-model counts
    lanes = []
    for _ in range(max(1, args.streams)):
        c = capi.Context(0)
        n_ = capi.InitNet(c, seed=1)
        n_.set_params(weights)
        n_.set_input_mode(input_mode)
        b_ = initnet.make_batch(c, B)
        if args.projection == "auto":
            b_.set_autotune(True)  # the projection strategy is picked during the warm-up
        else:
            b_.set_projection(int(args.projection))
        lanes.append((c, n_, b_))
    ctx, net, batch = lanes[0]

    def seeds_of(base, k):
        return np.arange(base + k * B, base + (k + 1) * B, dtype=np.uint64)

    # a step is the whole AlignerProjective::align of every pair, results included: the
    # statistics linearisation and Aligner::_computeStatistics run on the device, and every
    # pair's transform, H, b, error, inliers, Omega and validity are copied back
    def learned_step(lane, use_graph):
        _, n_, b_ = lanes[lane]
        b_.convert()
        T, res, _, used = b_.align_learned(n_, ref, cur, ap, use_graph=use_graph, want_results=True)
        return T, used

    def plain_step(lane, use_graph):
        _, _, b_ = lanes[lane]
        b_.convert()
        T, _ = b_.align(ref, cur, None, ap, use_graph=use_graph, want_results=True)
        return T

    # warm-up (separate seeds) and the set-up of the production path on every lane: the first
    # graph-mode call tunes the projection strategy (eagerly), the second captures the graphs
    for lane, (c, _, b_) in enumerate(lanes):
        for k in range(args.warmup):
            initnet.render_pairs(b_, seeds_of(WARMUP_SEED_BASE, k), max_rot, max_trans)
            learned_step(lane, True)
            plain_step(lane, True)
        initnet.render_pairs(b_, seeds_of(WARMUP_SEED_BASE, 0), max_rot, max_trans)
        learned_step(lane, True)
        learned_step(lane, True)
        c.synchronize()

    # inputs of the timed steps, resident in HBM before the timed region: every step's frames
    # rendered once (a new batch of pairs per step) into one packed device store
    n_pix = batch.num_pixels()
    store = capi.DeviceBuffer(ctx, (args.steps, n_pix), np.uint16)
    Tgts = []
    for k in range(args.steps):
        Tgts.append(initnet.render_pairs(batch, seeds_of(EVAL_SEED_BASE, k), max_rot, max_trans))
        batch.download_depth_device(store.row_ptr(k))
    ctx.synchronize()

    def run_lanes(step):
        """Every timed step once, step k on lane k % lanes, one host thread per lane; returns the
        wall time from the first upload to the last result and the transforms by step."""
        out = {}
        errors = []

        def work(lane):
            try:
                c, _, b_ = lanes[lane]
                for k in range(lane, args.steps, len(lanes)):
                    b_.upload_depth_device(store.row_ptr(k))
                    out[k] = step(lane, True)
                c.synchronize()
            except Exception as e:  # surfaced after the join
                errors.append(e)

        threads = [threading.Thread(target=work, args=(i,)) for i in range(len(lanes))]
        t0 = time.perf_counter()
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        dt = time.perf_counter() - t0
        if errors:
            raise errors[0]
        return dt, out

    # timed region: the steps' conversions, learned starts and refinements (replayed from their
    # hipGraphs, no per-launch events), lanes overlapping
    t_total, out = run_lanes(learned_step)
    n_ok = 0
    used_total = 0
    Ts = []
    for k in range(args.steps):
        T, used = out[k]
        ok, _ = aligned(T, Tgts[k])
        n_ok += int(ok.sum())
        used_total += int(used.sum())
        Ts.append(T)
    tuned_main = batch.autotune_state()

    # per-kernel times: the same step run eagerly with HIP events around every launch, over the
    # first batches of the same seeds (outside the timed region)
    psteps = min(4, args.steps)
    if not args.no_profile:
        ctx.profile(True)
    t_eager = 0.0
    eager_same = True
    for k in range(psteps):
        batch.upload_depth_device(store.row_ptr(k))
        ctx.synchronize()
        t0 = time.perf_counter()
        T, _ = learned_step(0, False)
        ctx.synchronize()
        t_eager += time.perf_counter() - t0
        eager_same = eager_same and np.array_equal(T, Ts[k], equal_nan=True)
    prof, prof_error = {}, None
    if not args.no_profile:
        try:
            prof = ctx.profile_read()
        except capi.NicpError as e:
            prof_error = str(e)
            log("kernel timing unavailable: %s" % e)
    ctx.profile(False)
    n_pairs = B * args.steps
    value = n_ok / t_total
    ms_step = 1000.0 * t_total / args.steps
    _, err_learned = aligned(np.concatenate(Ts), np.concatenate(Tgts))
    log("batched+learned: %.2f ms/step, %d/%d aligned, %.1f pairs/s, aligned %.1f pairs/s, network start kept %d" % (
        ms_step, n_ok, n_pairs, n_pairs / t_total, value, used_total))

    # the same pairs without the network (batched, identity start, graph replay); bit-identical
    # to the single-pair path, so its accuracy is the plain single-pair accuracy
    t_plain_b, out_plain = run_lanes(plain_step)
    n_ok_plain = sum(int(aligned(out_plain[k], Tgts[k])[0].sum()) for k in range(args.steps))
    frac_plain = n_ok_plain / n_pairs
    log("batched from identity: %.2f ms/step, %d/%d aligned" % (1000 * t_plain_b / args.steps, n_ok_plain, n_pairs))

    # per-frame / per-pair counts of the last timed batch for the byte model
    frames_info, pairs_info = [], []
    batch.upload_depth_device(store.row_ptr(args.steps - 1))
    batch.convert()
    _, res, _, _ = batch.align_learned(net, ref, cur, ap, use_graph=False, want_results=True)
    for f in range(2 * B):
        c = batch.frame_cloud(f, (ROWS, COLS))
        frames_info.append((ROWS * COLS, c.n))
        c.close()
    # per pair: projector pixels, reference points, both-valid pixels, reference frame pixels
    pairs_info = [(ROWS * COLS, frames_info[2 * i][1], res[i].overlap_pixels, ROWS * COLS) for i in range(B)]
    # the counter summaries hold the default workload only
    default = (args.pairs, args.iters, args.max_rot, args.max_trans, args.projection) == (
        128, 10, MAX_ROT_DEG, MAX_TRANS_M, "auto")
    kernels = {}
    total_ms = sum(v[0] for v in prof.values()) if prof else 0.0
    for name, (ms, cnt) in prof.items():
        if cnt == 0:
            continue
        avg = ms / cnt
        nb = kernel_bytes(name, pairs_info, frames_info)
        kernels[name] = {"avg_ms": round(avg, 4), "launches": cnt,
                         "GB/s": round(nb / (avg * 1e-3) / 1e9, 1) if nb else None,
                         "share": round(ms / max(total_ms, 1e-9), 3)}
        tr = measured_traffic(name) if default else None
        if tr:
            kernels[name]["traffic_GB"] = round(tr / 1e9, 3)
            kernels[name]["traffic_over_algorithmic"] = round(tr / nb, 3) if nb else None
    hbm = [k for k in kernels if kernel_bytes(k, pairs_info, frames_info)]
    dom = max(hbm, key=lambda k: prof[k][0]) if hbm else None
    roofline = None
    if dom:
        nb = kernel_bytes(dom, pairs_info, frames_info)
        avg = kernels[dom]["avg_ms"]
        ach = nb / (avg * 1e-3) / 1e9
        roofline = {"kernel": dom, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic(dom) if default else None, "bytes_per_launch": nb,
                    "avg_launch_ms": avg}
        own = kernel_bytes(dom + "(reads)", pairs_info, frames_info)
        if own:
            # the same launch against the bytes the kernel itself has to read
            roofline["bytes_model"] = "B_iter = 24 P + 12 N + 104 C2 per pair (one outer iteration)"
            roofline["kernel_reads_per_launch"] = own
            roofline["kernel_reads_frac"] = round(own / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    net_roofline = None
    if "initnet_forward" in kernels:
        avg = kernels["initnet_forward"]["avg_ms"]
        tf = NET_FLOPS_PER_PAIR * B / (avg * 1e-3) / 1e12
        net_roofline = {"stage": "initnet_forward", "bound": "mfma", "achieved": round(tf, 2),
                        "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(tf / BF16_PEAK_TFLOPS, 5),
                        "flops_per_launch": NET_FLOPS_PER_PAIR * B, "avg_launch_ms": avg}

    # plain single-pair GPU NICP: one pair per call, no graph, identity start, inputs resident
    plain = None
    if args.plain_pairs > 0:
        pb = initnet.make_batch(ctx, 1)
        one = np.array([0], np.int32), np.array([1], np.int32)
        times = []
        for i in range(args.plain_pairs + 1):
            initnet.render_pairs(pb, np.array([EVAL_SEED_BASE + i], np.uint64), max_rot, max_trans)
            ctx.synchronize()
            s0 = time.perf_counter()
            pb.convert()
            pb.align(one[0], one[1], None, ap, use_graph=False, want_results=True)
            ctx.synchronize()
            if i > 0:
                times.append(time.perf_counter() - s0)
        plain = 1.0 / float(np.median(times))
        pb.close()
        log("plain single-pair: %.1f pairs/s, aligned fraction %.4f -> %.1f aligned pairs/s" % (
            plain, frac_plain, plain * frac_plain))
    plain_aligned = plain * frac_plain if plain else None

    # batched configuration with uneven frames: pairs of 160x120 / 320x240 / 640x480 /
    # 1280x960 (uniform, seeded), motions <= 8 deg / 10 cm, 15 iterations replayed from a graph
    uneven = None
    if args.uneven_pairs > 0:
        uneven = uneven_batch(ctx, args.uneven_pairs, 15, args.uneven_steps)
        log("uneven batch: %s" % uneven)

    # kNN configuration: 16 pairs of 200k-point unstructured clouds, 20 AlignerNN iterations
    nn = None
    if args.nn_pairs > 0:
        nn = nn_batch(ctx, args.nn_pairs, args.nn_points, 20, args.nn_steps)
        log("kNN batch: %s" % nn)

    cpu = parity = None
    if not args.no_cpu:
        parity = parity_sample(ctx, np.arange(EVAL_SEED_BASE, EVAL_SEED_BASE + 2), max_rot, max_trans)
        log("parity sample: %s" % parity)
        cpu = cpu_baseline()
        log("cpu baseline: %s" % cpu)

    out = {
        "metric": "aligned pairs/s @640x480 (rot err<0.1deg): batched+learned init vs plain, 1 GPU",
        "value": round(value, 2),
        "unit": "aligned pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "%d held-out pairs of 640x480 uint16 depth per step, a new batch each step "
                               "(seeded 20-plane/box scenes, rot<=%gdeg, trans<=%gm, sigma=1.5e-3 z^2, 10%% holes); "
                               "both conversions + learned initial guess (bf16 CNN, chi2 fallback to identity) + "
                               "%d-iteration projective NICP per pair" % (B, max_rot, max_trans, args.iters),
                   "pairs_per_step": B, "eval_pairs": n_pairs, "resolution": "640x480", "iterations": args.iters,
                   "init": "learned", "weights": os.path.relpath(weights_path, ROOT), "streams": len(lanes)},
        "aligned_fraction": round(n_ok / n_pairs, 4),
        "pairs_per_s": round(n_pairs / t_total, 2),
        "network_start_kept": round(used_total / n_pairs, 4),
        "rot_err_median_deg": round(float(np.median(err_learned[:, 0])), 4),
        "diverged_pairs": int(np.sum(~np.isfinite(err_learned[:, 0]))),
        "timing": "wall time of all timed steps, %d batches in flight (one context/stream and host thread "
                  "each; step k on lane k %% %d), inputs resident in HBM; conversion and learned step replayed "
                  "from hipGraphs; per-kernel times from a separate single-stream eager pass with HIP events "
                  "around every launch" % (len(lanes), len(lanes)),
        "eager_ms_per_step": round(1000.0 * t_eager / psteps, 3),
        "eager_equals_graph": bool(eager_same),
        "plain_single_pair_pairs_per_s": round(plain, 2) if plain else None,
        "plain_aligned_fraction": round(frac_plain, 4),
        "plain_single_pair_aligned_pairs_per_s": round(plain_aligned, 2) if plain_aligned else None,
        "vs_plain": round(value / plain_aligned, 2) if plain_aligned else None,
        "batched_identity_ms_per_step": round(1000.0 * t_plain_b / args.steps, 3),
        "batched_identity_aligned_pairs_per_s": round(n_ok_plain / t_plain_b, 2),
        "uneven_batch": uneven,
        "nn_batch": nn,
        "roofline": roofline,
        "net_roofline": net_roofline,
        "kernels": kernels,
        "autotune": [{"projection_strategy": c, "ms_strategy0": round(m0, 4), "ms_strategy1": round(m1, 4),
                      "ms_strategy2": round(m2, 4)} for c, m0, m1, m2 in tuned_main],
        "kernel_timing_error": prof_error,
        "parity_sample": parity,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    store.close()
    for c, _, _ in lanes:
        c.close()


if __name__ == "__main__":
    main()
