"""Training and weights of the learned initial guess (the C ABI's nicp_initnet_*).

The training set is generated on the device: seeded synthetic pairs are rendered by
nicp_batch_render_synthetic into a batch of 640x480 frames (the second camera moved
by a random rigid motion of up to 30 deg / 0.5 m), converted, and the ground-truth
motion is refined by plain NICP started at it (the label; pairs whose refinement
leaves the ground truth by more than 1 deg / 5 cm keep the ground truth). The
network input of every pair (nicp_batch_net_input) and its label stay resident in
device memory, and training runs the network's own forward/backward kernels with
Adam (nicp_initnet_train_step) over mini-batches drawn from them.

Weights are stored as bf16 bit patterns (the precision the matrix cores consume)
in an .npy file, with the training record beside it in JSON.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import capi, se3, synth

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "initnet_640x480.npy")
ROWS, COLS = 480, 640
LOSS_WEIGHTS = (1.0, 1.0, 1.0, 6.0, 6.0, 6.0)


def to_bf16_bits(p):
    u = np.ascontiguousarray(p, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    return r.astype(np.uint16)


def from_bf16_bits(b):
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


def save_weights(path, params, record):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, to_bf16_bits(params))
    with open(os.path.splitext(path)[0] + ".json", "w") as f:
        json.dump(record, f, indent=1, sort_keys=True)


def load_weights(path=DEFAULT_WEIGHTS):
    """The stored parameters as float32, or None when no trained weights exist."""
    if not os.path.exists(path):
        return None
    return from_bf16_bits(np.load(path, allow_pickle=False))


def make_batch(ctx, num_pairs, sp=None):
    """A batch of 2 num_pairs 640x480 frames: frame 2i is the reference of pair i, 2i + 1 its current."""
    fx, fy, cx, cy = synth.camera_for(ROWS, COLS)
    cam = capi.camera(fx, fy, cx, cy)
    sp = sp or capi.stats_params()
    return capi.Batch(ctx, [(ROWS, COLS)] * (2 * num_pairs), [cam] * (2 * num_pairs), sp)


def render_pairs(batch, seeds, max_rot_deg, max_trans_m):
    """Renders the pairs of `seeds` into the batch (as synth.pair does on the host); returns T_gt (n, 16)."""
    n = len(seeds)
    T = np.stack([synth.motion(int(s), max_rot_deg, max_trans_m) for s in seeds])
    I = np.eye(4, dtype=np.float32).flatten(order="F")
    frames = np.arange(2 * n, dtype=np.int32)
    scene = np.repeat(np.asarray(seeds, np.uint64), 2)
    streams = np.tile(np.array([10, 11], np.uint64), n)
    poses = np.empty((2 * n, 16), np.float32)
    poses[0::2] = I
    poses[1::2] = T
    batch.render_synthetic(frames, scene, streams, poses)
    return T


def symmetry_targets(x6, flags):
    """Pose 6-vectors of pairs transformed by the symmetry flags of nicp_initnet_train_step_rows:
    bit 0 swaps the frames (T -> T^-1), bit 1 mirrors the columns (T -> F T F, F = diag(-1, 1, 1)),
    bit 2 mirrors the rows (F = diag(1, -1, 1)). Mirroring an image about its centre column is
    mirroring the scene about the camera's x = 0 plane when the principal point is the image
    centre, as for every camera of the synthetic scenes."""
    x = np.array(x6, np.float64).reshape(-1, 6)
    f = np.asarray(flags).reshape(-1)
    sw = (f & 1) != 0
    if np.any(sw):
        t, q = x[sw, :3], x[sw, 3:]
        w = np.sqrt(np.maximum(0.0, 1.0 - np.sum(q * q, axis=1)))
        # R^T t for the unit quaternion (w, q): t - 2 w (q x t) + 2 q x (q x t)
        c = np.cross(q, t)
        x[sw, :3] = -(t - 2.0 * w[:, None] * c + 2.0 * np.cross(q, c))
        x[sw, 3:] = -q
    h = (f & 2) != 0
    x[h] *= np.array([-1.0, 1.0, 1.0, 1.0, -1.0, -1.0])
    v = (f & 4) != 0
    x[v] *= np.array([1.0, -1.0, 1.0, -1.0, 1.0, -1.0])
    return x.astype(np.float32)


def generate(ctx, net, num_pairs, seed_base, chunk=256, max_rot_deg=30.0, max_trans_m=0.5, log=print):
    """Device-resident inputs (DeviceBuffer [n, 60, 80, 2] float32) and host labels (n, 6)."""
    X = capi.DeviceBuffer(ctx, (num_pairs, 60, 80, 2))
    Y = np.empty((num_pairs, 6), np.float32)
    batch = make_batch(ctx, chunk)
    ap = capi.align_params()
    kept_gt = 0
    t0 = time.time()
    for c0 in range(0, num_pairs, chunk):
        n = min(chunk, num_pairs - c0)
        seeds = np.arange(seed_base + c0, seed_base + c0 + n, dtype=np.uint64)
        Tgt = render_pairs(batch, seeds, max_rot_deg, max_trans_m)
        batch.convert()
        ref = np.arange(0, 2 * n, 2, dtype=np.int32)
        cur = ref + 1
        Tref, _ = batch.align(ref, cur, Tgt, ap, use_graph=False, want_results=True)
        labels = np.empty((n, 16), np.float32)
        for i in range(n):
            rot, tr = synth.pose_error(Tref[i], Tgt[i])
            ok = rot < 1.0 and tr < 0.05 and np.all(np.isfinite(Tref[i]))
            labels[i] = Tref[i] if ok else Tgt[i]
            kept_gt += 0 if ok else 1
        batch.net_input(net, ref, cur, X.row_ptr(c0))
        Y[c0:c0 + n] = se3.t2v(labels).astype(np.float32)
        if (c0 // chunk) % 50 == 0:
            log("  generated %d/%d pairs (%.1f s)" % (c0 + n, num_pairs, time.time() - t0))
    batch.close()
    return X, Y, {"pairs": num_pairs, "labels_kept_ground_truth": kept_gt, "seconds": time.time() - t0}


def predict(ctx, net, X, batch=512):
    """Network outputs (n, 6) for a DeviceBuffer of inputs."""
    n = X.shape[0]
    out = capi.DeviceBuffer(ctx, (batch, 6))
    pred = np.empty((n, 6), np.float32)
    for b0 in range(0, n, batch):
        m = min(batch, n - b0)
        net.forward_device(X.row_ptr(b0), m, out.ptr)
        pred[b0:b0 + m] = out.download(rows=m)
    out.close()
    return pred


def evaluate(ctx, net, X, Y):
    """Rotation / translation errors of the predictions and the fraction within 5 deg / 10 cm."""
    pred = predict(ctx, net, X).astype(np.float64)
    lab = np.asarray(Y, np.float64)
    qn = np.sum(pred[:, 3:] ** 2, axis=1)
    big = qn > 0.99
    pred[big, 3:] *= np.sqrt(0.99 / qn[big])[:, None]
    P, L = se3.v2t(pred), se3.v2t(lab)
    err = np.array([synth.pose_error(P[i], L[i]) for i in range(len(P))])
    rots, trs = err[:, 0], err[:, 1]
    return {"rot_err_mean_deg": float(rots.mean()), "rot_err_median_deg": float(np.median(rots)),
            "trans_err_mean_m": float(trs.mean()), "trans_err_median_m": float(np.median(trs)),
            "within_5deg_10cm": float(np.mean((rots < 5.0) & (trs < 0.1)))}


def train(ctx, net, X, Y, epochs, batch=256, lr=1e-3, lr_min=1e-5, seed=0, augment=True, log=print,
          on_epoch=None):
    """Adam with a cosine learning rate over `epochs` passes of random minibatches; with
    augment, every sample gets random symmetry flags (8 exact variants of each pair)."""
    n = X.shape[0]
    per_epoch = n // batch
    steps = epochs * per_epoch
    rng = np.random.default_rng(seed)
    step = 0
    t0 = time.time()
    losses = []
    for ep in range(epochs):
        perm = rng.permutation(n).astype(np.int32)
        ep_loss = 0.0
        for k in range(per_epoch):
            rows = perm[k * batch:(k + 1) * batch]
            flags = rng.integers(0, 8, batch).astype(np.int32) if augment else np.zeros(batch, np.int32)
            cur_lr = lr_min + 0.5 * (lr - lr_min) * (1.0 + np.cos(np.pi * step / max(1, steps)))
            ep_loss += net.train_step_rows(X, rows, symmetry_targets(Y[rows], flags), float(cur_lr), flags)
            step += 1
        losses.append(ep_loss / max(1, per_epoch))
        log("  epoch %d/%d loss %.5f (%.1f s)" % (ep + 1, epochs, losses[-1], time.time() - t0))
        if on_epoch:
            on_epoch(ep + 1)
    return {"epochs": epochs, "batch": batch, "lr": lr, "lr_min": lr_min, "steps": step, "losses": losses,
            "augment": bool(augment), "seconds": time.time() - t0}


def main(argv=None):
    ap = argparse.ArgumentParser(description="train the learned initial guess on device-rendered pairs")
    ap.add_argument("--pairs", type=int, default=200000)
    ap.add_argument("--val-pairs", type=int, default=4096)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed-base", type=int, default=10_000_000)
    ap.add_argument("--init-seed", type=int, default=1)
    ap.add_argument("--input-mode", type=int, default=1, help="0: depth, 1: inverse depth")
    ap.add_argument("--no-augment", action="store_true")
    ap.add_argument("--lr-min", type=float, default=1e-5)
    ap.add_argument("--init-weights", default=None,
                    help="start from these weights (.npy of bf16 bits, with their .json record) instead of the seed")
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    a = ap.parse_args(argv)

    def log(*m):
        print(*m, flush=True)

    ctx = capi.Context(0)
    net = capi.InitNet(ctx, seed=a.init_seed)
    net.set_loss_weights(LOSS_WEIGHTS)
    net.set_input_mode(a.input_mode)
    previous = None
    if a.init_weights:
        w = load_weights(a.init_weights)
        if w is None:
            raise SystemExit("no weights at %s" % a.init_weights)
        net.set_params(w)
        rec = os.path.splitext(a.init_weights)[0] + ".json"
        if os.path.exists(rec):
            with open(rec) as f:
                previous = json.load(f)
        if previous and previous.get("input_mode", a.input_mode) != a.input_mode:
            raise SystemExit("the initial weights were trained with input mode %d" % previous["input_mode"])
    log("generating %d training pairs" % a.pairs)
    X, Y, gen = generate(ctx, net, a.pairs, a.seed_base, log=log)
    log("generating %d validation pairs" % a.val_pairs)
    Xv, Yv, _ = generate(ctx, net, a.val_pairs, a.seed_base + a.pairs + 1_000_000, log=log)
    log("before training: %s" % evaluate(ctx, net, Xv, Yv))
    def on_epoch(ep):
        if ep % 10 == 0 and ep < a.epochs:
            log("  validation after %d epochs: %s" % (ep, evaluate(ctx, net, Xv, Yv)))

    tr = train(ctx, net, X, Y, a.epochs, a.batch, a.lr, a.lr_min, augment=not a.no_augment, log=log,
               on_epoch=on_epoch)
    val = evaluate(ctx, net, Xv, Yv)
    log("validation: %s" % val)
    record = {"generation": gen, "training": tr, "validation": val, "train_seeds": [a.seed_base, a.seed_base + a.pairs],
              "loss_weights": LOSS_WEIGHTS, "motion": {"max_rot_deg": 30.0, "max_trans_m": 0.5},
              "resolution": [COLS, ROWS], "init_seed": a.init_seed, "input_mode": a.input_mode}
    if previous is not None:
        # the run this one started from (its own predecessors stay in its record's "previous")
        record["previous"] = previous
    save_weights(a.out, net.get_params(), record)
    log("saved %s" % a.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
