"""Device parity of the learned initial guess: the CNN's forward and backward
kernels (bf16 matrix-core GEMMs with fp32 accumulation) layer by layer against
the CPU oracle in its bf16-operand mode, the network input against the oracle's
DepthImage_scale, the device scene renderer against the host one, and the
fallback policy of the learned-start aligner.

Bars: each GEMM result within 2e-5 (forward) / 5e-5 (gradients) of the sum of
the absolute products of its bf16 operands -- the fp32 accumulation error of
sums of up to 1281 (forward) or a few hundred per split (gradients) terms; the
loss gradient and the network input bit-exact; renders identical.
"""
import sys

import numpy as np
import pytest

import oracle as O
from nicp_amd import capi, synth
from oracle import initnet as net
from tests.helpers import aparams, cams, params

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _params(seed):
    rng = np.random.default_rng(seed)
    p = np.empty(net.num_params(), np.float32)
    off = net.offsets()
    for l in range(len(net.LAYERS)):
        p[off[l]:off[l + 1]] = rng.normal(0.0, np.sqrt(2.0 / net.layer_k(l)), off[l + 1] - off[l])
    return p


def _input(seed, B):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.5, 4.5, (B,) + net.IN_SHAPE).astype(np.float32)
    x[rng.uniform(size=x.shape) < 0.1] = 0.0
    return x


def _act_shape(l, B):
    if l == 0:
        return (B,) + net.IN_SHAPE
    ih, iw, ic, oh, ow, oc, conv, relu = net.LAYERS[l - 1]
    return (B, oh, ow, oc) if conv else (B, oc)


def _read_acts(d, B, what, first=0):
    """Layers first ... L of an activation (what = 0) or gradient (what = 1) buffer; layers below
    first are None (the backward pass writes no gradient for the network's input)."""
    return [None] * first + [
        d.read(what, l, int(np.prod(_act_shape(l, B)))).reshape(_act_shape(l, B)).astype(np.float64)
        for l in range(first, len(net.LAYERS) + 1)]


def test_forward_layers(ctx):
    say("case initnet forward layers B=4")
    B = 4
    d = capi.InitNet(ctx, seed=1)
    p = _params(1)
    d.set_params(p)
    x = _input(2, B)
    y = d.forward(x)
    acts = _read_acts(d, B, 0)
    assert np.array_equal(acts[0], x.astype(np.float64))
    assert np.array_equal(acts[-1], y.astype(np.float64))
    for l in range(len(net.LAYERS)):
        ref = net.layer_forward(p, acts[l], l, "bf16").reshape(acts[l + 1].shape)
        A = np.abs(net.bf16(net.layer_input_matrix(acts[l], l)))
        W = np.abs(net.bf16(net.layer_matrix(p, l)))
        bound = (A @ W).reshape(acts[l + 1].shape)
        err = np.abs(acts[l + 1] - ref)
        say("  layer %d max err %.3g, max err/bound %.3g" % (l, err.max(), (err / np.maximum(bound, 1e-30)).max()))
        assert np.all(err <= 2e-5 * bound + 1e-30), l


def test_forward_large_batch(ctx):
    """A forward pass of 4096 inputs: the first layer's GEMM has 76800 row tiles, more than a grid's
    y dimension holds (65535), so the row tiles run on x. Then forwards of 256 inputs on the same
    network, whose buffers were reserved for 4096: with fewer rows the fully connected layers split
    their reductions further and need more chunk room than the 4096-row call (the reservation
    covers every smaller batch). Every output row agrees with the same input's row of the
    256-input forwards to bf16 precision: the two batch sizes split the fully connected layers'
    reductions differently, and a different FP32 order can flip the bf16 rounding of an
    activation that the next layer reads (2^-8 relative). Rows beyond the first 3495 inputs (row
    tiles past 65535) would be garbage or missing if the grid were wrong."""
    B = 4096
    say("case initnet forward large batch B=%d" % B)
    d = capi.InitNet(ctx, seed=1)
    d.set_params(_params(1))
    x = _input(3, B)
    y = d.forward(x)
    ys = np.concatenate([d.forward(x[i:i + 256]) for i in range(0, B, 256)])
    assert np.all(np.isfinite(y))
    err = np.abs(y.astype(np.float64) - ys.astype(np.float64))
    scale = max(1.0, float(np.abs(ys).max()))
    say("  max |y - y_256| = %.3g, mean %.3g (max |y| %.3g)" % (err.max(), err.mean(), np.abs(ys).max()))
    assert err.max() <= 2e-2 * scale
    assert err.mean() <= 1e-3 * scale


def test_backward_layers(ctx):
    say("case initnet backward layers B=3")
    B = 3
    d = capi.InitNet(ctx, seed=2)
    p = _params(3)
    d.set_params(p)
    w = np.float32([1, 1, 1, 4, 4, 4])
    d.set_loss_weights(w)
    x = _input(4, B)
    t = np.random.default_rng(5).normal(size=(B, 6)).astype(np.float32)
    d.train_step(x, t, lr=0.0)
    acts = _read_acts(d, B, 0)
    dacts = _read_acts(d, B, 1, first=1)
    grads = d.read(2, 0, net.num_params()).astype(np.float64)
    # loss gradient: 2 w (y - t) / B in fp32, bit-exact
    y32 = acts[-1].astype(np.float32)
    g6 = (np.float32(2.0) * w) * (y32 - t) / np.float32(B)
    assert np.array_equal(dacts[-1].astype(np.float32), g6)
    off = net.offsets()
    for l in reversed(range(len(net.LAYERS))):
        ih, iw, ic, oh, ow, oc, conv, relu = net.LAYERS[l]
        G = dacts[l + 1].reshape(-1, oc)  # masked in place by the layer's ReLU
        dW, dX = net.layer_backward(p, acts[l], G, l, "bf16")
        A = np.abs(net.bf16(net.layer_input_matrix(acts[l], l)))
        Gb = np.abs(net.bf16(G))
        bW = A.T @ Gb
        gW = grads[off[l]:off[l + 1]].reshape(dW.shape)
        eW = np.abs(gW - dW)
        say("  layer %d dW max err %.3g, err/bound %.3g" % (l, eW.max(), (eW / np.maximum(bW, 1e-30)).max()))
        assert np.all(eW <= 5e-5 * bW + 1e-30), l
        if l > 0:
            Wb = np.abs(net.bf16(net.layer_matrix(p, l)[:-1]))
            bcol = Gb @ Wb.T
            bX = net.col2im(bcol, l, B) if conv else bcol.reshape(B, ic)
            mask = acts[l] > 0.0  # the ReLU of the layer below masks dX in place
            ref = np.where(mask, dX.reshape(acts[l].shape), 0.0)
            eX = np.abs(dacts[l] - ref)
            say("  layer %d dX max err %.3g, err/bound %.3g" % (l, eX.max(),
                                                                (eX / np.maximum(bX.reshape(eX.shape), 1e-30)).max()))
            assert np.all(eX <= 5e-5 * bX.reshape(eX.shape) + 1e-30), l


def test_full_chain(ctx):
    say("case initnet full chain B=5")
    B = 5
    d = capi.InitNet(ctx, seed=3)
    p = _params(6)
    d.set_params(p)
    x = _input(7, B)
    t = np.random.default_rng(8).normal(size=(B, 6)).astype(np.float32) * 0.3
    loss = d.train_step(x, t, lr=0.0)
    lo, go = net.loss_and_grad(p, x, t, mode="bf16")
    le, ge = net.loss_and_grad(p, x, t, mode="exact")
    g = d.read(2, 0, net.num_params()).astype(np.float64)
    rel_b = np.linalg.norm(g - go) / np.linalg.norm(go)
    rel_e = np.linalg.norm(g - ge) / np.linalg.norm(ge)
    say("  loss %.6g oracle(bf16) %.6g exact %.6g; grad rel err vs bf16 oracle %.3g, vs exact %.3g" %
        (loss, lo, le, rel_b, rel_e))
    assert abs(loss - lo) <= 1e-3 * abs(lo)
    assert rel_b < 1e-2
    assert rel_e < 5e-2


def test_training_reduces_loss(ctx):
    say("case initnet training B=32")
    d = capi.InitNet(ctx, seed=4)
    X = _input(10, 32)
    # a learnable target: differences of the two channels' mean depth
    m = X.reshape(32, -1, 2).mean(axis=1)
    T = np.stack([m[:, 0] - m[:, 1], 0.5 * (m[:, 0] - 2.5), 0.5 * (m[:, 1] - 2.5)] * 2, axis=1).astype(np.float32)
    first = d.train_step(X, T, lr=0.0)
    for _ in range(150):
        last = d.train_step(X, T, lr=1e-3)
    final = d.train_step(X, T, lr=0.0)
    say("  loss %.4g -> %.4g" % (first, final))
    assert final < 0.2 * first
    del last


def _host_symmetry(x, f):
    """The input transform of the symmetry flags, on the host."""
    y = x[..., ::-1] if f & 1 else x
    y = y[:, ::-1, :] if f & 2 else y
    y = y[::-1, :, :] if f & 4 else y
    return np.ascontiguousarray(y)


def test_train_step_rows_matches_host_batch(ctx):
    say("case initnet train_step_rows with symmetry flags")
    X = _input(11, 16)
    T = np.random.default_rng(12).normal(size=(8, 6)).astype(np.float32) * 0.1
    rows = np.array([3, 7, 7, 0, 15, 9, 2, 11], np.int32)
    flags = np.arange(8, dtype=np.int32)
    Xh = np.stack([_host_symmetry(X[r], f) for r, f in zip(rows, flags)])
    for mode in (0, 1):
        a, b = capi.InitNet(ctx, seed=7), capi.InitNet(ctx, seed=7)
        a.set_input_mode(mode)
        b.set_input_mode(mode)
        Xd = capi.DeviceBuffer(ctx, X.shape)
        Xd.upload(X)
        la = a.train_step(Xh, T, lr=1e-3)
        lb = b.train_step_rows(Xd, rows, T, 1e-3, flags)
        say("  input mode %d: loss %.6g / %.6g" % (mode, la, lb))
        # same minibatch, same forward pass; the atomics of the backward may round differently
        assert la == lb
        assert np.max(np.abs(a.get_params() - b.get_params())) < 1e-6
        np.testing.assert_array_equal(b.read(0, 0, Xh.size).reshape(Xh.shape), Xh)


def test_inverse_depth_first_layer(ctx):
    say("case initnet inverse-depth input B=3")
    B = 3
    d = capi.InitNet(ctx, seed=8)
    p = _params(13)
    d.set_params(p)
    d.set_input_mode(1)
    x = _input(14, B)
    d.forward(x)
    acts = _read_acts(d, B, 0)
    ref = net.layer_forward(p, acts[0], 0, "bf16", input_mode=1)
    A = np.abs(net.bf16(net.layer_input_matrix(acts[0], 0, 1, "bf16")))
    bound = (A @ np.abs(net.bf16(net.layer_matrix(p, 0)))).reshape(acts[1].shape)
    err = np.abs(acts[1] - ref)
    say("  layer 0 max err/bound %.3g" % (err / np.maximum(bound, 1e-30)).max())
    assert np.all(err <= 2e-5 * bound + 1e-30)


def test_net_input_matches_depth_scale(ctx):
    say("case initnet input 640x480 and 320x240")
    shapes = [(480, 640), (480, 640), (240, 320), (240, 320)]
    cm = [cams(r, c)[1] for r, c in shapes]
    _, sp = params()
    b = capi.Batch(ctx, shapes, cm, sp)
    d = capi.InitNet(ctx, seed=5)
    imgs = []
    for i, (r, c) in enumerate(shapes):
        dr, dc, _ = synth.pair(40 + i // 2, r, c)
        img = dr if i % 2 == 0 else dc
        imgs.append(img)
        b.upload(i, img)
    out = capi.DeviceBuffer(ctx, (2, 60, 80, 2))
    b.net_input(d, [0, 2], [1, 3], out.ptr)
    got = out.download()
    for p, (ri, ci) in enumerate([(0, 1), (2, 3)]):
        for ch, fi in enumerate((ri, ci)):
            step = shapes[fi][1] // 80
            dm = np.where(imgs[fi] > 0, np.float32(0.001) * imgs[fi].astype(np.float32), 0).astype(np.float32)
            ref = O.depth_scale(dm, step)
            np.testing.assert_array_equal(got[p, :, :, ch].view(np.uint32), ref.view(np.uint32))
    bad = capi.Batch(ctx, [(100, 100)], [cams(100, 100)[1]], sp)
    with pytest.raises(capi.NicpError):
        bad.net_input(d, [0], [0], out.ptr)


@pytest.mark.parametrize("res", [(120, 160), (480, 640)])
def test_render_matches_host(ctx, res):
    rows, cols = res
    say("case synthetic render %dx%d" % (rows, cols))
    _, sp = params()
    cm = cams(rows, cols)[1]
    b = capi.Batch(ctx, [res] * 2, [cm] * 2, sp)
    T = synth.motion(77, 5.0, 0.05)
    I = np.eye(4, dtype=np.float32).flatten(order="F")
    b.render_synthetic([0, 1], [77, 77], [10, 11], np.stack([I, T]))
    ref0, ref1, Tp = synth.pair(77, rows, cols)
    assert np.array_equal(Tp, T)
    # compared through the converted index images and points, which are exact functions of
    # the depth (the batch keeps no host copy of it)
    b.convert()
    ocam, _ = cams(rows, cols)
    osp, _ = params()
    for f, ref in ((0, ref0), (1, ref1)):
        cl = b.frame_cloud(f, res)
        idx = cl.index_image()
        fo = O.Frame(ref, 0.001, ocam, osp)
        say("  frame %d: points %d host %d" % (f, int((idx >= 0).sum()), fo.n))
        np.testing.assert_array_equal(idx, fo.index_image)
        np.testing.assert_array_equal(cl.download()[0].view(np.uint32), fo.points.view(np.uint32))


def _constant_net(ctx, x6):
    d = capi.InitNet(ctx, seed=6)
    p = np.zeros(net.num_params(), np.float32)
    p[net.offsets()[-1] - 6:] = x6  # the output bias
    d.set_params(p)
    return d


def test_align_learned_policy(ctx):
    say("case align_learned policy")
    rows, cols = 480, 640
    _, sp = params()
    cm = cams(rows, cols)[1]
    _, ap = aparams()
    # pair 0: large motion, pair 1: small motion
    dr0, dc0, T0 = synth.pair(501, rows, cols, max_rot_deg=20.0, max_trans_m=0.3)
    dr1, dc1, T1 = synth.pair(502, rows, cols, max_rot_deg=2.0, max_trans_m=0.02)
    b = capi.Batch(ctx, [(rows, cols)] * 4, [cm] * 4, sp)
    for i, img in enumerate((dr0, dc0, dr1, dc1)):
        b.upload(i, img)
    b.convert()
    x_good = O.t2v(T0).astype(np.float32)
    # a network that always predicts the large motion: kept for pair 0, rejected for pair 1
    d = _constant_net(ctx, x_good)
    T, res, starts, used = b.align_learned(d, [0, 2], [1, 3], ap, use_graph=False)
    # the probe: one linearisation from each start
    probe = capi.align_params(outer_iterations=1)
    I = np.eye(4, dtype=np.float32).flatten(order="F")
    # the prediction is the same constant for both pairs; pair 0 must have kept it
    assert used[0] == 1
    _, pr = b.align([0, 2, 0, 2], [1, 3, 1, 3], np.stack([starts[0], starts[0], I, I]), probe, use_graph=False)
    say("  used_network %s inliers net %d/%d identity %d/%d" % (used.tolist(), pr[0].inliers, pr[1].inliers,
                                                                 pr[2].inliers, pr[3].inliers))
    for i in range(2):
        a, f = pr[i], pr[2 + i]
        fb_wins = (f.inliers >= 0.5 * a.inliers and f.error / f.inliers < a.error / a.inliers) if a.inliers else True
        assert used[i] == int(not fb_wins)
    assert used[0] == 1 and used[1] == 0
    np.testing.assert_array_equal(starts[1], I)
    # the refinement is the plain aligner from the chosen start, bit for bit
    Tp, _ = b.align([0, 2], [1, 3], starts, ap, use_graph=False)
    np.testing.assert_array_equal(T.view(np.uint32), Tp.view(np.uint32))
    e0 = synth.pose_error(T[0], T0)
    e1 = synth.pose_error(T[1], T1)
    say("  errors: large motion %.4f deg %.4f m, small motion %.4f deg %.4f m" % (e0 + e1))
    # the small-motion pair aligns from the identity; the large-motion pair settles at NICP's
    # own fixed point near the ground truth (about 1 deg away for this 10 deg / 0.3 m motion,
    # the same from the CPU oracle) and far closer than from the identity
    assert e1[0] < 0.1 and e1[1] < 0.002
    Ti, _ = b.align([0], [1], None, ap, use_graph=False)
    ei = synth.pose_error(Ti[0], T0)
    say("  large motion from identity: %.3f deg %.4f m" % ei)
    assert e0[0] < 2.0 and e0[1] < 0.02
    assert ei[0] > 2.0 * e0[0]


@pytest.mark.parametrize("sub", [8, 4, 2])
def test_probe_twin_equals_separate(ctx, sub):
    """The start probe's one-pass projection of both candidate starts (k_zproject_twin) gives the
    linearisations of two separate projections, bit for bit: several seeds, ragged point counts
    (different hole rates and a nearly empty frame), the predicted and the fallback start."""
    say("case probe twin sub=%d" % sub)
    rows, cols = 480, 640
    _, sp = params()
    _, ap = aparams()
    cm = cams(rows, cols)[1]
    frames, starts_net = [], []
    for k, (seed, holes) in enumerate([(601, 0.1), (602, 0.3), (603, 0.0), (604, 0.97)]):
        dr, dc, Tgt = synth.pair(seed, rows, cols, max_rot_deg=20.0, max_trans_m=0.3, holes=holes)
        frames += [dr, dc]
        starts_net.append(Tgt if k % 2 == 0 else O.v2t(np.array([0.02, -0.01, 0.03, 0.01, 0.02, -0.01], np.float32)))
    b = capi.Batch(ctx, [(rows, cols)] * len(frames), [cm] * len(frames), sp)
    for i, f in enumerate(frames):
        b.upload(i, f)
    b.convert()
    n = len(frames) // 2
    ref = np.arange(0, 2 * n, 2)
    cur = ref + 1
    I = np.eye(4, dtype=np.float32).flatten(order="F")
    starts = np.concatenate([np.stack(starts_net), np.tile(I, (n, 1))])
    ref2, cur2 = np.concatenate([ref, ref]), np.concatenate([cur, cur])
    rt = b.probe(ref2, cur2, starts, ap, sub=sub, twin=True)
    rs = b.probe(ref2, cur2, starts, ap, sub=sub, twin=False)
    for x, y in zip(rt, rs):
        assert x.inliers == y.inliers and x.num_correspondences == y.num_correspondences
        np.testing.assert_array_equal(np.array(x.H).view(np.uint64), np.array(y.H).view(np.uint64))
        np.testing.assert_array_equal(np.array(x.b).view(np.uint64), np.array(y.b).view(np.uint64))
        assert np.float32(x.error).view(np.uint32) == np.float32(y.error).view(np.uint32)
    say("  inliers %s" % [x.inliers for x in rt])
    assert rt[0].inliers > 100
    with pytest.raises(capi.NicpError):
        b.probe(ref2, np.concatenate([cur, ref]), starts, ap, sub=sub, twin=True)
    b.close()
