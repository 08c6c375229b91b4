"""Batches fed from device memory and two batches in flight on two contexts (two streams).

- nicp_batch_download_depth_device / nicp_batch_upload_depth_device move every frame's depth
  of an uneven batch to and from one packed device array: the packed array equals the host
  images, and a batch loaded from it converts to the same clouds, bit for bit, as one loaded
  from the host.
- Two contexts driven from two host threads (the pipelined benchmark) give every pair the
  transform, bit for bit, that the same calls give one after the other on one context.
- A learned-start call with many pairs, then one with fewer on the same network (whose buffers,
  reserved for the larger call, are reused): the smaller call equals the same call on a fresh
  network, bit for bit, replayed from its graphs.

Reference: DepthImageConverterIntegralImage::compute (nicp/nicp/depthimageconverterintegralimage.cpp:18-64)
and AlignerProjective::align (nicp/nicp/alignerprojective.cpp:21-136), which both paths run.
"""
import sys
import threading

import numpy as np
import pytest

from nicp_amd import capi, initnet, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _batch(ctx, shapes):
    cams = [capi.camera(*synth.camera_for(r, c)) for r, c in shapes]
    return capi.Batch(ctx, shapes, cams, capi.stats_params())


def test_depth_device_roundtrip(ctx):
    # the odd-sized frame leaves the frames after it at odd offsets of the packed array, so
    # both the 8-byte and the 2-byte copy paths run
    shapes = [(120, 160), (480, 640), (121, 161), (240, 320), (960, 1280), (120, 160)]
    say("case depth_device_roundtrip frames=%d" % len(shapes))
    imgs = [synth.pair(40 + i, r, c)[0] for i, (r, c) in enumerate(shapes)]
    a = _batch(ctx, shapes)
    for i, d in enumerate(imgs):
        a.upload(i, d)
    n = a.num_pixels()
    assert n == sum(r * c for r, c in shapes)
    store = capi.DeviceBuffer(ctx, (1, n), np.uint16)
    a.download_depth_device(store.ptr)
    packed = store.download()[0]
    np.testing.assert_array_equal(packed, np.concatenate([d.ravel() for d in imgs]))
    b = _batch(ctx, shapes)
    b.upload_depth_device(store.ptr)
    a.convert()
    b.convert()
    for i, shape in enumerate(shapes):
        ca, cb = a.frame_cloud(i, shape), b.frame_cloud(i, shape)
        for xa, xb in zip(ca.download(), cb.download()):
            np.testing.assert_array_equal(np.asarray(xa).view(np.uint32), np.asarray(xb).view(np.uint32))
        np.testing.assert_array_equal(ca.index_image(), cb.index_image())
        ca.close()
        cb.close()
    store.close()
    a.close()
    b.close()


def test_two_contexts_equal_one(ctx):
    """Learned steps of two batches on two contexts from two threads == the same steps in turn."""
    w = initnet.load_weights()
    assert w is not None, "trained weights missing"
    B, steps = 32, 3
    say("case two_contexts pairs=%d steps=%d" % (B, steps))
    ap = capi.align_params()
    ref = np.arange(0, 2 * B, 2, dtype=np.int32)
    cur = ref + 1
    ctx2 = capi.Context(0)
    lanes = []
    for c in (ctx, ctx2):
        net = capi.InitNet(c, seed=1)
        net.set_params(w)
        net.set_input_mode(0)
        batch = initnet.make_batch(c, B)
        lanes.append((c, net, batch))
    # the depth of every step, rendered once into one packed store
    n = lanes[0][2].num_pixels()
    store = capi.DeviceBuffer(ctx, (2 * steps, n), np.uint16)
    for k in range(2 * steps):
        batch = lanes[0][2]
        initnet.render_pairs(batch, np.arange(5000 + k * B, 5000 + (k + 1) * B, dtype=np.uint64), 30.0, 0.5)
        batch.download_depth_device(store.row_ptr(k))
    ctx.synchronize()

    def run(lane, ks, out):
        c, net, batch = lanes[lane]
        for k in ks:
            batch.upload_depth_device(store.row_ptr(k))
            batch.convert()
            out[k] = batch.align_learned(net, ref, cur, ap, use_graph=True, want_results=True)[0]
        c.synchronize()

    seq = {}
    run(0, range(2 * steps), seq)
    par = {}
    th = [threading.Thread(target=run, args=(i, range(i, 2 * steps, 2), par)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for k in range(2 * steps):
        np.testing.assert_array_equal(par[k].view(np.uint32), seq[k].view(np.uint32))
    store.close()
    for c, net, batch in lanes:
        batch.close()
        net.close()
    ctx2.close()


def test_learned_fewer_pairs_after_more(ctx):
    """Learned starts for 64 pairs, then for the first 8 of them on the same network and batch,
    equal the same 8-pair call on a fresh network and batch, bit for bit (graph replay). The
    network's buffers stay reserved for 64 pairs; an 8-pair forward splits its fully connected
    layers' reductions further than a 64-pair one, so this covers the reservation of every
    smaller batch."""
    w = initnet.load_weights()
    assert w is not None, "trained weights missing"
    say("case learned_fewer_pairs_after_more pairs=64,8")
    ap = capi.align_params()
    seeds = np.arange(7000, 7064, dtype=np.uint64)

    def setup(B):
        net = capi.InitNet(ctx, seed=1)
        net.set_params(w)
        net.set_input_mode(0)
        batch = initnet.make_batch(ctx, B)
        initnet.render_pairs(batch, seeds[:B], 30.0, 0.5)
        batch.convert()
        return net, batch

    net_a, batch_a = setup(64)
    ref = np.arange(0, 128, 2, dtype=np.int32)
    batch_a.align_learned(net_a, ref, ref + 1, ap, use_graph=True, want_results=True)
    Ta = [batch_a.align_learned(net_a, ref[:8], ref[:8] + 1, ap, use_graph=True, want_results=True)[0]
          for _ in range(2)]
    net_b, batch_b = setup(8)
    Tb = batch_b.align_learned(net_b, ref[:8], ref[:8] + 1, ap, use_graph=True, want_results=True)[0]
    for T in Ta:
        np.testing.assert_array_equal(T.view(np.uint32), Tb.view(np.uint32))
    for o in (batch_a, net_a, batch_b, net_b):
        o.close()
