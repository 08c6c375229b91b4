"""CPU oracle of the initial-guess CNN -- TEST INFRASTRUCTURE ONLY.

Only the test suite may import this module, as the checker of the device network
(no reference counterpart: the reference's Aligner takes its initial guess from the
caller, nicp/nicp/aligner.h:102-106).

Network: four 3x3 stride-2 padding-1 convolutions (2->32->32->64->64 channels, ReLU)
over a (60, 80, 2) NHWC input, then fully connected 1280->256 (ReLU) and 256->6.
Parameters are one flat float32 vector, layer by layer a row-major (K + 1) x N matrix:
the reduction index K of a convolution is (ky, kx, ci) with ci fastest, the last row
is the bias.

The first layer reads the input as depth (input_mode 0) or inverse depth (1).

Two arithmetic modes:
  exact  -- float64 throughout (pinned against torch's conv2d/linear and autograd, and
            against finite differences);
  bf16   -- what the device computes: every GEMM operand is rounded to bf16 (round to
            nearest even), products and sums are exact (float64) and each GEMM result is
            stored as float32; the loss gradient is formed in float32.
"""
import numpy as np

# (ih, iw, ic, oh, ow, oc, conv, relu)
LAYERS = [
    (60, 80, 2, 30, 40, 32, True, True),
    (30, 40, 32, 15, 20, 32, True, True),
    (15, 20, 32, 8, 10, 64, True, True),
    (8, 10, 64, 4, 5, 64, True, True),
    (1, 1, 1280, 1, 1, 256, False, True),
    (1, 1, 256, 1, 1, 6, False, False),
]
IN_SHAPE = (60, 80, 2)


def layer_k(l):
    ih, iw, ic, oh, ow, oc, conv, relu = LAYERS[l]
    return 9 * ic if conv else ic


def offsets():
    off = [0]
    for l in range(len(LAYERS)):
        off.append(off[-1] + (layer_k(l) + 1) * LAYERS[l][5])
    return off


def num_params():
    return offsets()[-1]


def layer_matrix(params, l):
    """(K + 1, N) float64 matrix of layer l (the last row is the bias)."""
    off = offsets()
    N = LAYERS[l][5]
    return np.asarray(params[off[l]:off[l + 1]], np.float64).reshape(layer_k(l) + 1, N)


def bf16(x):
    """Rounds to bf16 (nearest even), returned as float64 values."""
    x32 = np.ascontiguousarray(x, np.float32)
    u = x32.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(special, (u >> 16) | np.where((u & 0xFFFF) != 0, np.uint64(0x40), np.uint64(0)), r)
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def im2col(x, l):
    """x (B, ih, iw, ic) -> (B * oh * ow, 9 * ic), column order (ky, kx, ci)."""
    ih, iw, ic, oh, ow, oc, conv, relu = LAYERS[l]
    B = x.shape[0]
    xp = np.zeros((B, ih + 2, iw + 2, ic), np.float64)
    xp[:, 1:ih + 1, 1:iw + 1, :] = x
    cols = np.empty((B, oh, ow, 3, 3, ic), np.float64)
    for ky in range(3):
        for kx in range(3):
            # input row 2 oy + ky - 1 is padded row 2 oy + ky
            cols[:, :, :, ky, kx, :] = xp[:, ky:ky + 2 * oh:2, kx:kx + 2 * ow:2, :]
    return cols.reshape(B * oh * ow, 9 * ic)


def col2im(dcols, l, B):
    """Adjoint of im2col: (B * oh * ow, 9 * ic) -> (B, ih, iw, ic)."""
    ih, iw, ic, oh, ow, oc, conv, relu = LAYERS[l]
    d = dcols.reshape(B, oh, ow, 3, 3, ic)
    xp = np.zeros((B, ih + 2 + 1, iw + 2 + 1, ic), np.float64)
    for ky in range(3):
        for kx in range(3):
            xp[:, ky:ky + 2 * oh:2, kx:kx + 2 * ow:2, :] += d[:, :, :, ky, kx, :]
    return xp[:, 1:ih + 1, 1:iw + 1, :]


def _operand(A, mode):
    return bf16(A) if mode == "bf16" else np.asarray(A, np.float64)


def _gemm(A, B, mode):
    out = _operand(A, mode) @ _operand(B, mode)
    return out.astype(np.float32).astype(np.float64) if mode == "bf16" else out


def input_feature(x, input_mode, mode="exact"):
    """The first layer's view of the input: depth (input_mode 0) or inverse depth (1, with
    0 kept for holes; a float32 division in the bf16 mode, as on the device)."""
    if input_mode == 0:
        return np.asarray(x, np.float64)
    if mode == "bf16":
        x32 = np.asarray(x, np.float32)
        with np.errstate(divide="ignore"):
            return np.where(x32 != 0, np.float32(1.0) / x32, np.float32(0.0)).astype(np.float64)
    x = np.asarray(x, np.float64)
    with np.errstate(divide="ignore"):
        return np.where(x != 0, 1.0 / x, 0.0)


def layer_input_matrix(x, l, input_mode=0, mode="exact"):
    """GEMM operand A of layer l with the bias column of ones appended."""
    ih, iw, ic, oh, ow, oc, conv, relu = LAYERS[l]
    B = x.shape[0]
    if l == 0:
        x = input_feature(x, input_mode, mode)
    A = im2col(x, l) if conv else np.asarray(x, np.float64).reshape(B, ic)
    return np.concatenate([A, np.ones((A.shape[0], 1))], axis=1)


def layer_forward(params, x, l, mode="exact", input_mode=0):
    """Output of layer l (after its ReLU) for input x; NHWC shaped for convolutions."""
    ih, iw, ic, oh, ow, oc, conv, relu = LAYERS[l]
    B = x.shape[0]
    y = _gemm(layer_input_matrix(x, l, input_mode, mode), layer_matrix(params, l), mode)
    if relu:
        y = np.maximum(y, 0.0)
    return y.reshape(B, oh, ow, oc) if conv else y


def layer_backward(params, x, g, l, mode="exact", input_mode=0):
    """(dW (K + 1, N), dX) of layer l for its input x and the gradient g of its output
    after the ReLU mask has been applied (shape (B * oh * ow, N)); for layer 0, dX is the
    gradient w.r.t. the input feature."""
    ih, iw, ic, oh, ow, oc, conv, relu = LAYERS[l]
    B = x.shape[0]
    A = layer_input_matrix(x, l, input_mode, mode)
    G = np.asarray(g, np.float64).reshape(-1, oc)
    dW = _operand(A, mode).T @ _operand(G, mode)
    W = layer_matrix(params, l)[:-1]
    dcol = _operand(G, mode) @ _operand(W, mode).T
    dX = col2im(dcol, l, B) if conv else dcol.reshape(B, ic)
    return dW, dX


def forward(params, x, mode="exact", input_mode=0):
    """Activations [input, layer 0 output, ..., layer 5 output]."""
    acts = [np.asarray(x, np.float64)]
    for l in range(len(LAYERS)):
        acts.append(layer_forward(params, acts[-1], l, mode, input_mode))
    return acts


def loss_and_grad(params, x, target, weights=(1, 1, 1, 1, 1, 1), mode="exact", input_mode=0):
    """Weighted squared error 1/B sum_b sum_j w_j (y_bj - t_bj)^2 and its parameter gradient."""
    acts = forward(params, x, mode, input_mode)
    y = acts[-1]
    B = y.shape[0]
    t = np.asarray(target, np.float64).reshape(B, 6)
    w = np.asarray(weights, np.float64)
    if mode == "bf16":
        d = (y.astype(np.float32) - t.astype(np.float32))
        w32 = w.astype(np.float32)
        g = ((np.float32(2.0) * w32) * d / np.float32(B)).astype(np.float64)
        loss = float(np.sum(w * d.astype(np.float64) ** 2) / B)
    else:
        d = y - t
        g = 2.0 * w * d / B
        loss = float(np.sum(w * d * d) / B)
    grads = np.zeros(num_params(), np.float64)
    off = offsets()
    for l in reversed(range(len(LAYERS))):
        ih, iw, ic, oh, ow, oc, conv, relu = LAYERS[l]
        out = acts[l + 1].reshape(-1, oc)
        G = g.reshape(-1, oc)
        if relu:
            G = np.where(out > 0.0, G, 0.0)
        dW, dX = layer_backward(params, acts[l], G, l, mode, input_mode)
        grads[off[l]:off[l + 1]] = dW.reshape(-1)
        if mode == "bf16":
            dX = dX.astype(np.float32).astype(np.float64)
        g = dX
    return loss, grads
