"""Pins the CPU oracle of the initial-guess CNN (oracle/initnet.py) against
independent computations: torch's float64 conv2d/linear and autograd, central
finite differences, torch's bf16 rounding and the im2col/col2im adjoint identity."""
import numpy as np
import pytest
import torch

from oracle import initnet as net


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.empty(net.num_params())
    off = net.offsets()
    for l in range(len(net.LAYERS)):
        K = net.layer_k(l)
        p[off[l]:off[l + 1]] = rng.normal(0.0, scale * np.sqrt(2.0 / K), off[l + 1] - off[l])
    return p


def _input(seed, B):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.5, 4.5, (B,) + net.IN_SHAPE)
    x[rng.uniform(size=x.shape) < 0.1] = 0.0
    return x


def _torch_forward(params, x):
    """The same network built from torch.nn.functional ops in float64."""
    h = torch.tensor(x, dtype=torch.float64).permute(0, 3, 1, 2)
    for l, (ih, iw, ic, oh, ow, oc, conv, relu) in enumerate(net.LAYERS):
        W = torch.tensor(net.layer_matrix(params, l), dtype=torch.float64, requires_grad=False)
        if conv:
            w = W[:-1].reshape(3, 3, ic, oc).permute(3, 2, 0, 1).contiguous()
            h = torch.nn.functional.conv2d(h, w, W[-1], stride=2, padding=1)
            assert h.shape[2:] == (oh, ow)
        else:
            if h.dim() == 4:
                h = h.permute(0, 2, 3, 1).reshape(h.shape[0], -1)  # NHWC flatten
            h = h @ W[:-1] + W[-1]
        if relu:
            h = torch.relu(h)
    return h


@pytest.mark.parametrize("input_mode", [0, 1])
def test_forward_matches_torch(input_mode):
    p, x = _params(1), _input(2, 3)
    ours = net.forward(p, x, input_mode=input_mode)[-1]
    feat = x if input_mode == 0 else np.where(x != 0, 1.0 / np.where(x != 0, x, 1.0), 0.0)
    ref = _torch_forward(p, feat).numpy()
    assert ours.shape == (3, 6)
    assert np.max(np.abs(ours - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_gradient_matches_torch_autograd():
    p, x = _params(3), _input(4, 2)
    t = np.random.default_rng(5).normal(size=(2, 6))
    w = np.array([1.0, 1.0, 1.0, 4.0, 4.0, 4.0])
    loss, g = net.loss_and_grad(p, x, t, w)
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    # rebuild the torch network on the flat parameter tensor so autograd reaches it
    h = torch.tensor(x, dtype=torch.float64).permute(0, 3, 1, 2)
    off = net.offsets()
    for l, (ih, iw, ic, oh, ow, oc, conv, relu) in enumerate(net.LAYERS):
        W = pt[off[l]:off[l + 1]].reshape(net.layer_k(l) + 1, oc)
        if conv:
            h = torch.nn.functional.conv2d(h, W[:-1].reshape(3, 3, ic, oc).permute(3, 2, 0, 1), W[-1], stride=2,
                                           padding=1)
        else:
            if h.dim() == 4:
                h = h.permute(0, 2, 3, 1).reshape(h.shape[0], -1)
            h = h @ W[:-1] + W[-1]
        if relu:
            h = torch.relu(h)
    lt = (torch.tensor(w) * (h - torch.tensor(t)) ** 2).sum() / 2
    lt.backward()
    assert abs(loss - lt.item()) < 1e-12 * abs(lt.item())
    ref = pt.grad.numpy()
    assert np.max(np.abs(g - ref)) < 1e-10 * np.max(np.abs(ref))


def test_gradient_finite_differences():
    p, x = _params(6), _input(7, 2)
    t = np.random.default_rng(8).normal(size=(2, 6)) * 0.1
    _, g = net.loss_and_grad(p, x, t)
    rng = np.random.default_rng(9)
    off = net.offsets()
    h = 1e-6
    for l in range(len(net.LAYERS)):
        # weights and the bias row of every layer
        idx = list(rng.integers(off[l], off[l + 1], 6)) + [off[l + 1] - 1]
        for i in idx:
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (net.loss_and_grad(pp, x, t)[0] - net.loss_and_grad(pm, x, t)[0]) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (l, i, fd, g[i])


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(10)
    v = np.concatenate([rng.normal(size=4096).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 4096),
                        np.float32([0.0, -0.0, 1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 3.0e38, -3.4e38,
                                    np.inf, -np.inf, 1e-40])])
    ours = net.bf16(v)
    ref = torch.tensor(v).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(ours, ref)
    assert np.isnan(net.bf16(np.float32([np.nan]))[0])


@pytest.mark.parametrize("l", [0, 1, 2, 3])
def test_im2col_col2im_adjoint(l):
    ih, iw, ic, oh, ow, oc, conv, relu = net.LAYERS[l]
    rng = np.random.default_rng(l)
    x = rng.normal(size=(2, ih, iw, ic))
    y = rng.normal(size=(2 * oh * ow, 9 * ic))
    lhs = np.sum(net.im2col(x, l) * y)
    rhs = np.sum(x * net.col2im(y, l, 2))
    assert abs(lhs - rhs) < 1e-9 * max(1.0, abs(lhs))


def test_bf16_mode_close_to_exact():
    # the device semantics differ from exact arithmetic only by operand rounding
    p, x = _params(11), _input(12, 2)
    ye = net.forward(p, x)[-1]
    yb = net.forward(p, x, "bf16")[-1]
    assert np.linalg.norm(yb - ye) < 3e-2 * np.linalg.norm(ye)
