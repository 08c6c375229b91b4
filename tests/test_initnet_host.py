"""Host side of the learned initial guess: pose 6-vectors, the pose transforms that go
with the input symmetries used in training, and the geometry that makes an image mirror
a scene mirror."""
import numpy as np
import pytest

import oracle as O
from nicp_amd import se3, synth
from nicp_amd.initnet import from_bf16_bits, symmetry_targets, to_bf16_bits


def _random_x(n, seed):
    rng = np.random.default_rng(seed)
    x = np.empty((n, 6))
    x[:, :3] = rng.uniform(-0.5, 0.5, (n, 3))
    ax = rng.normal(size=(n, 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    ang = rng.uniform(0, np.radians(30), n)
    x[:, 3:] = ax * np.sin(ang / 2)[:, None]
    return x


def test_se3_roundtrip_and_oracle():
    x = _random_x(50, 1)
    T = se3.v2t(x)
    np.testing.assert_allclose(se3.t2v(T), x, atol=2e-7)
    for i in range(5):
        np.testing.assert_allclose(T[i], O.v2t(x[i].astype(np.float32)), atol=2e-6)
        np.testing.assert_allclose(se3.t2v(T[i]), O.t2v(T[i]), atol=2e-6)


def _mat(T16):
    return np.asarray(T16, np.float64).reshape(4, 4, order="F")


@pytest.mark.parametrize("flags", range(8))
def test_symmetry_targets_are_the_pose_transforms(flags):
    x = _random_x(20, 2 + flags)
    y = symmetry_targets(x, np.full(20, flags))
    F = {2: np.diag([-1.0, 1.0, 1.0, 1.0]), 4: np.diag([1.0, -1.0, 1.0, 1.0])}
    for i in range(20):
        M = _mat(se3.v2t(x[i]))
        if flags & 1:
            M = np.linalg.inv(M)
        for bit in (2, 4):
            if flags & bit:
                M = F[bit] @ M @ F[bit]
        np.testing.assert_allclose(_mat(se3.v2t(y[i])), M, atol=2e-6)


def test_image_mirror_is_scene_mirror():
    # with the principal point at the image centre, the pixel of F p is the column-mirrored
    # pixel of p, and the pair's motion T becomes F T F
    fx, fy, cx, cy = synth.camera_for(480, 640)
    assert cx == (640 - 1) / 2 and cy == (480 - 1) / 2
    rng = np.random.default_rng(3)
    p = np.c_[rng.uniform(-1, 1, (100, 2)), rng.uniform(1, 4, 100), np.ones(100)]
    T = _mat(se3.v2t(_random_x(1, 4)[0]))
    F = np.diag([-1.0, 1.0, 1.0, 1.0])

    def pix(q):
        return np.c_[fx * q[:, 0] / q[:, 2] + cx, fy * q[:, 1] / q[:, 2] + cy]

    cur = (np.linalg.inv(T) @ p.T).T                      # p seen from the current camera
    cur_m = (np.linalg.inv(F @ T @ F) @ (F @ p.T)).T       # F p seen from the mirrored pair's current camera
    u, um = pix(cur), pix(cur_m)
    np.testing.assert_allclose(um[:, 0], 639 - u[:, 0], atol=1e-9)
    np.testing.assert_allclose(um[:, 1], u[:, 1], atol=1e-9)


def test_bf16_weight_storage():
    p = np.random.default_rng(5).normal(size=1000).astype(np.float32)
    q = from_bf16_bits(to_bf16_bits(p))
    assert np.max(np.abs(q - p) / np.abs(p)) <= 2.0 ** -8
    np.testing.assert_array_equal(from_bf16_bits(to_bf16_bits(q)), q)
