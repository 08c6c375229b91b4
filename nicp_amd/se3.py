"""Pose 6-vectors (t, qx, qy, qz) <-> 4x4 transforms, float64 host versions of
v2t / t2v (nicp/nicp/bm_se3.h:10-51): the quaternion's vector part with w >= 0
implied. Transforms are float[16] column-major, as everywhere in the C ABI."""
import numpy as np


def t2v(T16):
    """(..., 16) column-major transforms -> (..., 6) vectors."""
    T = np.asarray(T16, np.float64).reshape(-1, 4, 4).transpose(0, 2, 1)  # row-major view
    R = T[:, :3, :3]
    out = np.empty((T.shape[0], 6))
    out[:, :3] = T[:, :3, 3]
    # quaternion of R (Shepperd's method), normalised, w >= 0
    q = np.empty((T.shape[0], 4))  # w x y z
    tr = R[:, 0, 0] + R[:, 1, 1] + R[:, 2, 2]
    for i in range(T.shape[0]):
        m = R[i]
        if tr[i] > 0:
            s = 2.0 * np.sqrt(tr[i] + 1.0)
            q[i] = [0.25 * s, (m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s, (m[1, 0] - m[0, 1]) / s]
        elif m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
            s = 2.0 * np.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2])
            q[i] = [(m[2, 1] - m[1, 2]) / s, 0.25 * s, (m[0, 1] + m[1, 0]) / s, (m[0, 2] + m[2, 0]) / s]
        elif m[1, 1] > m[2, 2]:
            s = 2.0 * np.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2])
            q[i] = [(m[0, 2] - m[2, 0]) / s, (m[0, 1] + m[1, 0]) / s, 0.25 * s, (m[1, 2] + m[2, 1]) / s]
        else:
            s = 2.0 * np.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1])
            q[i] = [(m[1, 0] - m[0, 1]) / s, (m[0, 2] + m[2, 0]) / s, (m[1, 2] + m[2, 1]) / s, 0.25 * s]
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q[q[:, 0] < 0] *= -1.0
    out[:, 3:] = q[:, 1:]
    return out.reshape(np.shape(T16)[:-1] + (6,)) if np.ndim(T16) > 1 else out[0]


def v2t(x6):
    """(..., 6) vectors -> (..., 16) column-major transforms."""
    x = np.asarray(x6, np.float64).reshape(-1, 6)
    qx, qy, qz = x[:, 3], x[:, 4], x[:, 5]
    qw = np.sqrt(np.maximum(0.0, 1.0 - (qx * qx + qy * qy + qz * qz)))
    T = np.zeros((x.shape[0], 4, 4))
    T[:, 0, 0] = 1 - 2 * (qy * qy + qz * qz)
    T[:, 0, 1] = 2 * (qx * qy - qw * qz)
    T[:, 0, 2] = 2 * (qx * qz + qw * qy)
    T[:, 1, 0] = 2 * (qx * qy + qw * qz)
    T[:, 1, 1] = 1 - 2 * (qx * qx + qz * qz)
    T[:, 1, 2] = 2 * (qy * qz - qw * qx)
    T[:, 2, 0] = 2 * (qx * qz - qw * qy)
    T[:, 2, 1] = 2 * (qy * qz + qw * qx)
    T[:, 2, 2] = 1 - 2 * (qx * qx + qy * qy)
    T[:, :3, 3] = x[:, :3]
    T[:, 3, 3] = 1.0
    out = T.transpose(0, 2, 1).reshape(-1, 16).astype(np.float32)
    return out.reshape(np.shape(x6)[:-1] + (16,)) if np.ndim(x6) > 1 else out[0]
