"""Seeded synthetic depth pairs (bindings of ``csrc/synth.cpp``).

Scenes of 20 random planes/boxes rendered by a pinhole camera to uint16
millimetre depth, with depth noise sigma = 1.5e-3 * z^2 m and 10 % holes.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnicp_synth.so")
_lib = None

#: camera intrinsics per resolution: fx scaled by width / 640, principal point at the pixel centre
CAMERAS = {
    (120, 160): (131.25, 131.25, 79.5, 59.5),
    (240, 320): (262.5, 262.5, 159.5, 119.5),
    (480, 640): (525.0, 525.0, 319.5, 239.5),
    (960, 1280): (1050.0, 1050.0, 639.5, 479.5),
    (60, 80): (65.625, 65.625, 39.5, 29.5),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libnicp_synth.so is not built: run __graft_entry__.build()")
        _lib = C.CDLL(LIB_PATH)
        _lib.nicp_synth_pair.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                         C.c_float, C.c_float, C.c_float, C.c_float, C.c_void_p, C.c_void_p,
                                         C.c_void_p]
        _lib.nicp_synth_render.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_float, C.c_float,
                                           C.c_float, C.c_float, C.c_void_p, C.c_float, C.c_float, C.c_void_p]
        _lib.nicp_synth_motion.argtypes = [C.c_uint64, C.c_float, C.c_float, C.c_void_p]
        _lib.nicp_synth_cloud_pair.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                               C.c_float, C.c_float, C.c_float, C.c_float] + [C.c_void_p] * 8
    return _lib


def camera_for(rows, cols):
    if (rows, cols) in CAMERAS:
        return CAMERAS[(rows, cols)]
    f = 525.0 * cols / 640.0
    return (f, f, (cols - 1) / 2.0, (rows - 1) / 2.0)


def pair(seed, rows=480, cols=640, max_rot_deg=5.0, max_trans_m=0.05, noise=1.5e-3, holes=0.1, intrinsics=None):
    """Returns (depth_ref uint16[rows, cols], depth_cur, T_gt float32[4, 4] column-major flattened to 16)."""
    fx, fy, cx, cy = intrinsics or camera_for(rows, cols)
    dr = np.empty((rows, cols), np.uint16)
    dc = np.empty((rows, cols), np.uint16)
    T = np.empty(16, np.float32)
    rc = lib().nicp_synth_pair(seed, rows, cols, fx, fy, cx, cy, max_rot_deg, max_trans_m, noise, holes,
                               dr.ctypes.data, dc.ctypes.data, T.ctypes.data)
    if rc != 0:
        raise RuntimeError("nicp_synth_pair failed")
    return dr, dc, T


def render(scene_seed, pose16, rows, cols, stream=0, noise=1.5e-3, holes=0.1, intrinsics=None):
    fx, fy, cx, cy = intrinsics or camera_for(rows, cols)
    d = np.empty((rows, cols), np.uint16)
    pose = np.ascontiguousarray(pose16, np.float32)
    rc = lib().nicp_synth_render(scene_seed, stream, rows, cols, fx, fy, cx, cy, pose.ctypes.data, noise, holes,
                                 d.ctypes.data)
    if rc != 0:
        raise RuntimeError("nicp_synth_render failed")
    return d


def motion(seed, max_rot_deg, max_trans_m):
    T = np.empty(16, np.float32)
    lib().nicp_synth_motion(seed, max_rot_deg, max_trans_m, T.ctypes.data)
    return T


def pose_error(T16, Tgt16):
    """(rotation error in degrees, translation error in metres) of T against T_gt."""
    A = np.asarray(T16, np.float64).reshape(4, 4, order="F")
    B = np.asarray(Tgt16, np.float64).reshape(4, 4, order="F")
    D = np.linalg.inv(B) @ A
    R = D[:3, :3]
    # atan2 of the skew part and the trace: accurate for small angles, unlike arccos of the
    # trace, whose resolution near the identity is ~sqrt(machine epsilon) of the inputs
    s = 0.5 * np.linalg.norm([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
    c = 0.5 * (np.trace(R) - 1.0)
    ang = np.degrees(np.arctan2(s, c))
    return float(ang), float(np.linalg.norm(D[:3, 3]))


def cloud_pair(seed, n=200000, max_rot_deg=10.0, max_trans_m=0.2, noise=0.002, rows=480, cols=640):
    """Unstructured source/target clouds of the kNN configuration.

    Returns dict with src/tgt points and normals (n, 4), curvature (n,), the target's flat
    information matrices omega_p/omega_n (n, 16) and T_gt (16,) with p_src = T_gt p_tgt."""
    fx, fy, cx, cy = camera_for(rows, cols)
    out = {k: np.empty((n, 4), np.float32) for k in ("src_points", "src_normals", "tgt_points", "tgt_normals")}
    out["curvature"] = np.empty(n, np.float32)
    out["omega_p"] = np.empty((n, 16), np.float32)
    out["omega_n"] = np.empty((n, 16), np.float32)
    out["T_gt"] = np.empty(16, np.float32)
    rc = lib().nicp_synth_cloud_pair(seed, n, rows, cols, fx, fy, cx, cy, max_rot_deg, max_trans_m, noise,
                                     *[out[k].ctypes.data for k in ("src_points", "src_normals", "tgt_points",
                                                                     "tgt_normals", "curvature", "omega_p", "omega_n",
                                                                     "T_gt")])
    if rc != 0:
        raise RuntimeError("nicp_synth_cloud_pair failed (%d)" % rc)
    return out
