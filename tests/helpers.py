"""Shared test helpers: synthetic frames and oracle / device parameter sets."""
import numpy as np

import oracle as O
from nicp_amd import capi, synth

RES = {"160x120": (120, 160), "320x240": (240, 320), "640x480": (480, 640), "1280x960": (960, 1280)}


def cams(rows, cols, min_d=0.5, max_d=4.5):
    fx, fy, cx, cy = synth.camera_for(rows, cols)
    return O.camera(fx, fy, cx, cy, min_d, max_d), capi.camera(fx, fy, cx, cy, min_d, max_d)


def params(**kw):
    return O.stats_params(**kw), capi.stats_params(**kw)


def aparams(**kw):
    return O.align_params(**kw), capi.align_params(**kw)


def identity16():
    return np.eye(4, dtype=np.float32).flatten(order="F")


def angle_dist(T16, U16):
    return synth.pose_error(T16, U16)


def spherical_room(rows, cols, cam, pose=None, seed=0, noise=0.002, holes=0.05):
    """A panoramic depth image (uint16 mm) of a box room with a few box obstacles, seen by the
    spherical camera `cam` from `pose` (4x4 column-major camera-to-room transform, default the
    identity). A pixel (r, c) looks along (sin t, tan f, cos t) with t = (c - hc) / hres and
    f = (r - vc) / vres, so the depth is the distance from the camera's vertical axis."""
    hres = cols / (2.0 * cam.horizontal_fov)
    vres = rows / (2.0 * cam.vertical_fov)
    cc, rr = np.meshgrid(np.arange(cols, dtype=np.float64), np.arange(rows, dtype=np.float64))
    t = (cc - cam.horizontal_center) / hres
    f = (rr - cam.vertical_center) / vres
    v = np.stack([np.sin(t), np.tan(f), np.cos(t)], -1)
    P = np.eye(4) if pose is None else np.asarray(pose, np.float64).reshape(4, 4, order="F")
    R, o = P[:3, :3], P[:3, 3]
    w = v @ R.T  # ray directions in the room; depth d scales v
    lo, hi = np.array([-4.0, -1.5, -5.0]), np.array([5.0, 1.2, 4.5])
    boxes = [(np.array([1.0, -1.5, 2.0]), np.array([2.0, 0.2, 3.0])),
             (np.array([-3.0, -1.5, -2.5]), np.array([-2.2, -0.4, -1.0])),
             (np.array([-1.0, 0.6, -4.0]), np.array([0.5, 1.2, -3.2])),
             (np.array([2.5, -1.5, -3.5]), np.array([3.5, -0.9, -2.0]))]
    with np.errstate(divide="ignore", invalid="ignore"):
        tr = np.where(w > 0, (hi - o) / w, np.where(w < 0, (lo - o) / w, np.inf))
        d = tr.min(-1)
        for bl, bh in boxes:
            t1, t2 = (bl - o) / w, (bh - o) / w
            tn = np.nanmax(np.minimum(t1, t2), -1)
            tf = np.nanmin(np.maximum(t1, t2), -1)
            hit = (tn <= tf) & (tn > 0)
            d = np.where(hit & (tn < d), tn, d)
    rng = np.random.default_rng(seed)
    d = d * (1.0 + noise * rng.standard_normal(d.shape))
    d[rng.uniform(size=d.shape) < holes] = 0
    d[~np.isfinite(d)] = 0
    return np.clip(np.round(d * 1000.0), 0, 65535).astype(np.uint16)


def spherical_cams(rows, cols, hfov=np.pi, vfov=np.pi / 4, min_d=0.5, max_d=10.0):
    return (O.spherical_camera(rows, cols, hfov, vfov, min_d, max_d),
            capi.spherical_camera(rows, cols, hfov, vfov, min_d, max_d))


def offset16(deg, axis, t):
    """An isometry (column-major 16 floats) built with scipy: rotation by deg about axis, then t."""
    from scipy.spatial.transform import Rotation
    M = np.eye(4)
    M[:3, :3] = Rotation.from_rotvec(np.radians(deg) * np.asarray(axis, np.float64) / np.linalg.norm(axis)).as_matrix()
    M[:3, 3] = t
    return M.astype(np.float32).flatten(order="F")


def offset_frame_arrays(f, O16):
    """Copies of an oracle frame's arrays moved by the sensor offset (Cloud::transformInPlace,
    N/depthimageconverterintegralimage.cpp:63), with the FP64 transform written here:
    points R p + t, normals R n, information matrices R W R^T."""
    M = O16.reshape(4, 4, order="F").astype(np.float64)
    R, t = M[:3, :3], M[:3, 3]
    pts = f.points.astype(np.float64).copy()
    pts[:, :3] = pts[:, :3] @ R.T + t
    nrm = f.normals.astype(np.float64).copy()
    nrm[:, :3] = nrm[:, :3] @ R.T
    R4 = np.eye(4)
    R4[:3, :3] = R
    R4[3, 3] = 0.0

    def om(W16):
        W = W16.reshape(-1, 4, 4).transpose(0, 2, 1)  # column-major records
        W = R4[None] @ W @ R4.T[None]
        return np.ascontiguousarray(W.transpose(0, 2, 1).reshape(-1, 16), np.float32)

    return (np.ascontiguousarray(pts, np.float32), np.ascontiguousarray(nrm, np.float32),
            f.curvature.copy(), om(f.omega_p), om(f.omega_n))
