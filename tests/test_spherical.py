"""Pins the oracle's SphericalPointProjector restatement (oracle/nicp_oracle.cpp,
nicp/nicp/sphericalpointprojector.h:252-303, .cpp:19-141) with FP64 closed forms, round trips
and brute force, and checks the converter and aligner on panoramic frames. The reference
ships no tests or data for this projector."""
import numpy as np
import pytest

import oracle as O
from tests.helpers import identity16, spherical_cams, spherical_room

ROWS, COLS = 180, 360


def _metres(raw):
    return np.where(raw > 0, np.float32(0.001) * raw.astype(np.float32), 0).astype(np.float32)


def _unproject(depth, cam):
    rows, cols = depth.shape
    pts = np.zeros((rows * cols, 4), np.float32)
    idx = np.empty(rows * cols, np.int32)
    f = O.lib().orc_spherical_unproject
    f.restype = O.C.c_int
    n = f(O._p(depth), rows, cols, O.C.byref(cam), O._p(identity16()), O._p(pts), O._p(idx))
    return pts[:n].copy(), idx.reshape(rows, cols)


def _project(points, T16, cam, rows, cols):
    idx = np.empty(rows * cols, np.int32)
    dep = np.empty(rows * cols, np.float32)
    points = np.ascontiguousarray(points, np.float32)
    O.lib().orc_spherical_project(O._p(points), len(points), O._p(np.ascontiguousarray(T16, np.float32)),
                                  O.C.byref(cam), rows, cols, O._p(idx), O._p(dep))
    return idx.reshape(rows, cols), dep.reshape(rows, cols)


@pytest.mark.parametrize("hfov,vfov", [(np.pi, np.pi / 4), (2.0, 0.5)])
def test_unproject_closed_form(hfov, vfov):
    oc, _ = spherical_cams(ROWS, COLS, hfov, vfov)
    d = _metres(spherical_room(ROWS, COLS, oc, seed=1))
    pts, idx = _unproject(d, oc)
    hres = COLS / (2.0 * hfov)
    vres = ROWS / (2.0 * vfov)
    r, c = np.mgrid[0:ROWS, 0:COLS]
    th = (c - oc.horizontal_center) / hres
    ph = (r - oc.vertical_center) / vres
    inside = (np.abs(th) <= hfov * (1 - 1e-6)) & (np.abs(ph) <= vfov * (1 - 1e-6))
    outside = (np.abs(th) > hfov * (1 + 1e-6)) | (np.abs(ph) > vfov * (1 + 1e-6))
    valid = (d >= oc.min_distance) & (d <= oc.max_distance)
    got = idx >= 0
    assert np.array_equal(got[inside | outside], (valid & inside)[inside | outside])
    assert np.array_equal(idx[got], np.arange(got.sum()))
    z = d[got].astype(np.float64)
    expect = np.stack([np.sin(th[got]) * z, np.tan(ph[got]) * z, np.cos(th[got]) * z], -1)
    assert np.all(np.abs(pts[:, :3] - expect) <= 2e-6 * (1 + np.abs(expect)))
    assert np.all(pts[:, 3] == 1)
    # the cylinder distance of every point is its depth
    assert np.allclose(np.hypot(pts[:, 0], pts[:, 2]), z, rtol=1e-6)


def test_project_round_trip_and_zbuffer():
    oc, _ = spherical_cams(ROWS, COLS)
    d = _metres(spherical_room(ROWS, COLS, oc, seed=2))
    pts, idx = _unproject(d, oc)
    pidx, pdep = _project(pts, identity16(), oc, ROWS, COLS)
    # truncation puts a point on its own pixel or the one before it
    rr, cc = np.nonzero(pidx >= 0)
    own_r, own_c = np.nonzero(idx >= 0)
    where = {int(i): (r, c) for r, c, i in zip(own_r, own_c, idx[own_r, own_c])}
    moved = 0
    for r, c in zip(rr, cc):
        i = int(pidx[r, c])
        r0, c0 = where[i]
        assert r0 - 1 <= r <= r0 and c0 - 1 <= c <= c0
        moved += (r, c) != (r0, c0)
    assert pidx[rr, cc].size > 0.5 * len(pts)
    # brute-force z-buffer in FP64 from a moved projector (the round trip above sits on pixel
    # borders by construction): each point's pixel by truncation, the nearest per pixel (ties
    # to the lower index); pixels with a point within 1e-3 of a pixel border are skipped
    from nicp_amd import se3
    T = se3.v2t(np.array([0.03, 0.01, -0.02, 0.01, 0.015, -0.005])).astype(np.float32)
    pidx, pdep = _project(pts, T, oc, ROWS, COLS)
    Tm = T.reshape(4, 4, order="F").astype(np.float64)
    cp = pts.astype(np.float64) @ np.linalg.inv(Tm).T
    theta = np.arctan2(cp[:, 0], cp[:, 2])
    dist = np.hypot(cp[:, 0], cp[:, 2])
    phi = np.arctan2(cp[:, 1], dist)
    hres, vres = COLS / (2 * oc.horizontal_fov), ROWS / (2 * oc.vertical_fov)
    xf = hres * theta + oc.horizontal_center
    yf = vres * phi + oc.vertical_center
    near = (np.abs(xf - np.round(xf)) < 1e-3) | (np.abs(yf - np.round(yf)) < 1e-3)
    px, py = np.trunc(xf).astype(np.int64), np.trunc(yf).astype(np.int64)
    ok = (px >= 0) & (px < COLS) & (py >= 0) & (py < ROWS) & (dist >= oc.min_distance) & (dist <= oc.max_distance)
    best = np.full((ROWS, COLS), -1)
    unsure = np.zeros((ROWS, COLS), bool)
    for i in np.nonzero(ok)[0]:
        y, x = py[i], px[i]
        if near[i]:  # float rounding may put it on either side of the border
            unsure[max(y - 1, 0):y + 2, max(x - 1, 0):x + 2] = True
        j = best[y, x]
        if j < 0 or dist[i] < dist[j]:
            best[y, x] = i
    sure = ~unsure
    assert sure.mean() > 0.9
    assert np.array_equal(pidx[sure], best[sure])
    hit = pidx >= 0
    assert np.allclose(pdep[hit], dist[pidx[hit]], rtol=1e-6)


def test_projection_fov_quirk():
    """_project tests fabs(theta > fov): an azimuth below -fov is not rejected by the field of
    view (a point at -fov - 0.05 with the centre moved right lands inside the image)."""
    oc, _ = spherical_cams(ROWS, COLS, 1.0, 0.5)
    oc.horizontal_center = COLS / 2.0 + 40
    th = np.array([-1.05, 1.05, 0.0])
    p = np.stack([2 * np.sin(th), np.zeros(3), 2 * np.cos(th), np.ones(3)], -1).astype(np.float32)
    idx, _ = _project(p, identity16(), oc, ROWS, COLS)
    hit = set(idx[idx >= 0].tolist())
    assert 0 in hit and 1 not in hit and 2 in hit


def test_intervals_closed_form():
    oc, _ = spherical_cams(ROWS, COLS, 2.5, 0.7)
    d = _metres(spherical_room(ROWS, COLS, oc, seed=3))
    iv = np.empty(ROWS * COLS * 2, np.int32)
    O.lib().orc_spherical_intervals(O._p(d), ROWS, COLS, O.C.byref(oc), O.C.c_float(0.25), O._p(iv))
    iv = iv.reshape(ROWS, COLS, 2)
    valid = (d >= oc.min_distance) & (d <= oc.max_distance)
    assert np.all(iv[~valid] == -1)
    a = np.arctan2(0.25, d[valid].astype(np.float64))
    ex = COLS / (2 * 2.5) * a
    ey = ROWS / (2 * 0.7) * a
    sure = (np.abs(ex - np.round(ex)) > 1e-4) & (np.abs(ey - np.round(ey)) > 1e-4)
    assert np.array_equal(iv[valid][sure, 0], np.floor(ex[sure]).astype(np.int32))
    assert np.array_equal(iv[valid][sure, 1], np.floor(ey[sure]).astype(np.int32))


def test_conversion_suppresses_radius_limits():
    """The converter's stats ignore min/maxImageRadius for this projector: normals exist with a
    minimum radius no interval reaches, and agree with the FP64 brute force."""
    oc, _ = spherical_cams(ROWS, COLS)
    raw = spherical_room(ROWS, COLS, oc, seed=4)
    sp = O.stats_params(world_radius=0.25, min_image_radius=1000, max_image_radius=2, min_points=30,
                        curvature_threshold=0.075)
    f = O.Frame(raw, 0.001, oc, sp)
    good = np.linalg.norm(f.normals[:, :3], axis=1) > 0.5
    assert good.mean() > 0.5
    n64, c64, valid = O.stats_f64(raw, 0.001, oc, sp, f)
    both = good & (valid == 1)
    cosang = np.abs(np.sum(f.normals[both, :3] * n64[both], axis=1))
    # FP32 prefix sums lose digits over large windows; the windows themselves agree
    assert np.median(cosang) > 0.999999 and np.quantile(cosang, 0.05) > 0.999


def test_oracle_aligns_panoramic_pair():
    oc, _ = spherical_cams(ROWS, COLS)
    sp = O.stats_params(world_radius=0.25, min_image_radius=10, max_image_radius=100, min_points=30,
                        curvature_threshold=0.075)
    from nicp_amd import se3
    Tgt = se3.v2t(np.array([0.05, 0.0, -0.04, 0.0, 0.02, 0.0])).astype(np.float32)
    ref = O.Frame(spherical_room(ROWS, COLS, oc, seed=5), 0.001, oc, sp)
    cur = O.Frame(spherical_room(ROWS, COLS, oc, pose=Tgt, seed=6), 0.001, oc, sp)
    ap = O.align_params(outer_iterations=20)
    T, res = O.align(ref.cloud(), cur.cloud(), ROWS, COLS, oc, identity16(), ap, 1, fp64_solve=True)
    from nicp_amd import synth
    rot, tr = synth.pose_error(T, Tgt)
    assert rot < 0.2 and tr < 0.01, (rot, tr)
