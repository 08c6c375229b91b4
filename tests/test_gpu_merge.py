"""GPU parity of the Gaussian unprojection, Cloud::transformInPlace, Cloud::add and
Merger::merge (nicp/nicp/pinholepointprojector.cpp:93-134, cloud.cpp:163-205,
merger.cpp:15-109) against the oracle, bit for bit: the oracle starts from the device's own
converted clouds, so both sides see identical inputs. Information matrices are compared on
their upper triangles, which is what the device keeps."""
import numpy as np
import pytest

import oracle as O
from nicp_amd import capi, synth
from tests.helpers import cams, identity16, params

pytestmark = pytest.mark.gpu

UPPER = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


class HostCloud:
    """The oracle's copy of a device cloud: points, normals, curvature, full 4x4 Stats and
    information matrices (column-major), Gaussian records."""

    def __init__(self, cl, stats=False):
        p, nr, cu, op, on = cl.download()
        self.points, self.normals, self.curvature = p, nr, cu
        self.omega_p, self.omega_n = op, on
        self.gauss = cl.download_gaussians()
        self.stats16 = None
        if stats:
            ev, el, mn, npt = cl.download_stats()
            n = len(p)
            S = np.zeros((n, 16), np.float32)
            for c in range(3):
                S[:, 4 * c:4 * c + 3] = ev[:, 3 * c:3 * c + 3]
            S[:, 12:16] = mn
            self.stats16 = S

    def arrays(self):
        return dict(points=self.points, normals=self.normals, stats16=self.stats16, omega_p=self.omega_p,
                    omega_n=self.omega_n, gaussians=self.gauss)

    def append(self, other):
        for k in ("points", "normals", "curvature", "omega_p", "omega_n", "gauss"):
            setattr(self, k, np.ascontiguousarray(np.concatenate([getattr(self, k), getattr(other, k)])))
        if self.stats16 is not None and other.stats16 is not None:
            self.stats16 = np.ascontiguousarray(np.concatenate([self.stats16, other.stats16]))
        else:
            self.stats16 = None

    def truncate(self, k):
        for k_ in ("points", "normals", "curvature", "omega_p", "omega_n", "gauss", "stats16"):
            if getattr(self, k_) is not None:
                setattr(self, k_, np.ascontiguousarray(getattr(self, k_)[:k]))


def _check_equal(cl, hc, stats=False):
    h = HostCloud(cl, stats)
    assert len(h.points) == len(hc.points)
    assert np.array_equal(_bits(h.points), _bits(hc.points))
    assert np.array_equal(_bits(h.normals[:, :3]), _bits(hc.normals[:, :3]))
    assert np.array_equal(_bits(h.curvature), _bits(hc.curvature))
    for a, b in ((h.omega_p, hc.omega_p), (h.omega_n, hc.omega_n)):
        for r, c in UPPER:
            assert np.array_equal(_bits(a[:, c * 4 + r]), _bits(b[:, c * 4 + r])), (r, c)
    g, go = h.gauss, hc.gauss
    flags, flags_o = O.gaussian_flags(g), O.gaussian_flags(go)
    assert np.array_equal(flags, flags_o)
    mom, inf = (flags & 1) != 0, (flags & 2) != 0
    assert np.array_equal(_bits(g[mom][:, 0:3]), _bits(go[mom][:, 0:3]))
    assert np.array_equal(_bits(g[mom][:, 4:13]), _bits(go[mom][:, 4:13]))
    assert np.array_equal(_bits(g[inf][:, 13:25]), _bits(go[inf][:, 13:25]))
    if stats:
        assert np.array_equal(_bits(h.stats16[:, [0, 1, 2, 4, 5, 6, 8, 9, 10, 12, 13, 14, 15]]),
                              _bits(hc.stats16[:, [0, 1, 2, 4, 5, 6, 8, 9, 10, 12, 13, 14, 15]]))


def _metres(raw):
    return np.where(raw > 0, np.float32(0.001) * raw.astype(np.float32), 0).astype(np.float32)


@pytest.mark.parametrize("rows,cols", [(120, 160), (480, 640)])
@pytest.mark.parametrize("baseline,alpha", [(0.075, 0.1), (0.2, 0.03)])
def test_gaussians_match_oracle(ctx, rows, cols, baseline, alpha):
    print("case gaussians", rows, cols, baseline, alpha, flush=True)
    dr, _, _ = synth.pair(21, rows, cols)
    ocam, dcam = cams(rows, cols)
    _, dsp = params(min_image_radius=3)
    cl = ctx.cloud_from_depth(dr, dcam, dsp)
    cl.compute_gaussians(baseline, alpha)
    g = cl.download_gaussians()
    K = np.array(ocam.K, np.float32)
    _, iKRt = O.projector_matrices(K, identity16())
    p_o, g_o, idx_o = O.unproject_gaussians(_metres(dr), K, iKRt, ocam.min_distance, ocam.max_distance, baseline,
                                            alpha)
    assert np.array_equal(cl.index_image(), idx_o)
    assert np.array_equal(_bits(g), _bits(g_o))


@pytest.mark.parametrize("rows,cols", [(120, 160), (480, 640)])
def test_transform_matches_oracle(ctx, rows, cols):
    print("case transform", rows, cols, flush=True)
    dr, _, _ = synth.pair(22, rows, cols)
    _, dcam = cams(rows, cols)
    _, dsp = params(min_image_radius=3)
    cl = ctx.cloud_from_depth(dr, dcam, dsp, keep_stats=True)
    cl.compute_gaussians()
    hc = HostCloud(cl, stats=True)
    T = synth.motion(5, 25.0, 0.4)
    cl.transform(T)
    O.cloud_transform(T, **hc.arrays())
    _check_equal(cl, hc, stats=True)
    # the identity changes nothing
    before = HostCloud(cl, stats=True)
    cl.transform(identity16())
    _check_equal(cl, before, stats=True)


def _frames(scene, poses, rows, cols):
    return [synth.render(scene, P, rows, cols, stream=10 + i) for i, P in enumerate(poses)]


@pytest.mark.parametrize("rows,cols,thr,stats", [(120, 160, 0.1, False), (480, 640, 0.1, False),
                                                (240, 320, 0.02, False), (240, 320, 0.1, True)])
def test_add_and_merge_match_oracle(ctx, rows, cols, thr, stats):
    """nicp_aligner's map loop (nicp/nicp_aligner.cpp:200-210): three views of one scene added
    at their poses, merged after each add from the newest camera. The second merge meets
    Gaussians whose information form is current from the first."""
    print("case add+merge", rows, cols, thr, stats, flush=True)
    ocam, dcam = cams(rows, cols)
    _, dsp = params(min_image_radius=3)
    poses = [identity16(), synth.motion(31, 6.0, 0.08), synth.motion(32, 10.0, 0.15)]
    depths = _frames(7, poses, rows, cols)
    clouds = [ctx.cloud_from_depth(d, dcam, dsp, keep_stats=stats) for d in depths]
    for c in clouds:
        c.compute_gaussians()
    hosts = [HostCloud(c, stats) for c in clouds]
    mp = capi.merger_params(distance_threshold=thr)
    world = clouds[0]
    hw = hosts[0]
    total_folded = 0
    for k in (1, 2):
        world.add(clouds[k], poses[k])
        O.cloud_transform(poses[k], **hosts[k].arrays())
        hw.append(hosts[k])
        _check_equal(world, hw, stats)
        _check_equal(clouds[k], hosts[k], stats)  # Cloud::add transforms its argument in place
        ns, col = world.merge(poses[k], dcam, rows, cols, mp, want_collapsed=True)
        kk, col_o = O.merge(hw.points, hw.gauss, poses[k], ocam, rows, cols, normals=hw.normals,
                            curvature=hw.curvature, stats16=hw.stats16, omega_p=hw.omega_p, omega_n=hw.omega_n,
                            distance_threshold=thr)
        hw.truncate(kk)
        assert np.array_equal(col, col_o)
        assert world.n == kk and ns == len(col) - kk
        n = len(col)
        folded = (col >= 0) & (col != np.arange(n))
        total_folded += int(folded.sum())
        _check_equal(world, hw, stats)
    # the views overlap: the merges fold many points, some owners take several
    assert total_folded > 0.2 * clouds[1].n
    flags = O.gaussian_flags(world.download_gaussians())
    assert (flags == 3).sum() > 0 and (flags == 1).sum() > 0


def test_merge_needs_gaussians_and_empty_cloud(ctx):
    rows, cols = 60, 80
    dr, _, _ = synth.pair(23, rows, cols)
    _, dcam = cams(rows, cols)
    _, dsp = params(min_image_radius=3)
    cl = ctx.cloud_from_depth(dr, dcam, dsp)
    with pytest.raises(capi.NicpError):
        cl.merge(identity16(), dcam, rows, cols)
    host = ctx.cloud_from_host(np.zeros((0, 4), np.float32))
    with pytest.raises(capi.NicpError):
        host.compute_gaussians()
    assert host.merge(identity16(), dcam, rows, cols) == 0
    # adding a cloud without Gaussians to one with them is refused; the other way drops them
    cl.compute_gaussians()
    other = ctx.cloud_from_depth(dr, dcam, dsp)
    with pytest.raises(capi.NicpError):
        cl.add(other, identity16())
    other.add(cl, identity16())
    with pytest.raises(capi.NicpError):
        other.download_gaussians()
    with pytest.raises(capi.NicpError):
        cl.add(cl, identity16())


def test_merged_map_aligns(ctx):
    """The merged map serves as the reference cloud of the next alignment (nicp_aligner.cpp:
    the map is projected into the next view and registered against it); the device alignment
    against the map agrees with the oracle's on the same clouds."""
    rows, cols = 240, 320
    ocam, dcam = cams(rows, cols)
    oap, dap = O.align_params(), capi.align_params()
    _, dsp = params(min_image_radius=3)
    poses = [identity16(), synth.motion(41, 2.0, 0.02), synth.motion(42, 2.0, 0.02)]
    depths = _frames(9, poses, rows, cols)
    world = ctx.cloud_from_depth(depths[0], dcam, dsp)
    world.compute_gaussians()
    c1 = ctx.cloud_from_depth(depths[1], dcam, dsp)
    c1.compute_gaussians()
    world.add(c1, poses[1])
    world.merge(poses[1], dcam, rows, cols)
    c2 = ctx.cloud_from_depth(depths[2], dcam, dsp)
    T, res = ctx.align(world, c2, dcam, rows, cols, poses[1], dap)
    hw, h2 = world.download(), c2.download()
    To, _ = O.align(O.cloud_from_arrays(*hw), O.cloud_from_arrays(*h2), rows, cols, ocam, poses[1], oap, 1,
                    fp64_solve=True)
    rot, tr = synth.pose_error(T, To)
    assert np.radians(rot) < 1e-4 and tr < 1e-4, (rot, tr)
    rot, tr = synth.pose_error(T, poses[2])
    assert rot < 1.0 and tr < 0.02, (rot, tr)
