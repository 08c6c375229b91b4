"""The C++ drop-in front end (include/nicp/*.h) driven end to end by the
nicp_simple_aligner example: 16-bit PGM depth list + configuration in, a
trajectory out, compared with the poses the frames were rendered from."""
import os
import subprocess
import sys

import numpy as np
import pytest

from nicp_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "nicp_amd", "nicp_simple_aligner")

CONF = """// Projector
fx 525.0
fy 525.0
cx 319.5
cy 239.5
minDistance 0.5
maxDistance 4.5
// Stats calculator and information matrix calculators
minImageRadius 10
maxImageRadius 30
minPoints 50
curvatureThreshold 0.2
worldRadius 0.1
informationMatrixCurvatureThreshold 0.02
// Correspondence finder
inlierDistanceThreshold 1.0
inlierNormalAngularThreshold 0.95
inlierCurvatureRatioThreshold 1.3
flatCurvatureThreshold 0.02
// Linearizer
inlierMaxChi2 9000
robustKernel 1
// Aligner
outerIterations 10
innerIterations 1
minInliers 100
translationalMinEigenRatio 50
rotationalMinEigenRatio 50
"""


def write_pgm16(path, img):
    h, w = img.shape
    with open(path, "wb") as f:
        f.write(b"P5\n%d %d\n65535\n" % (w, h))
        f.write(img.astype(">u2").tobytes())


def test_app_builds_on_cpu():
    assert os.path.exists(APP)
    out = subprocess.run([APP], capture_output=True, text=True)
    assert "USAGE" in out.stdout


@pytest.mark.gpu
def test_simple_aligner_trajectory(tmp_path):
    print("case simple_aligner_trajectory", flush=True)
    rows, cols = 480, 640
    scene = 77
    poses = [np.eye(4)]
    for k in range(1, 4):
        step = synth.motion(500 + k, 3.0, 0.03).reshape(4, 4, order="F").astype(np.float64)
        poses.append(poses[-1] @ step)
    lines = []
    for k, P in enumerate(poses):
        d = synth.render(scene, P.astype(np.float32).flatten(order="F"), rows, cols, stream=100 + k)
        f = tmp_path / ("d%d.pgm" % k)
        write_pgm16(str(f), d)
        lines.append("%d.0 %s" % (k, f))
    (tmp_path / "list.txt").write_text("\n".join(lines) + "\n")
    (tmp_path / "c.conf").write_text(CONF)
    out = subprocess.run([APP, str(tmp_path / "c.conf"), str(tmp_path / "list.txt"), str(tmp_path / "odom.txt")],
                         capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    traj = [list(map(float, l.split()[1:])) for l in (tmp_path / "odom.txt").read_text().split("\n") if l.strip()]
    assert len(traj) == len(poses)
    from scipy.spatial.transform import Rotation
    for k, (t, P) in enumerate(zip(traj, poses)):
        R = Rotation.from_quat(t[3:7]).as_matrix()
        ang = np.degrees(np.arccos(np.clip((np.trace(R.T @ P[:3, :3]) - 1) / 2, -1, 1)))
        dist = np.linalg.norm(np.array(t[:3]) - P[:3, 3])
        print("frame %d: %.4f deg %.5f m" % (k, ang, dist), flush=True)
        assert ang < 0.1 * max(k, 1) + 1e-6 and dist < 0.003 * max(k, 1) + 1e-9
    _check_saved_frames(tmp_path, traj)


def _check_saved_frames(tmp_path, traj):
    """Each frame's cloud saved as <depth file>.nicp (nicp_simple_aligner.cpp:175, Cloud::save
    at cloud.cpp:99-150): header with the point count and step 1, the global transform as its
    6-vector (t, qx qy qz with qw >= 0), one 128-byte binary record per point."""
    for k, t in enumerate(traj):
        raw = (tmp_path / ("d%d.pgm.nicp" % k)).read_bytes()
        head, t2v, body = raw.split(b"\n", 2)
        tag, n, step = head.split()
        assert tag == b"NICPCLOUD" and int(step) == 1 and int(n) > 100000
        assert len(body) == 128 * int(n)
        q = np.array(t[3:7])
        q = q if q[3] >= 0 else -q
        np.testing.assert_allclose([float(v) for v in t2v.split()], list(t[:3]) + list(q[:3]), atol=2e-6)


APP2 = os.path.join(ROOT, "nicp_amd", "nicp_aligner")


def test_scene_app_builds_on_cpu():
    assert os.path.exists(APP2)
    out = subprocess.run([APP2], capture_output=True, text=True)
    assert "USAGE" in out.stdout


@pytest.mark.gpu
def test_scene_aligner_trajectory(tmp_path):
    """nicp_aligner (nicp/nicp/nicp_aligner.cpp): registration against the merged local scene,
    scenes restarted every chunkStep frames and saved, per-frame clouds saved."""
    print("case scene_aligner_trajectory", flush=True)
    rows, cols = 480, 640
    scene = 78
    poses = [np.eye(4)]
    for k in range(1, 6):
        step = synth.motion(600 + k, 3.0, 0.03).reshape(4, 4, order="F").astype(np.float64)
        poses.append(poses[-1] @ step)
    lines = []
    for k, P in enumerate(poses):
        d = synth.render(scene, P.astype(np.float32).flatten(order="F"), rows, cols, stream=200 + k)
        f = tmp_path / ("d%d.pgm" % k)
        write_pgm16(str(f), d)
        lines.append("%d.0 %s" % (k, f))
    (tmp_path / "list.txt").write_text("\n".join(lines) + "\n")
    (tmp_path / "c.conf").write_text(CONF + "chunkStep 3\ndistanceThreshold 0.1\n")
    out = subprocess.run([APP2, str(tmp_path / "c.conf"), str(tmp_path / "list.txt"), str(tmp_path / "odom.txt")],
                         capture_output=True, text=True, timeout=180, cwd=str(tmp_path))
    print(out.stdout[-3000:], out.stderr[-2000:], flush=True)
    assert out.returncode == 0
    traj = [list(map(float, l.split()[1:])) for l in (tmp_path / "odom.txt").read_text().split("\n") if l.strip()]
    assert len(traj) == len(poses)
    from scipy.spatial.transform import Rotation
    for k, (t, P) in enumerate(zip(traj, poses)):
        R = Rotation.from_quat(t[3:7]).as_matrix()
        ang = np.degrees(np.arccos(np.clip((np.trace(R.T @ P[:3, :3]) - 1) / 2, -1, 1)))
        dist = np.linalg.norm(np.array(t[:3]) - P[:3, 3])
        print("frame %d: %.4f deg %.5f m" % (k, ang, dist), flush=True)
        assert ang < 0.1 * max(k, 1) + 1e-6 and dist < 0.003 * max(k, 1) + 1e-9
    # merging keeps the scene near one frame's size: suppressed points are reported
    sizes = [int(l.split()[-1]) for l in out.stdout.splitlines() if l.startswith("scene points")]
    assert len(sizes) == len(poses)
    assert max(sizes) < 2.2 * sizes[0]
    assert "number of suppressed points" in out.stdout
    scenes = sorted(p.name for p in tmp_path.glob("scene-*.nicp"))
    assert scenes == ["scene-001.nicp", "scene-004.nicp", "scene-005.nicp"], scenes
    for k in range(len(poses)):
        assert (tmp_path / ("d%d.pgm.nicp" % k)).stat().st_size > 128 * 100000


@pytest.mark.gpu
def test_front_end_sensor_offset(tmp_path):
    """DepthImageConverterIntegralImage::compute(cloud, depth, offset) and
    Aligner::setSensorOffset through the C++ front end give, bit for bit, the C ABI's clouds moved
    by the offset (nicp_cloud_transform) aligned with nicp_align_offsets."""
    from nicp_amd import capi
    from tests.helpers import aparams, cams, identity16, offset16, params
    print("case front end sensor offset", flush=True)
    rows, cols = 480, 640
    dr, dc, _ = synth.pair(11, rows, cols, max_rot_deg=6.0, max_trans_m=0.06)
    write_pgm16(str(tmp_path / "a.pgm"), dr)
    write_pgm16(str(tmp_path / "b.pgm"), dc)
    S = offset16(20.0, (0.2, 1, -0.3), (0.1, -0.3, 0.8))
    exe = str(tmp_path / "off")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), "-o", exe,
                           os.path.join(ROOT, "tests", "cpp", "offset_front.cpp"), "-L", os.path.join(ROOT, "nicp_amd"),
                           "-lnicp_amd", "-Wl,-rpath," + os.path.join(ROOT, "nicp_amd")])
    out = subprocess.run([exe, str(tmp_path / "a.pgm"), str(tmp_path / "b.pgm")] + ["%.9g" % v for v in S],
                         capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    lines = dict((l.split()[0], l.split()[1:]) for l in out.stdout.splitlines() if l.strip())
    Tf = np.array([float(v) for v in lines["T"]], np.float32)
    ctx = capi.Context(0)
    _, dsp = params()
    _, dap = aparams()
    _, dcam = cams(rows, cols)
    S32 = np.array(["%.9g" % v for v in S], np.float32)
    cr, cc = ctx.cloud_from_depth(dr, dcam, dsp), ctx.cloud_from_depth(dc, dcam, dsp)
    cr.transform(S32)
    cc.transform(S32)
    T, res = ctx.align_offsets(cr, cc, dcam, rows, cols, identity16(), S32, S32, dap)
    np.testing.assert_array_equal(Tf.view(np.uint32), T.view(np.uint32))
    assert int(lines["inliers"][0]) == res.inliers
    ctx.close()
