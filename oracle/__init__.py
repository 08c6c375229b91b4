"""ctypes bindings of the CPU oracle — TEST INFRASTRUCTURE ONLY.

The oracle (``oracle/nicp_oracle.cpp``) restates the arithmetic of yorsh87/nicp
on the registration path. Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` may import this module, and only as the
checker or as the timed CPU reference. The product package ``nicp_amd`` never
imports it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libnicp_oracle.so")


def build():
    import subprocess
    subprocess.check_call(["sh", os.path.join(_HERE, "build_oracle.sh")])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run oracle/build_oracle.sh")
        _lib = C.CDLL(LIB_PATH)
    return _lib


class Camera(C.Structure):
    _fields_ = [("K", C.c_float * 9), ("min_distance", C.c_float), ("max_distance", C.c_float),
                ("model", C.c_int), ("horizontal_fov", C.c_float), ("vertical_fov", C.c_float),
                ("horizontal_center", C.c_float), ("vertical_center", C.c_float)]


class StatsParams(C.Structure):
    _fields_ = [("world_radius", C.c_float), ("min_image_radius", C.c_int), ("max_image_radius", C.c_int),
                ("min_points", C.c_int), ("curvature_threshold", C.c_float),
                ("point_flat", C.c_float * 3), ("point_nonflat", C.c_float * 3),
                ("normal_flat", C.c_float * 3), ("normal_nonflat", C.c_float * 3),
                ("point_curvature_threshold", C.c_float), ("normal_curvature_threshold", C.c_float)]


class AlignParams(C.Structure):
    _fields_ = [("inlier_distance_threshold", C.c_float), ("inlier_normal_angular_threshold", C.c_float),
                ("flat_curvature_threshold", C.c_float), ("inlier_curvature_ratio_threshold", C.c_float),
                ("row_search_region", C.c_int), ("col_search_region", C.c_int), ("demoted_to_gicp", C.c_int),
                ("inlier_max_chi2", C.c_float), ("robust_kernel", C.c_int), ("normal_omega_scale", C.c_float),
                ("outer_iterations", C.c_int), ("inner_iterations", C.c_int), ("lambda_", C.c_float),
                ("translational_min_eigen_ratio", C.c_float), ("rotational_min_eigen_ratio", C.c_float)]


class CloudC(C.Structure):
    _fields_ = [("n", C.c_int), ("points", C.c_void_p), ("normals", C.c_void_p), ("curvature", C.c_void_p),
                ("omega_p", C.c_void_p), ("omega_n", C.c_void_p)]


class AlignResult(C.Structure):
    _fields_ = [("error", C.c_float), ("inliers", C.c_int), ("num_correspondences", C.c_int),
                ("H", C.c_float * 36), ("b", C.c_float * 6), ("mean", C.c_double * 6),
                ("omega", C.c_double * 36), ("translational_ratio", C.c_double),
                ("rotational_ratio", C.c_double), ("solution_valid", C.c_int)]


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def camera(fx, fy, cx, cy, min_d=0.5, max_d=4.5):
    K = np.array([[fx, 0, cx], [0, fy, cy], [0, 0, 1]], dtype=np.float32)
    c = Camera()
    c.K[:] = K.flatten(order="F").tolist()
    c.min_distance = min_d
    c.max_distance = max_d
    return c


def spherical_camera(rows, cols, horizontal_fov=float(np.pi), vertical_fov=float(np.pi / 4), min_d=0.01, max_d=6.0,
                     horizontal_center=None, vertical_center=None):
    """SphericalPointProjector: half fields of view (defaults pi, pi/4, sphericalpointprojector.cpp:9-16)
    and the image centre (default the middle of the image)."""
    c = Camera()
    c.K[:] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    c.min_distance, c.max_distance = min_d, max_d
    c.model = 1
    c.horizontal_fov, c.vertical_fov = horizontal_fov, vertical_fov
    c.horizontal_center = cols / 2.0 if horizontal_center is None else horizontal_center
    c.vertical_center = rows / 2.0 if vertical_center is None else vertical_center
    return c


def stats_params(world_radius=0.1, min_image_radius=10, max_image_radius=30, min_points=50,
                 curvature_threshold=0.2, infomat_curvature_threshold=0.02):
    """Defaults of nicp/nicp_conf/nicp_aligner_1_1.conf:22-27."""
    s = StatsParams()
    s.world_radius = world_radius
    s.min_image_radius = min_image_radius
    s.max_image_radius = max_image_radius
    s.min_points = min_points
    s.curvature_threshold = curvature_threshold
    s.point_flat[:] = [1000.0, 0.001, 0.001]
    s.point_nonflat[:] = [1.0, 0.1, 0.1]
    s.normal_flat[:] = [1000.0, 0.001, 0.001]
    s.normal_nonflat[:] = [0.1, 1.0, 1.0]
    s.point_curvature_threshold = infomat_curvature_threshold
    s.normal_curvature_threshold = infomat_curvature_threshold
    return s


def align_params(outer_iterations=10, inner_iterations=1, **kw):
    """Defaults of nicp/nicp_conf/nicp_aligner_1_1.conf:30-44 and N/aligner.cpp:21-41."""
    a = AlignParams()
    a.inlier_distance_threshold = 1.0
    a.inlier_normal_angular_threshold = 0.95
    a.flat_curvature_threshold = 0.02
    a.inlier_curvature_ratio_threshold = 1.3
    a.row_search_region = 0
    a.col_search_region = 0
    a.demoted_to_gicp = 0
    a.inlier_max_chi2 = 9000.0
    a.robust_kernel = 1
    a.normal_omega_scale = 1.0
    a.outer_iterations = outer_iterations
    a.inner_iterations = inner_iterations
    a.lambda_ = 1e3
    a.translational_min_eigen_ratio = 50.0
    a.rotational_min_eigen_ratio = 50.0
    for k, v in kw.items():
        setattr(a, k, v)
    return a


class Frame:
    """Outputs of the oracle's DepthImageConverterIntegralImage::compute."""

    def __init__(self, raw, depth_scale, cam, sp):
        raw = np.ascontiguousarray(raw, dtype=np.uint16)
        rows, cols = raw.shape
        P = rows * cols
        self.rows, self.cols = rows, cols
        self.index_image = np.empty(P, np.int32)
        pts = np.empty((P, 4), np.float32)
        nrm = np.empty((P, 4), np.float32)
        curv = np.empty(P, np.float32)
        evec = np.empty((P, 9), np.float32)
        eval_ = np.empty((P, 3), np.float32)
        mean = np.empty((P, 4), np.float32)
        npts = np.empty(P, np.int32)
        op = np.empty((P, 16), np.float32)
        on = np.empty((P, 16), np.float32)
        f = lib().orc_frame_from_depth
        f.restype = C.c_int
        n = f(_p(raw), rows, cols, C.c_float(depth_scale), C.byref(cam), C.byref(sp), _p(self.index_image),
              _p(pts), _p(nrm), _p(curv), _p(evec), _p(eval_), _p(mean), _p(npts), _p(op), _p(on))
        self.n = n
        self.index_image = self.index_image.reshape(rows, cols)
        self.points, self.normals, self.curvature = pts[:n].copy(), nrm[:n].copy(), curv[:n].copy()
        self.eigvecs, self.eigvals, self.means, self.npts = evec[:n].copy(), eval_[:n].copy(), mean[:n].copy(), npts[:n].copy()
        self.omega_p, self.omega_n = op[:n].copy(), on[:n].copy()

    def cloud(self):
        c = CloudC()
        c.n = self.n
        c.points, c.normals, c.curvature = _p(self.points), _p(self.normals), _p(self.curvature)
        c.omega_p, c.omega_n = _p(self.omega_p), _p(self.omega_n)
        return c


def cloud_from_arrays(points, normals, curvature, omega_p, omega_n):
    c = CloudC()
    c.n = len(points)
    c._keep = (points, normals, curvature, omega_p, omega_n)
    c.points, c.normals, c.curvature = _p(points), _p(normals), _p(curvature)
    c.omega_p, c.omega_n = _p(omega_p), _p(omega_n)
    return c


def projector_matrices(K9, transform16):
    KRt = np.empty(16, np.float32)
    iKRt = np.empty(16, np.float32)
    lib().orc_projector_matrices(_p(np.ascontiguousarray(K9, np.float32)), _p(np.ascontiguousarray(transform16, np.float32)),
                                 _p(KRt), _p(iKRt))
    return KRt, iKRt


def project(points, KRt, rows, cols, min_d, max_d):
    idx = np.empty(rows * cols, np.int32)
    dep = np.empty(rows * cols, np.float32)
    points = np.ascontiguousarray(points, np.float32)
    lib().orc_project(_p(points), len(points), _p(np.ascontiguousarray(KRt, np.float32)), rows, cols,
                      C.c_float(min_d), C.c_float(max_d), _p(idx), _p(dep))
    return idx.reshape(rows, cols), dep.reshape(rows, cols)


def correspondences(ref_cloud, cur_cloud, ref_index, cur_index, T16, ap, num_threads=1):
    rows, cols = ref_index.shape
    pairs = np.empty(2 * rows * cols, np.int32)
    f = lib().orc_correspondences
    f.restype = C.c_int
    n = f(C.byref(ref_cloud), C.byref(cur_cloud), _p(np.ascontiguousarray(ref_index, np.int32)),
          _p(np.ascontiguousarray(cur_index, np.int32)), rows, cols, _p(np.ascontiguousarray(T16, np.float32)),
          C.byref(ap), num_threads, _p(pairs))
    return pairs[:2 * n].reshape(n, 2).copy()


def linearize(ref_cloud, cur_cloud, pairs, T16, ap, num_threads=1):
    H = np.empty(36, np.float32)
    b = np.empty(6, np.float32)
    e = C.c_float()
    i = C.c_int()
    pairs = np.ascontiguousarray(pairs, np.int32)
    lib().orc_linearize(C.byref(ref_cloud), C.byref(cur_cloud), _p(pairs), len(pairs),
                        _p(np.ascontiguousarray(T16, np.float32)), C.byref(ap), num_threads, _p(H), _p(b),
                        C.byref(e), C.byref(i))
    return H.reshape(6, 6, order="F"), b, e.value, i.value


def linearize_f64(ref_cloud, cur_cloud, pairs, T16, ap):
    H = np.empty(36, np.float64)
    b = np.empty(6, np.float64)
    e = C.c_double()
    i = C.c_int()
    pairs = np.ascontiguousarray(pairs, np.int32)
    lib().orc_linearize_f64(C.byref(ref_cloud), C.byref(cur_cloud), _p(pairs), len(pairs),
                            _p(np.ascontiguousarray(T16, np.float32)), C.byref(ap), _p(H), _p(b),
                            C.byref(e), C.byref(i))
    return H.reshape(6, 6, order="F"), b, e.value, i.value


def align(ref_cloud, cur_cloud, rows, cols, cam, init16, ap, num_threads=1, fp64_solve=False):
    """AlignerProjective::align; fp64_solve runs H, b and the LDLT in FP64 (the device's design)."""
    T = np.empty(16, np.float32)
    res = AlignResult()
    lib().orc_align_mode(C.byref(ref_cloud), C.byref(cur_cloud), rows, cols, C.byref(cam),
                         _p(np.ascontiguousarray(init16, np.float32)), C.byref(ap), num_threads, int(fp64_solve),
                         _p(T), C.byref(res))
    return T, res


def align_offsets(ref_cloud, cur_cloud, rows, cols, cam, init16, ref_offset16, cur_offset16, ap, num_threads=1,
                  fp64_solve=False):
    """AlignerProjective::align with the aligner's reference and current sensor offsets."""
    T = np.empty(16, np.float32)
    res = AlignResult()
    ro = None if ref_offset16 is None else np.ascontiguousarray(ref_offset16, np.float32)
    co = None if cur_offset16 is None else np.ascontiguousarray(cur_offset16, np.float32)
    lib().orc_align_offsets(C.byref(ref_cloud), C.byref(cur_cloud), rows, cols, C.byref(cam),
                            _p(np.ascontiguousarray(init16, np.float32)), _p(ro) if ro is not None else None,
                            _p(co) if co is not None else None, C.byref(ap), num_threads, int(fp64_solve), _p(T),
                            C.byref(res))
    return T, res


def stats_f64(raw, depth_scale, cam, sp, frame):
    n = frame.n
    nrm = np.empty((n, 3), np.float64)
    curv = np.empty(n, np.float64)
    valid = np.empty(n, np.int32)
    raw = np.ascontiguousarray(raw, np.uint16)
    lib().orc_stats_f64_bruteforce(_p(raw), frame.rows, frame.cols, C.c_float(depth_scale), C.byref(cam), C.byref(sp),
                                   _p(np.ascontiguousarray(frame.index_image.ravel(), np.int32)),
                                   _p(frame.points), n, _p(nrm), _p(curv), _p(valid))
    return nrm, curv, valid.astype(bool)


def eigen3_direct(cov):
    cov = np.ascontiguousarray(np.asarray(cov, np.float32).flatten(order="F"))
    ev = np.empty(3, np.float32)
    V = np.empty(9, np.float32)
    lib().orc_eigen3_direct(_p(cov), _p(ev), _p(V))
    return ev, V.reshape(3, 3, order="F")


def get_region(ii, rows, cols, xmin, xmax, ymin, ymax):
    out = np.empty(10, np.float32)
    lib().orc_get_region(_p(ii), int(rows), int(cols), int(xmin), int(xmax), int(ymin), int(ymax), _p(out))
    return out


def integral_image(index_image, points):
    rows, cols = index_image.shape
    ii = np.empty(rows * cols * 10, np.float32)
    lib().orc_integral_image(_p(np.ascontiguousarray(index_image, np.int32)), rows, cols,
                             _p(np.ascontiguousarray(points, np.float32)), _p(ii))
    return ii


def unproject(depth, iKRt, min_d, max_d):
    depth = np.ascontiguousarray(depth, np.float32)
    rows, cols = depth.shape
    pts = np.empty((rows * cols, 4), np.float32)
    idx = np.empty(rows * cols, np.int32)
    f = lib().orc_unproject
    f.restype = C.c_int
    n = f(_p(depth), rows, cols, _p(np.ascontiguousarray(iKRt, np.float32)), C.c_float(min_d), C.c_float(max_d),
          _p(pts), _p(idx))
    return pts[:n].copy(), idx.reshape(rows, cols)


def intervals(depth, K9, min_d, max_d, world_radius):
    depth = np.ascontiguousarray(depth, np.float32)
    rows, cols = depth.shape
    out = np.empty(rows * cols * 2, np.int32)
    lib().orc_project_intervals(_p(depth), rows, cols, _p(np.ascontiguousarray(K9, np.float32)), C.c_float(min_d),
                                C.c_float(max_d), C.c_float(world_radius), _p(out))
    return out.reshape(rows, cols, 2)


def set_num_threads(n):
    """OpenMP threads of the conversion's loops (the aligner's count is an argument of align)."""
    lib().orc_set_num_threads(int(n))


def max_threads():
    f = lib().orc_get_max_threads
    f.restype = C.c_int
    return int(f())


def v2t(x):
    T = np.empty(16, np.float32)
    lib().orc_v2t(_p(np.ascontiguousarray(x, np.float32)), _p(T))
    return T


def t2v(T):
    x = np.empty(6, np.float32)
    lib().orc_t2v(_p(np.ascontiguousarray(T, np.float32)), _p(x))
    return x


def ldlt_solve(H, rhs, dtype=np.float32):
    Hc = np.ascontiguousarray(np.asarray(H, dtype).flatten(order="F"))
    r = np.ascontiguousarray(rhs, dtype)
    x = np.empty(6, dtype)
    f = lib().orc_ldlt_solve_f32 if dtype == np.float32 else lib().orc_ldlt_solve_f64
    f.restype = C.c_int
    ok = f(_p(Hc), _p(r), _p(x))
    return x, ok


def compute_statistics(H, T16, ap):
    res = AlignResult()
    Hc = np.ascontiguousarray(np.asarray(H, np.float64).flatten(order="F"))
    lib().orc_compute_statistics(_p(Hc), _p(np.ascontiguousarray(T16, np.float32)), C.byref(ap), C.byref(res))
    return res


def nn_correspondences(ref_cloud, cur_cloud, T16, ap, normal_scaling=1.0):
    """Returns (pairs (k, 2), nearest neighbour per reference point)."""
    pairs = np.empty(2 * max(ref_cloud.n, 1), np.int32)
    nn = np.empty(max(ref_cloud.n, 1), np.int32)
    f = lib().orc_nn_correspondences
    f.restype = C.c_int
    k = f(C.byref(ref_cloud), C.byref(cur_cloud), _p(np.ascontiguousarray(T16, np.float32)), C.byref(ap),
          C.c_float(normal_scaling), _p(pairs), _p(nn))
    return pairs[:2 * k].reshape(k, 2).copy(), nn[:ref_cloud.n].copy()


def align_nn(ref_cloud, cur_cloud, init16, ap, normal_scaling=1.0, fp64_solve=False):
    T = np.empty(16, np.float32)
    res = AlignResult()
    lib().orc_align_nn_mode(C.byref(ref_cloud), C.byref(cur_cloud), _p(np.ascontiguousarray(init16, np.float32)),
                            C.byref(ap), C.c_float(normal_scaling), int(fp64_solve), _p(T), C.byref(res))
    return T, res


def depth_scale(depth_m, step, max_depth_cov=0.01):
    d = np.ascontiguousarray(depth_m, np.float32)
    rows, cols = d.shape
    out = np.empty((rows // step, cols // step), np.float32)
    lib().orc_depth_scale(_p(d), rows, cols, step, C.c_float(max_depth_cov), _p(out))
    return out


#: columns of a Gaussian record (orc_gaussian / nicp_gaussian) in a (n, 28) float32 array;
#: column 3 holds the int32 flags (bit 0 moments current, bit 1 information form current)
G_MEAN, G_FLAGS, G_COV, G_INFO, G_IVEC = slice(0, 3), 3, slice(4, 13), slice(13, 22), slice(22, 25)


def gaussian_flags(g):
    return np.ascontiguousarray(g[:, G_FLAGS]).view(np.int32)


def unproject_gaussians(depth_m, K9, iKRt, min_d, max_d, baseline=0.075, alpha=0.1):
    """PinholePointProjector::unProject with Gaussians: (points, gaussians (n, 28), index image)."""
    depth_m = np.ascontiguousarray(depth_m, np.float32)
    rows, cols = depth_m.shape
    pts = np.zeros((rows * cols, 4), np.float32)
    g = np.zeros((rows * cols, 28), np.float32)
    idx = np.empty(rows * cols, np.int32)
    f = lib().orc_unproject_gaussians
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float,
                  C.c_void_p, C.c_void_p, C.c_void_p]
    n = f(_p(depth_m), rows, cols, _p(np.ascontiguousarray(K9, np.float32)), _p(np.ascontiguousarray(iKRt, np.float32)),
          min_d, max_d, baseline, alpha, _p(pts), _p(g), _p(idx))
    return pts[:n].copy(), g[:n].copy(), idx.reshape(rows, cols)


def _f32(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous, "arrays are changed in place: float32, contiguous"
    return a


def cloud_transform(T16, points=None, normals=None, stats16=None, omega_p=None, omega_n=None, gaussians=None):
    """Cloud::transformInPlace on float32 arrays, in place (stats16 / omegas: (n, 16) column-major)."""
    arrs = [points, normals, stats16, omega_p, omega_n, gaussians]
    n = next(len(a) for a in arrs if a is not None)
    f = lib().orc_cloud_transform
    f.argtypes = [C.c_int] + [C.c_void_p] * 7
    f(n, _p(np.ascontiguousarray(T16, np.float32)), *[_p(_f32(a)) for a in arrs])


def merge(points, gaussians, T16, cam, rows, cols, normals=None, curvature=None, stats16=None, omega_p=None,
          omega_n=None, distance_threshold=0.1, max_point_depth=10.0):
    """Merger::merge in place; returns (new size, per-point decisions)."""
    n = len(points)
    collapsed = np.empty(n, np.int32)
    f = lib().orc_merge
    f.restype = C.c_int
    f.argtypes = [C.c_int] + [C.c_void_p] * 9 + [C.c_int, C.c_int, C.c_float, C.c_float, C.c_void_p]
    k = f(n, _p(_f32(points)), _p(_f32(normals)), _p(_f32(curvature)), _p(_f32(stats16)), _p(_f32(omega_p)),
          _p(_f32(omega_n)), _p(_f32(gaussians)), _p(np.ascontiguousarray(T16, np.float32)), C.byref(cam), rows, cols,
          distance_threshold, max_point_depth, _p(collapsed))
    return k, collapsed
