"""ctypes bindings of the nicp_amd C ABI (``include/nicp_amd.h``).

The library ``libnicp_amd.so`` is the gfx950 build of ``csrc/``. Every call
goes to the device; there is no CPU fallback: a missing library or a failing
HIP call raises :class:`NicpError`.
"""
import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnicp_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "nicp_amd.h")

NICP_OK = 0


class NicpError(RuntimeError):
    pass


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


class MergerParams(C.Structure):
    _fields_ = [("distance_threshold", C.c_float), ("normal_threshold", C.c_float), ("max_point_depth", C.c_float)]


def merger_params(distance_threshold=0.1, normal_threshold=float(np.cos(np.radians(10.0))), max_point_depth=10.0):
    """Merger defaults (nicp/nicp/merger.cpp:5-14)."""
    m = MergerParams()
    m.distance_threshold, m.normal_threshold, m.max_point_depth = distance_threshold, normal_threshold, max_point_depth
    return m


class AlignResult(C.Structure):
    _fields_ = [("error", C.c_float), ("inliers", C.c_int), ("num_correspondences", C.c_int),
                ("H", C.c_double * 36), ("b", C.c_double * 6), ("mean", C.c_double * 6),
                ("omega", C.c_double * 36), ("translational_ratio", C.c_double),
                ("rotational_ratio", C.c_double), ("solution_valid", C.c_int), ("total_time_ms", C.c_float),
                ("overlap_pixels", C.c_int)]


#: every entry point declared in include/nicp_amd.h
EXPORTS = [
    "nicp_last_error", "nicp_ctx_create", "nicp_ctx_destroy", "nicp_ctx_synchronize", "nicp_ctx_stream",
    "nicp_cloud_from_depth_u16", "nicp_cloud_from_host", "nicp_cloud_destroy", "nicp_cloud_size",
    "nicp_cloud_download", "nicp_cloud_download_stats", "nicp_cloud_index_image", "nicp_project",
    "nicp_correspondences", "nicp_linearize", "nicp_align", "nicp_align_offsets", "nicp_ctx_set_fp64_eigen_gap", "nicp_last_correspondences", "nicp_batch_create",
    "nicp_batch_destroy", "nicp_batch_upload_depth", "nicp_batch_convert", "nicp_batch_align",
    "nicp_batch_frame_cloud", "nicp_kernel_count", "nicp_kernel_name", "nicp_ctx_profile",
    "nicp_ctx_profile_read", "nicp_nn_correspondences", "nicp_align_nn", "nicp_cloud_from_depth_f32",
    "nicp_depth_scale", "nicp_batch_render_synthetic", "nicp_initnet_num_params", "nicp_initnet_create",
    "nicp_initnet_destroy", "nicp_initnet_get_params", "nicp_initnet_set_params", "nicp_initnet_set_loss_weights",
    "nicp_initnet_forward", "nicp_initnet_train_step", "nicp_initnet_read", "nicp_batch_net_input",
    "nicp_batch_align_learned", "nicp_device_alloc", "nicp_device_free", "nicp_device_copy",
    "nicp_initnet_train_step_rows", "nicp_initnet_set_input_mode", "nicp_cloud_gaussians", "nicp_cloud_set_gaussians",
    "nicp_cloud_download_gaussians", "nicp_cloud_transform", "nicp_cloud_add", "nicp_merge",
    "nicp_batch_set_autotune", "nicp_batch_autotune_read", "nicp_align_nn_batch", "nicp_solution_statistics",
    "nicp_align_noalloc", "nicp_batch_set_projection", "nicp_batch_set_bin_capacity", "nicp_batch_probe",
    "nicp_batch_upload_depth_device", "nicp_batch_download_depth_device", "nicp_batch_num_pixels",
]

_lib = None
vp = C.c_void_p


def lib():
    """Loads libnicp_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NicpError("libnicp_amd.so is not built: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.nicp_last_error.restype = C.c_char_p
        L.nicp_ctx_stream.restype = vp
        L.nicp_kernel_name.restype = C.c_char_p
        L.nicp_kernel_name.argtypes = [C.c_int]
        L.nicp_kernel_count.restype = C.c_int
        sig = {
            "nicp_ctx_create": [C.c_int, vp], "nicp_ctx_destroy": [vp], "nicp_ctx_synchronize": [vp],
            "nicp_ctx_stream": [vp], "nicp_ctx_set_fp64_eigen_gap": [vp, C.c_float],
            "nicp_cloud_from_depth_u16": [vp, vp, C.c_int, C.c_int, C.c_float, vp, vp, C.c_int, vp],
            "nicp_cloud_from_host": [vp, C.c_int, vp, vp, vp, vp, vp, vp],
            "nicp_cloud_destroy": [vp], "nicp_cloud_size": [vp, vp],
            "nicp_cloud_download": [vp, vp, vp, vp, vp, vp],
            "nicp_cloud_download_stats": [vp, vp, vp, vp, vp],
            "nicp_cloud_index_image": [vp, vp],
            "nicp_project": [vp, vp, vp, vp, C.c_int, C.c_int, vp, vp],
            "nicp_correspondences": [vp, vp, vp, vp, vp, C.c_int, C.c_int, vp, vp, vp, vp],
            "nicp_linearize": [vp, vp, vp, vp, C.c_int, vp, vp, vp, vp, vp, vp],
            "nicp_align": [vp, vp, vp, vp, C.c_int, C.c_int, vp, vp, vp, vp],
            "nicp_last_correspondences": [vp, vp, C.c_int, vp],
            "nicp_solution_statistics": [vp, vp, vp, vp, vp],
            "nicp_align_noalloc": [vp, vp, vp, vp, C.c_int, C.c_int, vp, vp, vp, vp],
            "nicp_batch_set_projection": [vp, C.c_int], "nicp_batch_set_bin_capacity": [vp, C.c_int],
            "nicp_batch_probe": [vp, C.c_int, vp, vp, vp, vp, C.c_int, C.c_int, vp],
            "nicp_align_offsets": [vp, vp, vp, vp, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp],
            "nicp_batch_create": [vp, C.c_int, vp, vp, vp, vp, vp],
            "nicp_batch_destroy": [vp], "nicp_batch_upload_depth": [vp, C.c_int, vp, C.c_float],
            "nicp_batch_upload_depth_device": [vp, vp, C.c_float], "nicp_batch_download_depth_device": [vp, vp],
            "nicp_batch_num_pixels": [vp, vp],
            "nicp_batch_convert": [vp],
            "nicp_batch_align": [vp, C.c_int, vp, vp, vp, vp, C.c_int, vp, vp],
            "nicp_batch_frame_cloud": [vp, C.c_int, vp],
            "nicp_ctx_profile": [vp, C.c_int], "nicp_ctx_profile_read": [vp, vp, vp],
            "nicp_nn_correspondences": [vp, vp, vp, vp, vp, C.c_float, C.c_float, vp, vp, vp],
            "nicp_align_nn": [vp, vp, vp, vp, vp, C.c_float, C.c_float, vp, vp],
            "nicp_align_nn_batch": [vp, C.c_int, vp, vp, vp, vp, C.c_float, C.c_float, vp, vp],
            "nicp_cloud_from_depth_f32": [vp, vp, C.c_int, C.c_int, vp, vp, C.c_int, vp],
            "nicp_depth_scale": [vp, vp, vp, C.c_int, C.c_int, C.c_float, C.c_int, C.c_float, vp],
            "nicp_batch_render_synthetic": [vp, C.c_int, vp, vp, vp, vp, C.c_float, C.c_float],
            "nicp_initnet_num_params": [],
            "nicp_initnet_create": [vp, C.c_uint64, vp], "nicp_initnet_destroy": [vp],
            "nicp_initnet_get_params": [vp, vp], "nicp_initnet_set_params": [vp, vp],
            "nicp_initnet_set_loss_weights": [vp, vp],
            "nicp_initnet_forward": [vp, vp, C.c_int, C.c_int, vp],
            "nicp_initnet_train_step": [vp, vp, vp, C.c_int, C.c_int, C.c_float, vp],
            "nicp_initnet_read": [vp, C.c_int, C.c_int, vp, C.c_size_t],
            "nicp_batch_net_input": [vp, vp, C.c_int, vp, vp, C.c_float, vp],
            "nicp_batch_align_learned": [vp, vp, C.c_int, vp, vp, vp, vp, C.c_int, vp, vp, vp, vp],
            "nicp_device_alloc": [vp, C.c_size_t, vp], "nicp_device_free": [vp, vp],
            "nicp_device_copy": [vp, vp, vp, C.c_size_t, C.c_int],
            "nicp_initnet_train_step_rows": [vp, vp, C.c_int, vp, vp, vp, C.c_int, C.c_float, vp],
            "nicp_initnet_set_input_mode": [vp, C.c_int],
            "nicp_cloud_gaussians": [vp, C.c_float, C.c_float], "nicp_cloud_set_gaussians": [vp, vp],
            "nicp_cloud_download_gaussians": [vp, vp], "nicp_cloud_transform": [vp, vp],
            "nicp_cloud_add": [vp, vp, vp],
            "nicp_merge": [vp, vp, vp, C.c_int, C.c_int, vp, vp, vp],
            "nicp_batch_set_autotune": [vp, C.c_int], "nicp_batch_autotune_read": [vp, C.c_int, vp, vp, vp],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            if name != "nicp_ctx_stream":
                f.restype = C.c_int
        _lib = L
    return _lib


def _check(rc):
    if rc != NICP_OK:
        raise NicpError("nicp_amd error %d: %s" % (rc, lib().nicp_last_error().decode(errors="replace")))


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def camera(fx, fy, cx, cy, min_d=0.5, max_d=4.5):
    c = Camera()
    c.K[:] = [fx, 0.0, 0.0, 0.0, fy, 0.0, cx, cy, 1.0]
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
    """Values of nicp/nicp_conf/nicp_aligner_1_1.conf:22-27 (information matrices: informationmatrixcalculator.h)."""
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
    """Values of nicp/nicp_conf/nicp_aligner_1_1.conf:30-44 and the Aligner constructor (aligner.cpp:21-41)."""
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


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().nicp_ctx_create(device, C.byref(h)))
        self.h = h
        self._children = weakref.WeakSet()  # objects whose device memory belongs to this context

    def adopt(self, obj):
        self._children.add(obj)

    def close(self):
        """Releases the context and, first, every object still alive on it."""
        if self.h:
            for c in list(self._children):
                c.close()
            lib().nicp_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(lib().nicp_ctx_synchronize(self.h))

    def set_fp64_eigen_gap(self, gap):
        """Later conversions solve windows whose FP32 eigenvalue gap (l1 - l0) / l2 is below
        `gap` again in FP64 (0: the reference's FP32 solve everywhere)."""
        _check(lib().nicp_ctx_set_fp64_eigen_gap(self.h, float(gap)))

    def profile(self, enable=True):
        """Starts (and resets) or stops per-kernel HIP-event timing on this context."""
        _check(lib().nicp_ctx_profile(self.h, int(enable)))

    def profile_read(self):
        """{kernel name: (total ms, launches)} accumulated since profile(True)."""
        n = lib().nicp_kernel_count()
        ms = np.zeros(n, np.float64)
        cnt = np.zeros(n, np.int64)
        _check(lib().nicp_ctx_profile_read(self.h, _p(ms), _p(cnt)))
        return {lib().nicp_kernel_name(i).decode(): (float(ms[i]), int(cnt[i])) for i in range(n)}

    # -- clouds --------------------------------------------------------------------------
    def cloud_from_depth(self, depth_u16, cam, sp, depth_scale=0.001, keep_stats=False):
        d = np.ascontiguousarray(depth_u16, dtype=np.uint16)
        h = C.c_void_p()
        _check(lib().nicp_cloud_from_depth_u16(self.h, _p(d), d.shape[0], d.shape[1], depth_scale, C.byref(cam),
                                               C.byref(sp), int(keep_stats), C.byref(h)))
        return Cloud(self, h, d.shape)

    def cloud_from_depth_f32(self, depth_m, cam, sp, keep_stats=False):
        d = np.ascontiguousarray(depth_m, dtype=np.float32)
        h = C.c_void_p()
        _check(lib().nicp_cloud_from_depth_f32(self.h, _p(d), d.shape[0], d.shape[1], C.byref(cam), C.byref(sp),
                                               int(keep_stats), C.byref(h)))
        return Cloud(self, h, d.shape)

    def depth_scale(self, depth, step, max_depth_cov=0.01, depth_scale=0.001):
        """DepthImage_scale of a uint16 (times depth_scale) or float (metres) image."""
        rows, cols = depth.shape
        out = np.empty((rows // step, cols // step), np.float32)
        if depth.dtype == np.uint16:
            d = np.ascontiguousarray(depth)
            _check(lib().nicp_depth_scale(self.h, _p(d), None, rows, cols, depth_scale, step, max_depth_cov, _p(out)))
        else:
            d = np.ascontiguousarray(depth, np.float32)
            _check(lib().nicp_depth_scale(self.h, None, _p(d), rows, cols, 1.0, step, max_depth_cov, _p(out)))
        return out

    def cloud_from_host(self, points, normals=None, curvature=None, omega_p=None, omega_n=None):
        pts = np.ascontiguousarray(points, np.float32)
        arrs = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (normals, curvature, omega_p, omega_n)]
        h = C.c_void_p()
        _check(lib().nicp_cloud_from_host(self.h, len(pts), _p(pts), *[_p(a) for a in arrs], C.byref(h)))
        return Cloud(self, h, None)

    # -- operations ----------------------------------------------------------------------
    def project(self, cloud, transform16, cam, rows, cols):
        idx = np.empty((rows, cols), np.int32)
        dep = np.empty((rows, cols), np.float32)
        _check(lib().nicp_project(self.h, cloud.h, _p(np.ascontiguousarray(transform16, np.float32)), C.byref(cam),
                                  rows, cols, _p(idx), _p(dep)))
        return idx, dep

    def correspondences(self, ref, cur, ref_index, cur_index, T16, ap):
        rows, cols = ref_index.shape
        pairs = np.empty((rows * cols, 2), np.int32)
        n = C.c_int()
        _check(lib().nicp_correspondences(self.h, ref.h, cur.h, _p(np.ascontiguousarray(ref_index, np.int32)),
                                          _p(np.ascontiguousarray(cur_index, np.int32)), rows, cols,
                                          _p(np.ascontiguousarray(T16, np.float32)), C.byref(ap), _p(pairs),
                                          C.byref(n)))
        return pairs[:n.value].copy()

    def linearize(self, ref, cur, pairs, T16, ap):
        pairs = np.ascontiguousarray(pairs, np.int32)
        H = np.empty(36, np.float64)
        b = np.empty(6, np.float64)
        e = C.c_double()
        i = C.c_int()
        _check(lib().nicp_linearize(self.h, ref.h, cur.h, _p(pairs), len(pairs),
                                    _p(np.ascontiguousarray(T16, np.float32)), C.byref(ap), _p(H), _p(b), C.byref(e),
                                    C.byref(i)))
        return H.reshape(6, 6, order="F"), b, e.value, i.value

    def solution_statistics(self, H, T16, ap):
        """Aligner::_computeStatistics on the device for a given H (6x6) at T: an AlignResult whose
        mean, omega, ratios and solution_valid are set."""
        res = AlignResult()
        Hc = np.ascontiguousarray(np.asarray(H, np.float64).reshape(6, 6).flatten(order="F"))
        _check(lib().nicp_solution_statistics(self.h, _p(Hc), _p(np.ascontiguousarray(T16, np.float32)),
                                              C.byref(ap), C.byref(res)))
        return res

    def align(self, ref, cur, cam, rows, cols, init16, ap):
        T = np.empty(16, np.float32)
        res = AlignResult()
        _check(lib().nicp_align(self.h, ref.h, cur.h, C.byref(cam), rows, cols,
                                _p(np.ascontiguousarray(init16, np.float32)), C.byref(ap), _p(T), C.byref(res)))
        return T, res

    def align_offsets(self, ref, cur, cam, rows, cols, init16, ref_offset16, cur_offset16, ap):
        """nicp_align with the aligner's reference and current sensor offsets (None: identity)."""
        T = np.empty(16, np.float32)
        res = AlignResult()
        ro = None if ref_offset16 is None else np.ascontiguousarray(ref_offset16, np.float32)
        co = None if cur_offset16 is None else np.ascontiguousarray(cur_offset16, np.float32)
        _check(lib().nicp_align_offsets(self.h, ref.h, cur.h, C.byref(cam), rows, cols,
                                        _p(np.ascontiguousarray(init16, np.float32)), _p(ro), _p(co), C.byref(ap),
                                        _p(T), C.byref(res)))
        return T, res

    def align_noalloc(self, ref, cur, cam, rows, cols, init16, ap):
        """nicp_align in the library's fixed device workspace (no device allocation)."""
        T = np.empty(16, np.float32)
        res = AlignResult()
        _check(lib().nicp_align_noalloc(self.h, ref.h, cur.h, C.byref(cam), rows, cols,
                                        _p(np.ascontiguousarray(init16, np.float32)), C.byref(ap), _p(T),
                                        C.byref(res)))
        return T, res

    def nn_correspondences(self, ref, cur, T16, ap, normal_scaling=1.0, cell_size=0.0):
        """Returns (pairs (k, 2), raw nearest neighbour per reference point)."""
        pairs = np.empty((max(ref.n, 1), 2), np.int32)
        nn = np.empty(max(ref.n, 1), np.int32)
        k = C.c_int()
        _check(lib().nicp_nn_correspondences(self.h, ref.h, cur.h, _p(np.ascontiguousarray(T16, np.float32)),
                                             C.byref(ap), normal_scaling, cell_size, _p(pairs), C.byref(k), _p(nn)))
        return pairs[:k.value].copy(), nn[:ref.n].copy()

    def align_nn(self, ref, cur, init16, ap, normal_scaling=1.0, cell_size=0.0):
        T = np.empty(16, np.float32)
        res = AlignResult()
        _check(lib().nicp_align_nn(self.h, ref.h, cur.h, _p(np.ascontiguousarray(init16, np.float32)), C.byref(ap),
                                   normal_scaling, cell_size, _p(T), C.byref(res)))
        return T, res

    def align_nn_batch(self, refs, curs, inits, ap, normal_scaling=1.0, cell_size=0.0):
        """nicp_align_nn_batch over the pairs (refs[i], curs[i]); inits: (num, 16) or None.
        Returns (T (num, 16), [AlignResult])."""
        num = len(refs)
        assert len(curs) == num
        rh = (C.c_void_p * num)(*[c.h for c in refs])
        ch = (C.c_void_p * num)(*[c.h for c in curs])
        T = np.empty((num, 16), np.float32)
        res = (AlignResult * num)()
        ini = None if inits is None else np.ascontiguousarray(inits, np.float32).reshape(num, 16)
        _check(lib().nicp_align_nn_batch(self.h, num, rh, ch, None if ini is None else _p(ini), C.byref(ap),
                                         normal_scaling, cell_size, _p(T), res))
        return T, list(res)

    def last_correspondences(self):
        n = C.c_int()
        _check(lib().nicp_last_correspondences(self.h, None, 0, C.byref(n)))
        pairs = np.empty((max(n.value, 1), 2), np.int32)
        _check(lib().nicp_last_correspondences(self.h, _p(pairs), n.value, C.byref(n)))
        return pairs[:n.value].copy()


class Cloud:
    def __init__(self, ctx, h, shape):
        self.ctx, self.h, self.shape = ctx, h, shape
        ctx.adopt(self)
        n = C.c_int()
        _check(lib().nicp_cloud_size(h, C.byref(n)))
        self.n = n.value

    def close(self):
        if self.h:
            lib().nicp_cloud_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def download(self):
        n = self.n
        pts = np.empty((n, 4), np.float32)
        nrm = np.empty((n, 4), np.float32)
        curv = np.empty(n, np.float32)
        op = np.empty((n, 16), np.float32)
        on = np.empty((n, 16), np.float32)
        _check(lib().nicp_cloud_download(self.h, _p(pts), _p(nrm), _p(curv), _p(op), _p(on)))
        return pts, nrm, curv, op, on

    def download_stats(self):
        n = self.n
        ev = np.empty((n, 9), np.float32)
        el = np.empty((n, 3), np.float32)
        mn = np.empty((n, 4), np.float32)
        npt = np.empty(n, np.int32)
        _check(lib().nicp_cloud_download_stats(self.h, _p(ev), _p(el), _p(mn), _p(npt)))
        return ev, el, mn, npt

    def index_image(self):
        idx = np.empty(self.shape, np.int32)
        _check(lib().nicp_cloud_index_image(self.h, _p(idx)))
        return idx

    def _size(self):
        n = C.c_int()
        _check(lib().nicp_cloud_size(self.h, C.byref(n)))
        self.n = n.value

    def compute_gaussians(self, baseline=0.075, alpha=0.1):
        """The Gaussians of PinholePointProjector::unProject (needs a cloud made from depth)."""
        _check(lib().nicp_cloud_gaussians(self.h, baseline, alpha))

    def set_gaussians(self, g):
        """g: (n, 28) float32 records (nicp_gaussian)."""
        g = np.ascontiguousarray(g, np.float32)
        assert g.shape == (self.n, 28)
        _check(lib().nicp_cloud_set_gaussians(self.h, _p(g)))

    def download_gaussians(self):
        """(n, 28) float32 records: mean 0:3, int32 flags at 3, covariance 4:13, information 13:22,
        information vector 22:25 (3x3 matrices column-major)."""
        g = np.empty((self.n, 28), np.float32)
        _check(lib().nicp_cloud_download_gaussians(self.h, _p(g)))
        return g

    def transform(self, T16):
        """Cloud::transformInPlace."""
        _check(lib().nicp_cloud_transform(self.h, _p(np.ascontiguousarray(T16, np.float32))))

    def add(self, other, T16):
        """Cloud::add: transforms `other` in place by T16 and appends it."""
        _check(lib().nicp_cloud_add(self.h, other.h, _p(np.ascontiguousarray(T16, np.float32))))
        self._size()
        self.shape = (1, self.n)

    def merge(self, T16, cam, rows, cols, params=None, want_collapsed=False):
        """Merger::merge seen from T16; returns the number of suppressed points (and the
        per-point decisions)."""
        params = params or merger_params()
        ns = C.c_int()
        col = np.empty(self.n, np.int32) if want_collapsed else None
        _check(lib().nicp_merge(self.h, _p(np.ascontiguousarray(T16, np.float32)), C.byref(cam), rows, cols,
                                C.byref(params), C.byref(ns), _p(col)))
        self._size()
        self.shape = (1, self.n)
        return (ns.value, col) if want_collapsed else ns.value


class Batch:
    """Frames of uneven size converted and aligned pair-wise in single launches per stage."""

    def __init__(self, ctx, shapes, cams, sp):
        self.ctx = ctx
        n = len(shapes)
        rows = np.array([s[0] for s in shapes], np.int32)
        cols = np.array([s[1] for s in shapes], np.int32)
        cam_arr = (Camera * n)(*cams)
        h = C.c_void_p()
        _check(lib().nicp_batch_create(ctx.h, n, _p(rows), _p(cols), cam_arr, C.byref(sp), C.byref(h)))
        self.h = h
        self.n = n
        ctx.adopt(self)

    def close(self):
        if self.h:
            lib().nicp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, frame, depth_u16, depth_scale=0.001):
        d = np.ascontiguousarray(depth_u16, np.uint16)
        _check(lib().nicp_batch_upload_depth(self.h, frame, _p(d), depth_scale))

    def num_pixels(self):
        n = C.c_longlong()
        _check(lib().nicp_batch_num_pixels(self.h, C.byref(n)))
        return n.value

    def upload_depth_device(self, src_ptr, depth_scale=0.001):
        """Every frame's depth from device memory at src_ptr (uint16, frames back to back); enqueued
        on the context's stream, no wait."""
        _check(lib().nicp_batch_upload_depth_device(self.h, C.c_void_p(src_ptr), depth_scale))

    def download_depth_device(self, dst_ptr):
        """Every frame's depth into device memory at dst_ptr (the layout of upload_depth_device)."""
        _check(lib().nicp_batch_download_depth_device(self.h, C.c_void_p(dst_ptr)))

    def convert(self):
        _check(lib().nicp_batch_convert(self.h))

    def align(self, ref, cur, init16, ap, use_graph=True, want_results=True):
        ref = np.ascontiguousarray(ref, np.int32)
        cur = np.ascontiguousarray(cur, np.int32)
        n = len(ref)
        init = None if init16 is None else np.ascontiguousarray(init16, np.float32).reshape(n, 16)
        T = np.empty((n, 16), np.float32)
        res = (AlignResult * n)() if want_results else None
        _check(lib().nicp_batch_align(self.h, n, _p(ref), _p(cur), _p(init), C.byref(ap), int(use_graph), _p(T), res))
        return T, res

    def set_autotune(self, enable=True):
        """Times the projection strategies on each new plan shape's first align call (eagerly) and
        uses the faster one afterwards."""
        _check(lib().nicp_batch_set_autotune(self.h, int(enable)))

    def set_projection(self, strategy):
        """Forces projection strategy 0, 1 or 2 (tile-binned) on later align calls; -1: default."""
        _check(lib().nicp_batch_set_projection(self.h, int(strategy)))

    def probe(self, ref, cur, starts16, ap, sub=4, twin=False):
        """One linearisation per pair at its start over every sub-th row and column (the learned
        aligner's start probe); results hold H, b, error, inliers and the correspondence count."""
        ref = np.ascontiguousarray(ref, np.int32)
        cur = np.ascontiguousarray(cur, np.int32)
        n = len(ref)
        st = np.ascontiguousarray(starts16, np.float32).reshape(n, 16)
        res = (AlignResult * n)()
        _check(lib().nicp_batch_probe(self.h, n, _p(ref), _p(cur), _p(st), C.byref(ap), int(sub), int(twin), res))
        return res

    def set_bin_capacity(self, capacity):
        """Source tiles a projector tile of the tile-binned projection lists (0..256)."""
        _check(lib().nicp_batch_set_bin_capacity(self.h, int(capacity)))

    def autotune_state(self):
        """[(choice, ms strategy 0, ms strategy 1, ms strategy 2)] per tuned shape."""
        n = C.c_int()
        _check(lib().nicp_batch_autotune_read(self.h, 0, C.byref(n), None, None))
        ch = np.empty(max(n.value, 1), np.int32)
        ms = np.empty(3 * max(n.value, 1), np.float32)
        _check(lib().nicp_batch_autotune_read(self.h, n.value, C.byref(n), _p(ch), _p(ms)))
        return [(int(ch[i]), float(ms[3 * i]), float(ms[3 * i + 1]), float(ms[3 * i + 2])) for i in range(n.value)]

    def frame_cloud(self, frame, shape):
        h = C.c_void_p()
        _check(lib().nicp_batch_frame_cloud(self.h, frame, C.byref(h)))
        return Cloud(self.ctx, h, shape)

    def render_synthetic(self, frames, scene_seeds, streams, poses16, noise=1.5e-3, holes=0.1):
        """Renders seeded synthetic scenes into the depth of the given frames on the device."""
        frames = np.ascontiguousarray(frames, np.int32)
        n = len(frames)
        seeds = np.ascontiguousarray(scene_seeds, np.uint64)
        streams = np.ascontiguousarray(streams, np.uint64)
        poses = np.ascontiguousarray(poses16, np.float32).reshape(n, 16)
        _check(lib().nicp_batch_render_synthetic(self.h, n, _p(frames), _p(seeds), _p(streams), _p(poses), noise,
                                                 holes))

    def net_input(self, net, ref, cur, dev_ptr, max_depth_cov=0.01):
        """Writes the network input of the pairs to the device pointer dev_ptr ([n][60][80][2] float)."""
        ref = np.ascontiguousarray(ref, np.int32)
        cur = np.ascontiguousarray(cur, np.int32)
        _check(lib().nicp_batch_net_input(self.h, net.h, len(ref), _p(ref), _p(cur), max_depth_cov,
                                          C.c_void_p(int(dev_ptr))))

    def align_learned(self, net, ref, cur, ap, fallback16=None, use_graph=True, want_results=True):
        """Returns (T, results, starts, used_network)."""
        ref = np.ascontiguousarray(ref, np.int32)
        cur = np.ascontiguousarray(cur, np.int32)
        n = len(ref)
        fb = None if fallback16 is None else np.ascontiguousarray(fallback16, np.float32).reshape(n, 16)
        T = np.empty((n, 16), np.float32)
        starts = np.empty((n, 16), np.float32)
        used = np.empty(n, np.int32)
        res = (AlignResult * n)() if want_results else None
        _check(lib().nicp_batch_align_learned(self.h, net.h, n, _p(ref), _p(cur), _p(fb), C.byref(ap), int(use_graph),
                                              _p(T), res, _p(starts), _p(used)))
        return T, res, starts, used


class DeviceBuffer:
    """Device memory of a context holding a float32 (or other dtype) array of a fixed shape."""

    def __init__(self, ctx, shape, dtype=np.float32):
        self.ctx = ctx
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(lib().nicp_device_alloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p.value
        ctx.adopt(self)

    def row_ptr(self, row):
        """Address of row `row` along the first axis."""
        return self.ptr + row * (self.nbytes // self.shape[0])

    def upload(self, a, row=0):
        a = np.ascontiguousarray(a, self.dtype)
        if row * (self.nbytes // self.shape[0]) + a.nbytes > self.nbytes:
            raise NicpError("upload past the end of the buffer")
        _check(lib().nicp_device_copy(self.ctx.h, C.c_void_p(self.row_ptr(row)), _p(a), a.nbytes, 0))

    def download(self, rows=None, row=0):
        shape = self.shape if rows is None else (rows,) + self.shape[1:]
        out = np.empty(shape, self.dtype)
        if row * (self.nbytes // self.shape[0]) + out.nbytes > self.nbytes:
            raise NicpError("download past the end of the buffer")
        _check(lib().nicp_device_copy(self.ctx.h, _p(out), C.c_void_p(self.row_ptr(row)), out.nbytes, 1))
        return out

    def close(self):
        if self.ptr:
            lib().nicp_device_free(self.ctx.h, C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class InitNet:
    """The initial-guess CNN (device weights; bf16 matrix-core layers, fp32 accumulation)."""

    IN_SHAPE = (60, 80, 2)

    def __init__(self, ctx, seed=1):
        self.ctx = ctx
        h = C.c_void_p()
        _check(lib().nicp_initnet_create(ctx.h, seed, C.byref(h)))
        self.h = h
        self.num_params = lib().nicp_initnet_num_params()
        ctx.adopt(self)

    def close(self):
        if self.h:
            lib().nicp_initnet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        out = np.empty(self.num_params, np.float32)
        _check(lib().nicp_initnet_get_params(self.h, _p(out)))
        return out

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.num_params:
            raise NicpError("expected %d parameters, got %d" % (self.num_params, p.size))
        _check(lib().nicp_initnet_set_params(self.h, _p(p)))

    def set_loss_weights(self, w6):
        w = np.ascontiguousarray(w6, np.float32)
        _check(lib().nicp_initnet_set_loss_weights(self.h, _p(w)))

    def forward(self, x):
        """x: host array (B, 60, 80, 2) -> (B, 6)."""
        x = np.ascontiguousarray(x, np.float32)
        B = x.shape[0]
        out = np.empty((B, 6), np.float32)
        _check(lib().nicp_initnet_forward(self.h, _p(x), B, 0, _p(out)))
        return out

    def forward_device(self, x_ptr, batch, out_ptr):
        _check(lib().nicp_initnet_forward(self.h, C.c_void_p(int(x_ptr)), batch, 1, C.c_void_p(int(out_ptr))))

    def train_step(self, x, target, lr=0.0):
        """Host arrays; returns the loss (lr > 0 applies one Adam step)."""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        loss = C.c_float()
        _check(lib().nicp_initnet_train_step(self.h, _p(x), _p(t), x.shape[0], 0, lr, C.byref(loss)))
        return loss.value

    def train_step_device(self, x_ptr, t_ptr, batch, lr):
        loss = C.c_float()
        _check(lib().nicp_initnet_train_step(self.h, C.c_void_p(int(x_ptr)), C.c_void_p(int(t_ptr)), batch, 1, lr,
                                             C.byref(loss)))
        return loss.value

    def train_step_rows(self, inputs, rows, targets, lr, flags=None):
        """One step on the minibatch `rows` of the device-resident DeviceBuffer `inputs`, each
        sample transformed by its symmetry flags, against host `targets` (batch, 6)."""
        r = np.ascontiguousarray(rows, np.int32)
        f = None if flags is None else np.ascontiguousarray(flags, np.int32)
        t = np.ascontiguousarray(targets, np.float32)
        loss = C.c_float()
        _check(lib().nicp_initnet_train_step_rows(self.h, C.c_void_p(inputs.ptr), inputs.shape[0], _p(r), _p(f),
                                                  _p(t), len(r), lr, C.byref(loss)))
        return loss.value

    def set_input_mode(self, mode):
        _check(lib().nicp_initnet_set_input_mode(self.h, int(mode)))

    def read(self, what, index, count):
        out = np.empty(count, np.float32)
        _check(lib().nicp_initnet_read(self.h, what, index, _p(out), count))
        return out
