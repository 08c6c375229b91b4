// Seeded synthetic depth scenes for the NICP registration loop.
//
// A scene is 20 planar rectangles and boxes in front of a pinhole camera
// (a back wall, a floor and a side wall, then 17 random objects with
// extents of 0.5-4 m). Depth images are rendered by exact ray casting in
// double precision, perturbed with gaussian noise of sigma = k * z^2 m,
// quantised to uint16 millimetres and punched with uniform random holes
// (value 0). Every random draw comes from a counter-based hash of
// (seed, stream, index), so the output does not depend on the thread count.
//
// The second frame of a pair is the same scene seen from a camera moved by a
// random rigid motion T_gt (uniform axis, angle <= max_rot, translation
// uniform in a ball of radius max_trans). T_gt is the pose of the second
// camera in the first camera's frame, which is the transform the reference's
// Aligner::T() returns for (reference = first, current = second)
// (nicp/nicp/alignerprojective.cpp:51,61).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
inline double uniform(uint64_t seed, uint64_t stream, uint64_t i) {
  uint64_t h = mix64(mix64(seed ^ mix64(stream)) + i);
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}
inline double gaussian(uint64_t seed, uint64_t stream, uint64_t i) {
  double u1 = uniform(seed, stream, 2 * i), u2 = uniform(seed, stream, 2 * i + 1);
  if (u1 < 1e-300) u1 = 1e-300;
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

struct Prim {
  int box;
  double c[3];
  double R[9];  // column-major; columns are the local axes
  double h[3];
};

struct Rng {
  uint64_t seed, stream, i = 0;
  double u() { return uniform(seed, stream, i++); }
  double u(double a, double b) { return a + (b - a) * u(); }
  double g() { return gaussian(seed, stream + 0x5151, i++); }
};

void rot_from_axis_angle(const double *axis, double ang, double *R) {
  double c = std::cos(ang), s = std::sin(ang), t = 1.0 - c;
  double x = axis[0], y = axis[1], z = axis[2];
  R[0] = t * x * x + c;     R[3] = t * x * y - s * z; R[6] = t * x * z + s * y;
  R[1] = t * x * y + s * z; R[4] = t * y * y + c;     R[7] = t * y * z - s * x;
  R[2] = t * x * z - s * y; R[5] = t * y * z + s * x; R[8] = t * z * z + c;
}
void random_axis(Rng &r, double *a) {
  double n = 0.0;
  do {
    a[0] = r.g(); a[1] = r.g(); a[2] = r.g();
    n = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  } while (n < 1e-9);
  a[0] /= n; a[1] /= n; a[2] /= n;
}
void mul3(const double *A, const double *B, double *C) {
  double o[9];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) o[c * 3 + r] = A[r] * B[c * 3] + A[3 + r] * B[c * 3 + 1] + A[6 + r] * B[c * 3 + 2];
  std::memcpy(C, o, sizeof(o));
}
// a random rotation that tilts the local z axis onto `normal` within max_tilt radians
void facing(Rng &r, const double *normal, double max_tilt, double *R) {
  // base frame with z = normal
  double z[3] = {normal[0], normal[1], normal[2]};
  double t[3] = {1, 0, 0};
  if (std::fabs(z[0]) > 0.9) { t[0] = 0; t[1] = 1; }
  double x[3] = {t[1] * z[2] - t[2] * z[1], t[2] * z[0] - t[0] * z[2], t[0] * z[1] - t[1] * z[0]};
  double nx = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  for (int i = 0; i < 3; ++i) x[i] /= nx;
  double y[3] = {z[1] * x[2] - z[2] * x[1], z[2] * x[0] - z[0] * x[2], z[0] * x[1] - z[1] * x[0]};
  double B[9] = {x[0], x[1], x[2], y[0], y[1], y[2], z[0], z[1], z[2]};
  double ax[3];
  random_axis(r, ax);
  double Rt[9];
  rot_from_axis_angle(ax, r.u(0.0, max_tilt), Rt);
  mul3(Rt, B, R);
}

void make_scene(uint64_t seed, double fx, double fy, double cx, double cy, int cols, int rows, std::vector<Prim> &prims) {
  Rng r{seed, 1};
  const double d2r = M_PI / 180.0;
  prims.clear();
  Prim p;
  // back wall
  p.box = 0;
  p.c[0] = r.u(-0.3, 0.3); p.c[1] = r.u(-0.3, 0.3); p.c[2] = r.u(3.3, 4.0);
  { double n[3] = {0, 0, -1}; facing(r, n, 20 * d2r, p.R); }
  p.h[0] = 4.0; p.h[1] = 3.0; p.h[2] = 0.0;
  prims.push_back(p);
  // floor
  p.c[0] = 0.0; p.c[1] = r.u(0.9, 1.4); p.c[2] = 2.5;
  { double n[3] = {0, -1, 0}; facing(r, n, 10 * d2r, p.R); }
  p.h[0] = 3.0; p.h[1] = 3.0;
  prims.push_back(p);
  // side wall
  double side = r.u() < 0.5 ? -1.0 : 1.0;
  p.c[0] = side * r.u(1.5, 2.2); p.c[1] = 0.0; p.c[2] = 2.5;
  { double n[3] = {-side, 0, 0}; facing(r, n, 15 * d2r, p.R); }
  p.h[0] = 3.0; p.h[1] = 2.0;
  prims.push_back(p);
  const double hx = 0.5 * cols / fx, hy = 0.5 * rows / fy;
  (void)cx; (void)cy;
  for (int i = 3; i < 20; ++i) {
    Prim q;
    q.box = r.u() < 0.5;
    double z = q.box ? r.u(1.2, 3.2) : r.u(1.0, 3.2);
    q.c[2] = z;
    q.c[0] = r.u(-0.8, 0.8) * z * hx;
    q.c[1] = r.u(-0.8, 0.8) * z * hy;
    double ax[3];
    random_axis(r, ax);
    rot_from_axis_angle(ax, r.u(0.0, M_PI), q.R);
    if (q.box) {
      q.h[0] = r.u(0.1, 0.5); q.h[1] = r.u(0.1, 0.5); q.h[2] = r.u(0.1, 0.5);
    } else {
      q.h[0] = r.u(0.25, 1.0); q.h[1] = r.u(0.25, 1.0); q.h[2] = 0.0;
    }
    prims.push_back(q);
  }
}

// ray (o + t d) against the primitives; returns the smallest t > 0 or -1
double cast(const std::vector<Prim> &prims, const double *o, const double *d) {
  double best = -1.0;
  for (const Prim &p : prims) {
    double oc[3] = {o[0] - p.c[0], o[1] - p.c[1], o[2] - p.c[2]};
    double ol[3], dl[3];
    for (int k = 0; k < 3; ++k) {
      ol[k] = p.R[k * 3] * oc[0] + p.R[k * 3 + 1] * oc[1] + p.R[k * 3 + 2] * oc[2];
      dl[k] = p.R[k * 3] * d[0] + p.R[k * 3 + 1] * d[1] + p.R[k * 3 + 2] * d[2];
    }
    double t = -1.0;
    if (!p.box) {
      if (std::fabs(dl[2]) < 1e-12) continue;
      double tt = -ol[2] / dl[2];
      if (tt <= 0.0) continue;
      double x = ol[0] + tt * dl[0], y = ol[1] + tt * dl[1];
      if (std::fabs(x) > p.h[0] || std::fabs(y) > p.h[1]) continue;
      t = tt;
    } else {
      double t0 = -1e300, t1 = 1e300;
      bool miss = false;
      for (int k = 0; k < 3; ++k) {
        if (std::fabs(dl[k]) < 1e-15) {
          if (std::fabs(ol[k]) > p.h[k]) { miss = true; break; }
          continue;
        }
        double a = (-p.h[k] - ol[k]) / dl[k], b = (p.h[k] - ol[k]) / dl[k];
        if (a > b) std::swap(a, b);
        t0 = std::max(t0, a);
        t1 = std::min(t1, b);
      }
      if (miss || t0 > t1 || t0 <= 0.0) continue;
      t = t0;
    }
    if (t > 0.0 && (best < 0.0 || t < best)) best = t;
  }
  return best;
}

// like cast() but also returns the unit normal of the hit surface
double cast_normal(const std::vector<Prim> &prims, const double *o, const double *d, double *nrm) {
  double best = -1.0;
  for (const Prim &p : prims) {
    double oc[3] = {o[0] - p.c[0], o[1] - p.c[1], o[2] - p.c[2]};
    double ol[3], dl[3];
    for (int k = 0; k < 3; ++k) {
      ol[k] = p.R[k * 3] * oc[0] + p.R[k * 3 + 1] * oc[1] + p.R[k * 3 + 2] * oc[2];
      dl[k] = p.R[k * 3] * d[0] + p.R[k * 3 + 1] * d[1] + p.R[k * 3 + 2] * d[2];
    }
    double t = -1.0;
    int axis = 2;
    if (!p.box) {
      if (std::fabs(dl[2]) < 1e-12) continue;
      double tt = -ol[2] / dl[2];
      if (tt <= 0.0) continue;
      double x = ol[0] + tt * dl[0], y = ol[1] + tt * dl[1];
      if (std::fabs(x) > p.h[0] || std::fabs(y) > p.h[1]) continue;
      t = tt;
    } else {
      double t0 = -1e300, t1 = 1e300;
      bool miss = false;
      for (int k = 0; k < 3; ++k) {
        if (std::fabs(dl[k]) < 1e-15) {
          if (std::fabs(ol[k]) > p.h[k]) { miss = true; break; }
          continue;
        }
        double a = (-p.h[k] - ol[k]) / dl[k], b = (p.h[k] - ol[k]) / dl[k];
        if (a > b) std::swap(a, b);
        if (a > t0) { t0 = a; axis = k; }
        t1 = std::min(t1, b);
      }
      if (miss || t0 > t1 || t0 <= 0.0) continue;
      t = t0;
    }
    if (t > 0.0 && (best < 0.0 || t < best)) {
      best = t;
      double n[3] = {p.R[axis * 3], p.R[axis * 3 + 1], p.R[axis * 3 + 2]};
      if (n[0] * d[0] + n[1] * d[1] + n[2] * d[2] > 0.0) { n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2]; }
      nrm[0] = n[0]; nrm[1] = n[1]; nrm[2] = n[2];
    }
  }
  return best;
}

// flat information matrix U diag(a, b, b) U^T = b I + (a - b) n n^T, 4x4 column-major
void flat_omega(const float *n, float a, float b, float *O) {
  for (int k = 0; k < 16; ++k) O[k] = 0.0f;
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) O[c * 4 + r] = (float)((r == c ? (double)b : 0.0) + ((double)a - b) * n[r] * n[c]);
}

void render(const std::vector<Prim> &prims, int rows, int cols, double fx, double fy, double cx, double cy,
            const double *R /*3x3 cm*/, const double *t, double noise, double holes, uint64_t seed,
            uint64_t stream, uint16_t *out) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int r = 0; r < rows; ++r) {
    for (int c = 0; c < cols; ++c) {
      double dc[3] = {(c - cx) / fx, (r - cy) / fy, 1.0};
      double d[3] = {R[0] * dc[0] + R[3] * dc[1] + R[6] * dc[2], R[1] * dc[0] + R[4] * dc[1] + R[7] * dc[2],
                     R[2] * dc[0] + R[5] * dc[1] + R[8] * dc[2]};
      double z = cast(prims, t, d);
      uint64_t pix = (uint64_t)r * cols + c;
      uint16_t v = 0;
      if (z > 0.0) {
        double zn = z + gaussian(seed, stream, pix) * noise * z * z;
        long q = std::lround(zn * 1000.0);
        if (q >= 1 && q <= 65535) v = (uint16_t)q;
      }
      if (uniform(seed, stream + 0x77, pix) < holes) v = 0;
      out[pix] = v;
    }
  }
}

}  // namespace

extern "C" {

// Renders the scene of `scene_seed` from the camera pose `pose` (float[16]
// column-major, pose of the camera in the scene frame). `stream` salts the
// noise and hole draws so that two renders of one scene are independent.
int nicp_synth_render(uint64_t scene_seed, uint64_t stream, int rows, int cols, float fx, float fy, float cx,
                      float cy, const float *pose, float noise_coeff, float hole_fraction, uint16_t *depth) {
  if (rows <= 0 || cols <= 0 || !depth || !pose) return -1;
  std::vector<Prim> prims;
  make_scene(scene_seed, fx, fy, cx, cy, cols, rows, prims);
  double R[9], t[3];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) R[c * 3 + r] = pose[c * 4 + r];
  for (int r = 0; r < 3; ++r) t[r] = pose[12 + r];
  render(prims, rows, cols, fx, fy, cx, cy, R, t, noise_coeff, hole_fraction, scene_seed, stream, depth);
  return 0;
}

// Samples the rigid motion of pair `seed` (see the file comment).
void nicp_synth_motion(uint64_t seed, float max_rot_deg, float max_trans_m, float *T_gt) {
  Rng r{seed, 2};
  double ax[3];
  random_axis(r, ax);
  double ang = r.u(0.0, max_rot_deg * M_PI / 180.0);
  double R[9];
  rot_from_axis_angle(ax, ang, R);
  double dir[3];
  random_axis(r, dir);
  double rad = max_trans_m * std::cbrt(r.u());
  for (int i = 0; i < 16; ++i) T_gt[i] = 0.0f;
  for (int c = 0; c < 3; ++c)
    for (int rr = 0; rr < 3; ++rr) T_gt[c * 4 + rr] = (float)R[c * 3 + rr];
  for (int k = 0; k < 3; ++k) T_gt[12 + k] = (float)(dir[k] * rad);
  T_gt[15] = 1.0f;
}

// An unstructured cloud pair (the kNN configuration): n points sampled from the
// visible surfaces of scene `seed` (rendered noise-free at rows x cols), with the
// exact surface normals, curvature 0 and flat information matrices; the target is
// the source moved by T_gt^-1 with gaussian position noise `noise` (m), so that
// p_source = T_gt p_target. Arrays: points/normals n*4, curvature n, omegas n*16.
int nicp_synth_cloud_pair(uint64_t seed, int n, int rows, int cols, float fx, float fy, float cx, float cy,
                          float max_rot_deg, float max_trans_m, float noise, float *src_points, float *src_normals,
                          float *tgt_points, float *tgt_normals, float *curvature, float *omega_p, float *omega_n,
                          float *T_gt) {
  if (n <= 0 || rows <= 0 || cols <= 0) return -1;
  std::vector<Prim> prims;
  make_scene(seed, fx, fy, cx, cy, cols, rows, prims);
  const size_t P = (size_t)rows * cols;
  std::vector<float> hit(P * 6);
  std::vector<char> ok(P, 0);
#pragma omp parallel for schedule(dynamic, 4)
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c) {
      double d[3] = {(c - cx) / fx, (r - cy) / fy, 1.0}, o[3] = {0, 0, 0}, nn[3] = {0, 0, 0};
      double t = cast_normal(prims, o, d, nn);
      size_t i = (size_t)r * cols + c;
      if (t > 0.5 && t < 4.5) {
        ok[i] = 1;
        for (int k = 0; k < 3; ++k) { hit[6 * i + k] = (float)(t * d[k]); hit[6 * i + 3 + k] = (float)nn[k]; }
      }
    }
  std::vector<size_t> valid;
  for (size_t i = 0; i < P; ++i)
    if (ok[i]) valid.push_back(i);
  if ((int)valid.size() < n) return -2;
  // partial Fisher-Yates with the counter-based hash
  for (int i = 0; i < n; ++i) {
    size_t j = i + (size_t)(uniform(seed, 0x3131, i) * (valid.size() - i));
    if (j >= valid.size()) j = valid.size() - 1;
    std::swap(valid[i], valid[j]);
  }
  nicp_synth_motion(seed, max_rot_deg, max_trans_m, T_gt);
  // inverse motion for the target: R^T (p - t)
  double R[9], t[3];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) R[c * 3 + r] = T_gt[c * 4 + r];
  for (int r = 0; r < 3; ++r) t[r] = T_gt[12 + r];
  for (int i = 0; i < n; ++i) {
    const float *h = &hit[6 * valid[i]];
    for (int k = 0; k < 3; ++k) {
      src_points[4 * i + k] = h[k];
      src_normals[4 * i + k] = h[3 + k];
    }
    src_points[4 * i + 3] = 1.0f;
    src_normals[4 * i + 3] = 0.0f;
    for (int k = 0; k < 3; ++k) {
      double pk = 0.0, nk = 0.0;
      for (int m = 0; m < 3; ++m) {
        pk += R[k * 3 + m] * ((double)h[m] - t[m]);
        nk += R[k * 3 + m] * (double)h[3 + m];
      }
      tgt_points[4 * i + k] = (float)(pk + noise * gaussian(seed, 0x4141 + k, i));
      tgt_normals[4 * i + k] = (float)nk;
    }
    tgt_points[4 * i + 3] = 1.0f;
    tgt_normals[4 * i + 3] = 0.0f;
    curvature[i] = 0.0f;
    flat_omega(&tgt_normals[4 * i], 1000.0f, 0.001f, omega_p + 16 * (size_t)i);
    flat_omega(&tgt_normals[4 * i], 1000.0f, 0.001f, omega_n + 16 * (size_t)i);
  }
  return 0;
}

// A registration pair: the reference frame at the identity, the current frame
// at T_gt. Both are uint16 rows x cols, values in millimetres.
int nicp_synth_pair(uint64_t seed, int rows, int cols, float fx, float fy, float cx, float cy, float max_rot_deg,
                    float max_trans_m, float noise_coeff, float hole_fraction, uint16_t *depth_ref,
                    uint16_t *depth_cur, float *T_gt) {
  if (rows <= 0 || cols <= 0 || !depth_ref || !depth_cur || !T_gt) return -1;
  nicp_synth_motion(seed, max_rot_deg, max_trans_m, T_gt);
  float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  nicp_synth_render(seed, 10, rows, cols, fx, fy, cx, cy, I, noise_coeff, hole_fraction, depth_ref);
  nicp_synth_render(seed, 11, rows, cols, fx, fy, cx, cy, T_gt, noise_coeff, hole_fraction, depth_cur);
  return 0;
}

}  // extern "C"
