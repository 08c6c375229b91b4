// Seeded synthetic scenes shared by the host generator (synth.cpp, built with
// g++) and the device renderer (synth_gpu.hip, built with hipcc): the scene
// description, the counter-based random draws and the exact ray cast. Scene
// construction runs on the host only; the draws and the cast are compiled for
// both sides so that a GPU render reproduces the host render pixel for pixel
// (up to the last-bit behaviour of log/cos in the noise draw).
//
// A scene is 20 planar rectangles and boxes in front of a pinhole camera
// (a back wall, a floor and a side wall, then 17 random objects with extents
// of 0.5-4 m). Every random draw is a hash of (seed, stream, index), so the
// output does not depend on the thread count or the device.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#if defined(__HIPCC__)
#define SYN_HD __host__ __device__
#else
#define SYN_HD
#endif

namespace synth {

constexpr int kPrims = 20;

SYN_HD inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
SYN_HD inline double uniform(uint64_t seed, uint64_t stream, uint64_t i) {
  uint64_t h = mix64(mix64(seed ^ mix64(stream)) + i);
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}
SYN_HD inline double gaussian(uint64_t seed, uint64_t stream, uint64_t i) {
  double u1 = uniform(seed, stream, 2 * i), u2 = uniform(seed, stream, 2 * i + 1);
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

struct Prim {
  int box;
  double c[3];
  double R[9];  // column-major; columns are the local axes
  double h[3];
};

// smallest t > 0 of the ray o + t d against the primitives, or -1; when nrm is
// not null it receives the unit normal of the hit surface, facing the ray
SYN_HD inline double cast(const Prim *prims, int n, const double *o, const double *d, double *nrm = nullptr) {
  double best = -1.0;
  for (int pi = 0; pi < n; ++pi) {
    const Prim &p = prims[pi];
    double oc[3] = {o[0] - p.c[0], o[1] - p.c[1], o[2] - p.c[2]};
    double ol[3], dl[3];
    for (int k = 0; k < 3; ++k) {
      ol[k] = p.R[k * 3] * oc[0] + p.R[k * 3 + 1] * oc[1] + p.R[k * 3 + 2] * oc[2];
      dl[k] = p.R[k * 3] * d[0] + p.R[k * 3 + 1] * d[1] + p.R[k * 3 + 2] * d[2];
    }
    double t = -1.0;
    int axis = 2;
    if (!p.box) {
      if (fabs(dl[2]) < 1e-12) continue;
      double tt = -ol[2] / dl[2];
      if (tt <= 0.0) continue;
      double x = ol[0] + tt * dl[0], y = ol[1] + tt * dl[1];
      if (fabs(x) > p.h[0] || fabs(y) > p.h[1]) continue;
      t = tt;
    } else {
      double t0 = -1e300, t1 = 1e300;
      bool miss = false;
      for (int k = 0; k < 3; ++k) {
        if (fabs(dl[k]) < 1e-15) {
          if (fabs(ol[k]) > p.h[k]) {
            miss = true;
            break;
          }
          continue;
        }
        double a = (-p.h[k] - ol[k]) / dl[k], b = (p.h[k] - ol[k]) / dl[k];
        if (a > b) {
          const double s = a;
          a = b;
          b = s;
        }
        if (a > t0) {
          t0 = a;
          axis = k;
        }
        t1 = fmin(t1, b);
      }
      if (miss || t0 > t1 || t0 <= 0.0) continue;
      t = t0;
    }
    if (t > 0.0 && (best < 0.0 || t < best)) {
      best = t;
      if (nrm) {
        double nn[3] = {p.R[axis * 3], p.R[axis * 3 + 1], p.R[axis * 3 + 2]};
        if (nn[0] * d[0] + nn[1] * d[1] + nn[2] * d[2] > 0.0) {
          nn[0] = -nn[0];
          nn[1] = -nn[1];
          nn[2] = -nn[2];
        }
        nrm[0] = nn[0];
        nrm[1] = nn[1];
        nrm[2] = nn[2];
      }
    }
  }
  return best;
}

// depth sample of pixel (r, c) of a camera at pose (R column-major, t): the ray's
// hit distance along the optical axis with gaussian noise sigma = noise * z^2,
// rounded to millimetres, 0 for no hit, out of range or a hole
SYN_HD inline uint16_t depth_sample(const Prim *prims, int n, int r, int c, int cols, double fx, double fy, double cx,
                                    double cy, const double *R, const double *t, double noise, double holes,
                                    uint64_t seed, uint64_t stream) {
  double dc[3] = {(c - cx) / fx, (r - cy) / fy, 1.0};
  double d[3] = {R[0] * dc[0] + R[3] * dc[1] + R[6] * dc[2], R[1] * dc[0] + R[4] * dc[1] + R[7] * dc[2],
                 R[2] * dc[0] + R[5] * dc[1] + R[8] * dc[2]};
  double z = cast(prims, n, t, d);
  uint64_t pix = (uint64_t)r * cols + c;
  uint16_t v = 0;
  if (z > 0.0) {
    double zn = z + gaussian(seed, stream, pix) * noise * z * z;
    double q = round(zn * 1000.0);
    if (q >= 1.0 && q <= 65535.0) v = (uint16_t)q;
  }
  if (uniform(seed, stream + 0x77, pix) < holes) v = 0;
  return v;
}

// ---------------------------------------------------------------------------------
// host-side scene and motion sampling
// ---------------------------------------------------------------------------------
struct Rng {
  uint64_t seed, stream, i = 0;
  double u() { return uniform(seed, stream, i++); }
  double u(double a, double b) { return a + (b - a) * u(); }
  double g() { return gaussian(seed, stream + 0x5151, i++); }
};

inline void rot_from_axis_angle(const double *axis, double ang, double *R) {
  double c = cos(ang), s = sin(ang), t = 1.0 - c;
  double x = axis[0], y = axis[1], z = axis[2];
  R[0] = t * x * x + c;     R[3] = t * x * y - s * z; R[6] = t * x * z + s * y;
  R[1] = t * x * y + s * z; R[4] = t * y * y + c;     R[7] = t * y * z - s * x;
  R[2] = t * x * z - s * y; R[5] = t * y * z + s * x; R[8] = t * z * z + c;
}
inline void random_axis(Rng &r, double *a) {
  double n = 0.0;
  do {
    a[0] = r.g(); a[1] = r.g(); a[2] = r.g();
    n = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  } while (n < 1e-9);
  a[0] /= n; a[1] /= n; a[2] /= n;
}
inline void mul3(const double *A, const double *B, double *C) {
  double o[9];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) o[c * 3 + r] = A[r] * B[c * 3] + A[3 + r] * B[c * 3 + 1] + A[6 + r] * B[c * 3 + 2];
  memcpy(C, o, sizeof(o));
}
// a random rotation that tilts the local z axis onto `normal` within max_tilt radians
inline void facing(Rng &r, const double *normal, double max_tilt, double *R) {
  double z[3] = {normal[0], normal[1], normal[2]};
  double t[3] = {1, 0, 0};
  if (fabs(z[0]) > 0.9) { t[0] = 0; t[1] = 1; }
  double x[3] = {t[1] * z[2] - t[2] * z[1], t[2] * z[0] - t[0] * z[2], t[0] * z[1] - t[1] * z[0]};
  double nx = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  for (int i = 0; i < 3; ++i) x[i] /= nx;
  double y[3] = {z[1] * x[2] - z[2] * x[1], z[2] * x[0] - z[0] * x[2], z[0] * x[1] - z[1] * x[0]};
  double B[9] = {x[0], x[1], x[2], y[0], y[1], y[2], z[0], z[1], z[2]};
  double ax[3];
  random_axis(r, ax);
  double Rt[9];
  rot_from_axis_angle(ax, r.u(0.0, max_tilt), Rt);
  mul3(Rt, B, R);
}

// the 20 primitives of scene `seed` for a camera of the given field of view
inline void make_scene(uint64_t seed, double fx, double fy, int cols, int rows, std::vector<Prim> &prims) {
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
  for (int i = 3; i < kPrims; ++i) {
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

// rigid motion of pair `seed`: uniform axis, angle uniform in [0, max_rot], translation
// uniform in the ball of radius max_trans; T column-major float[16]
inline void motion(uint64_t seed, float max_rot_deg, float max_trans_m, float *T) {
  Rng r{seed, 2};
  double ax[3];
  random_axis(r, ax);
  double ang = r.u(0.0, max_rot_deg * M_PI / 180.0);
  double R[9];
  rot_from_axis_angle(ax, ang, R);
  double dir[3];
  random_axis(r, dir);
  double rad = max_trans_m * cbrt(r.u());
  for (int i = 0; i < 16; ++i) T[i] = 0.0f;
  for (int c = 0; c < 3; ++c)
    for (int rr = 0; rr < 3; ++rr) T[c * 4 + rr] = (float)R[c * 3 + rr];
  for (int k = 0; k < 3; ++k) T[12 + k] = (float)(dir[k] * rad);
  T[15] = 1.0f;
}

}  // namespace synth
