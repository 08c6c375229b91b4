/*
 * CPU oracle for the NICP registration loop — TEST INFRASTRUCTURE ONLY.
 *
 * A restatement, in plain C++, of the arithmetic of yorsh87/nicp on the path
 * depth -> cloud -> (project -> correspond -> linearize -> solve)^iters.
 * Each function cites the reference lines it follows (N/ = nicp/nicp/).
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
 * this library, as the checker or as the timed CPU reference. It is never a
 * product path.
 *
 * Build: oracle/build_oracle.sh (g++ -O3 -msse4.2 -ffp-contract=off -fopenmp).
 * -msse4.2 is the reference's Release flag (CMakeLists.txt:49): no FMA exists
 * in the reference binary, so every product and sum here is rounded on its own.
 *
 * Floating-point association orders follow the expression templates of the
 * Eigen version the reference targets (3.3, Ubuntu 18.04). Eigen is not
 * installed here, so these orders are restated from Eigen's published source
 * and are NOT checked against an Eigen build:
 *   P1  Matrix4f * Vector4f (lazy packet product): ((a0 v0 + a1 v1) + a2 v2) + a3 v3
 *   P3  Isometry3f * Vector4f (3x4 affine block, scalar redux): (a0 v0 + a1 v1) + (a2 v2 + a3 v3)
 *   P4  Vector4f dot / squaredNorm (SSE predux): (x0 y0 + x2 y2) + (x1 y1 + x3 y3)
 *   R3  3-term scalar redux (Matrix3f products, trace, cross-norms): x0 + (x1 + x2)
 *
 * Parity status per function (what pins it, besides the GPU comparison):
 *   A1 convert, A3 unproject/index image, A4 intervals, A6/A7 integral image and
 *   box query, A14 z-buffer, A15/A16 gates: pinned by closed forms and brute
 *   force in tests/test_oracle.py.
 *   A8/A9 eigensolver: restatement of Eigen's computeDirect; pinned to
 *   LAPACK (numpy.linalg.eigh) within tolerance, and by the FP64 companion.
 *   A10/A11 curvature, information matrices: closed forms.
 *   A17 linearizer: FP64 twin pinned by brute force and finite differences;
 *   the FP32 association order of the 4x4 products is parity unpinned.
 *   A19 LDLT (FP32 and FP64), v2t/t2v: pinned by numpy.linalg.solve and
 *   round trips; Eigen's LDLT summation order is parity unpinned.
 *   A20 statistics: invariants only (symmetry, ratio >= 1, degenerate scene);
 *   the JacobiSVD/LLT bit patterns are parity unpinned.
 */
#include "nicp_oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include <omp.h>

namespace {

inline float mc(const float *m, int r, int c) { return m[c * 4 + r]; }
inline float &mr(float *m, int r, int c) { return m[c * 4 + r]; }

/* P1 */
inline void mul4_packet(const float *A, const float *v, float *out) {
  float o[4];
  for (int i = 0; i < 4; ++i) {
    float r = mc(A, i, 0) * v[0];
    r = mc(A, i, 1) * v[1] + r;
    r = mc(A, i, 2) * v[2] + r;
    r = mc(A, i, 3) * v[3] + r;
    o[i] = r;
  }
  std::memcpy(out, o, sizeof(o));
}

/* P3: Eigen's transform_right_product_impl for a homogeneous 4-vector */
inline void iso_mul4(const float *A, const float *v, float *out) {
  float o[4];
  for (int i = 0; i < 3; ++i)
    o[i] = (mc(A, i, 0) * v[0] + mc(A, i, 1) * v[1]) + (mc(A, i, 2) * v[2] + mc(A, i, 3) * v[3]);
  o[3] = v[3];
  std::memcpy(out, o, sizeof(o));
}

/* P4 */
inline float dot4(const float *a, const float *b) {
  return (a[0] * b[0] + a[2] * b[2]) + (a[1] * b[1] + a[3] * b[3]);
}

/* R3 */
inline float red3(float x0, float x1, float x2) { return x0 + (x1 + x2); }

void set_identity4(float *m) {
  std::memset(m, 0, 16 * sizeof(float));
  m[0] = m[5] = m[10] = m[15] = 1.0f;
}

/* Eigen Transform<float,3,Isometry>::inverse(): R^T, -(R^T t) */
void iso_inverse(const float *A, float *B) {
  float o[16];
  set_identity4(o);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) mr(o, i, j) = mc(A, j, i);
  for (int i = 0; i < 3; ++i) {
    float s = red3(mc(o, i, 0) * mc(A, 0, 3), mc(o, i, 1) * mc(A, 1, 3), mc(o, i, 2) * mc(A, 2, 3));
    mr(o, i, 3) = -s;
  }
  std::memcpy(B, o, sizeof(o));
}

/* Eigen transform_transform_product_impl (Isometry * Isometry) */
void iso_mul(const float *A, const float *B, float *C) {
  float o[16];
  set_identity4(o);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      mr(o, i, j) = red3(mc(A, i, 0) * mc(B, 0, j), mc(A, i, 1) * mc(B, 1, j), mc(A, i, 2) * mc(B, 2, j));
  for (int i = 0; i < 3; ++i) {
    float s = red3(mc(A, i, 0) * mc(B, 0, 3), mc(A, i, 1) * mc(B, 1, 3), mc(A, i, 2) * mc(B, 2, 3));
    mr(o, i, 3) = s + mc(A, i, 3);
  }
  std::memcpy(C, o, sizeof(o));
}

/* ---------------- Eigen SelfAdjointEigenSolver<Matrix3f>::computeDirect ----------------
 * Restated from Eigen 3.3's published algorithm (third-party, not vendored in the
 * reference): shift by trace/3, scale by max |coeff|, trigonometric roots of the
 * characteristic cubic, eigenvectors from cross products of the shifted rows with
 * the most distinct eigenvalue first. Called at N/statscalculatorintegralimage.cpp:60-61. */
inline float k3(const float *m, int r, int c) { return m[c * 3 + r]; }
inline float &k3r(float *m, int r, int c) { return m[c * 3 + r]; }

void eig_compute_roots(const float *m, float *roots) {
  const float s_inv3 = 1.0f / 3.0f;
  const float s_sqrt3 = std::sqrt(3.0f);
  float m00 = k3(m, 0, 0), m11 = k3(m, 1, 1), m22 = k3(m, 2, 2);
  float m10 = k3(m, 1, 0), m20 = k3(m, 2, 0), m21 = k3(m, 2, 1);
  float c0 = m00 * m11 * m22 + 2.0f * m10 * m20 * m21 - m00 * m21 * m21 - m11 * m20 * m20 - m22 * m10 * m10;
  float c1 = m00 * m11 - m10 * m10 + m00 * m22 - m20 * m20 + m11 * m22 - m21 * m21;
  float c2 = m00 + m11 + m22;
  float c2_over_3 = c2 * s_inv3;
  float a_over_3 = (c2 * c2_over_3 - c1) * s_inv3;
  a_over_3 = std::max(a_over_3, 0.0f);
  float half_b = 0.5f * (c0 + c2_over_3 * (2.0f * c2_over_3 * c2_over_3 - c1));
  float q = a_over_3 * a_over_3 * a_over_3 - half_b * half_b;
  q = std::max(q, 0.0f);
  float rho = std::sqrt(a_over_3);
  float theta = std::atan2(std::sqrt(q), half_b) * s_inv3;
  float cos_theta = std::cos(theta);
  float sin_theta = std::sin(theta);
  roots[0] = c2_over_3 - rho * (cos_theta + s_sqrt3 * sin_theta);
  roots[1] = c2_over_3 - rho * (cos_theta - s_sqrt3 * sin_theta);
  roots[2] = c2_over_3 + 2.0f * rho * cos_theta;
}

inline void cross3(const float *a, const float *b, float *o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
inline float sqn3(const float *a) { return red3(a[0] * a[0], a[1] * a[1], a[2] * a[2]); }

void eig_extract_kernel(const float *mat, float *res, float *representative) {
  int i0 = 0;
  float best = std::fabs(k3(mat, 0, 0));
  for (int i = 1; i < 3; ++i) {
    float v = std::fabs(k3(mat, i, i));
    if (v > best) { best = v; i0 = i; }
  }
  const float *rep = mat + 3 * i0;
  float repc[3] = {rep[0], rep[1], rep[2]};
  float c0[3], c1[3];
  cross3(repc, mat + 3 * ((i0 + 1) % 3), c0);
  cross3(repc, mat + 3 * ((i0 + 2) % 3), c1);
  float n0 = sqn3(c0), n1 = sqn3(c1);
  if (n0 > n1) {
    float s = std::sqrt(n0);
    for (int i = 0; i < 3; ++i) res[i] = c0[i] / s;
  } else {
    float s = std::sqrt(n1);
    for (int i = 0; i < 3; ++i) res[i] = c1[i] / s;
  }
  if (representative)
    for (int i = 0; i < 3; ++i) representative[i] = repc[i];
}

void eig3_direct(const float *mat_in, float *evals, float *evecs) {
  float mat[9];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) k3r(mat, r, c) = (r >= c) ? k3(mat_in, r, c) : k3(mat_in, c, r);
  float shift = red3(k3(mat, 0, 0), k3(mat, 1, 1), k3(mat, 2, 2)) / 3.0f;
  float scaled[9];
  std::memcpy(scaled, mat, sizeof(scaled));
  for (int i = 0; i < 3; ++i) k3r(scaled, i, i) -= shift;
  float scale = 0.0f;
  for (int i = 0; i < 9; ++i) scale = std::max(scale, std::fabs(scaled[i]));
  if (scale > 0.0f)
    for (int i = 0; i < 9; ++i) scaled[i] /= scale;
  eig_compute_roots(scaled, evals);
  const float eps = FLT_EPSILON;
  if ((evals[2] - evals[0]) <= eps) {
    std::memset(evecs, 0, 9 * sizeof(float));
    evecs[0] = evecs[4] = evecs[8] = 1.0f;
  } else {
    float tmp[9];
    std::memcpy(tmp, scaled, sizeof(tmp));
    float d0 = evals[2] - evals[1];
    float d1 = evals[1] - evals[0];
    int k = 0, l = 2;
    if (d0 > d1) { std::swap(k, l); d0 = d1; }
    for (int i = 0; i < 3; ++i) k3r(tmp, i, i) -= evals[k];
    eig_extract_kernel(tmp, evecs + 3 * k, evecs + 3 * l);
    if (d0 <= 2.0f * eps * d1) {
      /* Eigen 3.3 form of the near-degenerate branch (SelfAdjointEigenSolver.h) */
      float *vl = evecs + 3 * l, *vk = evecs + 3 * k;
      float d = red3(vk[0] * vl[0], vk[1] * vl[1], vk[2] * vl[2]);
      float t[3] = {d * vl[0], d * vl[1], d * vl[2]};
      for (int i = 0; i < 3; ++i) vl[i] -= t[i];
      float z = sqn3(vl);
      if (z > 0.0f) {
        float s = std::sqrt(z);
        for (int i = 0; i < 3; ++i) vl[i] /= s;
      }
    } else {
      std::memcpy(tmp, scaled, sizeof(tmp));
      for (int i = 0; i < 3; ++i) k3r(tmp, i, i) -= evals[l];
      eig_extract_kernel(tmp, evecs + 3 * l, nullptr);
    }
    float c[3];
    cross3(evecs + 6, evecs + 0, c);
    float z = sqn3(c);
    if (z > 0.0f) {
      float s = std::sqrt(z);
      for (int i = 0; i < 3; ++i) c[i] /= s;
    }
    std::memcpy(evecs + 3, c, sizeof(c));
  }
  for (int i = 0; i < 3; ++i) evals[i] = evals[i] * scale;
  for (int i = 0; i < 3; ++i) evals[i] = evals[i] + shift;
}

/* N/stats.h:93-98 — the 1e-9 literal promotes the sum's last add and the divide to double */
inline float stats_curvature(const float *ev) {
  return (float)((double)ev[0] / ((double)((ev[0] + ev[1]) + ev[2]) + 1e-9));
}

/* cyclic Jacobi eigen-decomposition of a symmetric matrix (FP64 companion) */
void jacobi_eigen(int n, double *a /* n*n col-major, destroyed */, double *evals, double *evecs) {
  for (int i = 0; i < n * n; ++i) evecs[i] = 0.0;
  for (int i = 0; i < n; ++i) evecs[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int c = 0; c < n; ++c)
      for (int r = 0; r < n; ++r) {
        tot += a[c * n + r] * a[c * n + r];
        if (r != c) off += a[c * n + r] * a[c * n + r];
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = a[q * n + p];
        if (apq == 0.0) continue;
        double app = a[p * n + p], aqq = a[q * n + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          double akp = a[p * n + k], akq = a[q * n + k];
          a[p * n + k] = c * akp - s * akq;
          a[q * n + k] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = a[k * n + p], aqk = a[k * n + q];
          a[k * n + p] = c * apk - s * aqk;
          a[k * n + q] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = evecs[p * n + k], vkq = evecs[q * n + k];
          evecs[p * n + k] = c * vkp - s * vkq;
          evecs[q * n + k] = s * vkp + c * vkq;
        }
      }
  }
  /* sort ascending */
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::vector<double> d(n);
  for (int i = 0; i < n; ++i) d[i] = a[i * n + i];
  std::sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
  std::vector<double> v(evecs, evecs + n * n);
  for (int i = 0; i < n; ++i) {
    evals[i] = d[idx[i]];
    for (int k = 0; k < n; ++k) evecs[i * n + k] = v[idx[i] * n + k];
  }
}

/* ---------------- SE(3), N/bm_se3.h:10-66 ---------------- */
template <typename S>
void quat2mat(const S *q, S *R /* 3x3 col-major */) {
  S qx = q[0], qy = q[1], qz = q[2];
  S qw = std::sqrt(S(1) - (qx * qx + (qy * qy + qz * qz)));
  auto at = [&](int r, int c) -> S & { return R[c * 3 + r]; };
  at(0, 0) = qw * qw + qx * qx - qy * qy - qz * qz;
  at(0, 1) = S(2) * (qx * qy - qw * qz);
  at(0, 2) = S(2) * (qx * qz + qw * qy);
  at(1, 0) = S(2) * (qx * qy + qz * qw);
  at(1, 1) = qw * qw - qx * qx + qy * qy - qz * qz;
  at(1, 2) = S(2) * (qy * qz - qx * qw);
  at(2, 0) = S(2) * (qx * qz - qy * qw);
  at(2, 1) = S(2) * (qy * qz + qx * qw);
  at(2, 2) = qw * qw - qx * qx - qy * qy + qz * qz;
}

/* Eigen Quaternion(const Matrix3&) (Shepperd's method), then normalize, w >= 0 */
template <typename S>
void mat2quat(const S *R /* 3x3 col-major */, S *rq) {
  auto m = [&](int r, int c) { return R[c * 3 + r]; };
  S q[4];  // x y z w
  S t = m(0, 0) + (m(1, 1) + m(2, 2));
  if (t > S(0)) {
    t = std::sqrt(t + S(1));
    q[3] = S(0.5) * t;
    t = S(0.5) / t;
    q[0] = (m(2, 1) - m(1, 2)) * t;
    q[1] = (m(0, 2) - m(2, 0)) * t;
    q[2] = (m(1, 0) - m(0, 1)) * t;
  } else {
    int i = 0;
    if (m(1, 1) > m(0, 0)) i = 1;
    if (m(2, 2) > m(i, i)) i = 2;
    int j = (i + 1) % 3, k = (j + 1) % 3;
    t = std::sqrt(m(i, i) - m(j, j) - m(k, k) + S(1));
    q[i] = S(0.5) * t;
    t = S(0.5) / t;
    q[3] = (m(k, j) - m(j, k)) * t;
    q[j] = (m(j, i) + m(i, j)) * t;
    q[k] = (m(k, i) + m(i, k)) * t;
  }
  S z = (q[0] * q[0] + q[2] * q[2]) + (q[1] * q[1] + q[3] * q[3]);
  if (z > S(0)) {
    S s = std::sqrt(z);
    for (int i = 0; i < 4; ++i) q[i] = q[i] / s;
  }
  rq[0] = q[0]; rq[1] = q[1]; rq[2] = q[2];
  if (q[3] < S(0)) { rq[0] = -rq[0]; rq[1] = -rq[1]; rq[2] = -rq[2]; }
}

template <typename S>
void v2t_t(const S *x, S *T /* 4x4 col-major */) {
  S R[9];
  quat2mat(x + 3, R);
  for (int i = 0; i < 16; ++i) T[i] = S(0);
  T[15] = S(1);
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) T[c * 4 + r] = R[c * 3 + r];
  T[12] = x[0]; T[13] = x[1]; T[14] = x[2];
}

template <typename S>
void t2v_t(const S *T, S *x) {
  S R[9];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) R[c * 3 + r] = T[c * 4 + r];
  x[0] = T[12]; x[1] = T[13]; x[2] = T[14];
  mat2quat(R, x + 3);
}

/* Eigen LDLT<Matrix6> with diagonal pivoting (ldlt_inplace<Lower>::unblocked) and the
 * pseudo-inverse-of-D solve. Summation orders are sequential. */
template <typename S>
int ldlt_solve(const S *Hin, const S *rhs, S *x) {
  const int n = 6;
  S m[36];
  std::memcpy(m, Hin, sizeof(m));
  auto A = [&](int r, int c) -> S & { return m[c * n + r]; };
  int tr[6];
  S temp[6];
  int ok = 1;
  for (int k = 0; k < n; ++k) {
    int idx = k;
    S best = std::fabs(A(k, k));
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A(i, i)) > best) { best = std::fabs(A(i, i)); idx = i; }
    tr[k] = idx;
    if (k != idx) {
      for (int j = 0; j < k; ++j) std::swap(A(k, j), A(idx, j));
      for (int i = idx + 1; i < n; ++i) std::swap(A(i, k), A(i, idx));
      std::swap(A(k, k), A(idx, idx));
      for (int i = k + 1; i < idx; ++i) {
        S t = A(i, k);
        A(i, k) = A(idx, i);
        A(idx, i) = t;
      }
    }
    int rs = n - k - 1;
    if (k > 0) {
      for (int j = 0; j < k; ++j) temp[j] = A(j, j) * A(k, j);
      S s = S(0);
      for (int j = 0; j < k; ++j) s = (j == 0) ? A(k, j) * temp[j] : s + A(k, j) * temp[j];
      A(k, k) -= s;
      for (int i = k + 1; i < n; ++i) {
        S t = S(0);
        for (int j = 0; j < k; ++j) t = (j == 0) ? A(i, j) * temp[j] : t + A(i, j) * temp[j];
        A(i, k) -= t;
      }
    }
    S akk = A(k, k);
    bool valid = std::fabs(akk) > S(0);
    if (k == 0 && !valid) {
      for (int j = 0; j < n; ++j) tr[j] = j;
      ok = 0;
      break;
    }
    if (rs > 0 && valid)
      for (int i = k + 1; i < n; ++i) A(i, k) /= akk;
  }
  S d[6];
  for (int i = 0; i < n; ++i) d[i] = rhs[i];
  for (int i = 0; i < n; ++i) std::swap(d[i], d[tr[i]]);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) d[i] -= A(i, j) * d[j];
  const S tol = std::numeric_limits<S>::min();
  for (int i = 0; i < n; ++i) {
    if (std::fabs(A(i, i)) > tol) d[i] /= A(i, i);
    else d[i] = S(0);
  }
  for (int i = n - 1; i >= 0; --i)
    for (int j = i + 1; j < n; ++j) d[i] -= A(j, i) * d[j];
  for (int i = n - 1; i >= 0; --i) std::swap(d[i], d[tr[i]]);
  for (int i = 0; i < n; ++i) x[i] = d[i];
  return ok;
}

/* N/bm_se3.h:55-66: skew(v) = -2 [v]x, embedded in a 4x4 */
void skew4(const float *v, float *S) {
  float tx = 2.0f * v[0], ty = 2.0f * v[1], tz = 2.0f * v[2];
  std::memset(S, 0, 16 * sizeof(float));
  mr(S, 0, 1) = tz;  mr(S, 1, 0) = -tz;
  mr(S, 0, 2) = -ty; mr(S, 2, 0) = ty;
  mr(S, 1, 2) = tx;  mr(S, 2, 1) = -tx;
}

/* lazy packet product of two col-major 4x4 (lhs vectorizable) */
void mul44_packet(const float *A, const float *B, float *C) {
  float o[16];
  for (int j = 0; j < 4; ++j) mul4_packet(A, B + 4 * j, o + 4 * j);
  std::memcpy(C, o, sizeof(o));
}
/* transpose(A) * B for col-major 4x4: coefficient path with vectorised redux (P4) */
void mul44_tn(const float *A, const float *B, float *C) {
  float o[16];
  for (int j = 0; j < 4; ++j)
    for (int i = 0; i < 4; ++i) o[j * 4 + i] = dot4(A + 4 * i, B + 4 * j);
  std::memcpy(C, o, sizeof(o));
}
void mul4_tn(const float *A, const float *v, float *out) {
  float o[4];
  for (int i = 0; i < 4; ++i) o[i] = dot4(A + 4 * i, v);
  std::memcpy(out, o, sizeof(o));
}

/* JacobiSVD-based pseudo-inverse of a symmetric 6x6 (FP64), threshold as Eigen's
 * default for Matrix6f: 6 * FLT_EPSILON relative to the largest singular value. */
void pinv6_sym(const double *H, double *P) {
  double a[36], ev[6], V[36];
  std::memcpy(a, H, sizeof(a));
  jacobi_eigen(6, a, ev, V);
  double smax = 0.0;
  for (int i = 0; i < 6; ++i) smax = std::max(smax, std::fabs(ev[i]));
  double thr = std::max(smax * 6.0 * (double)FLT_EPSILON, (double)std::numeric_limits<double>::min());
  for (int i = 0; i < 36; ++i) P[i] = 0.0;
  for (int k = 0; k < 6; ++k) {
    if (std::fabs(ev[k]) < thr) continue;
    double inv = 1.0 / ev[k];
    for (int c = 0; c < 6; ++c)
      for (int r = 0; r < 6; ++r) P[c * 6 + r] += V[k * 6 + r] * inv * V[k * 6 + c];
  }
}

bool llt6(const double *A, double *L) {
  for (int i = 0; i < 36; ++i) L[i] = 0.0;
  for (int j = 0; j < 6; ++j) {
    double s = A[j * 6 + j];
    for (int k = 0; k < j; ++k) s -= L[k * 6 + j] * L[k * 6 + j];
    if (!(s > 0.0)) return false;
    double d = std::sqrt(s);
    L[j * 6 + j] = d;
    for (int i = j + 1; i < 6; ++i) {
      double t = A[j * 6 + i];
      for (int k = 0; k < j; ++k) t -= L[k * 6 + i] * L[k * 6 + j];
      L[j * 6 + i] = t / d;
    }
  }
  return true;
}

bool inverse6(const double *A, double *Ai) {
  double m[36], inv[36];
  std::memcpy(m, A, sizeof(m));
  for (int i = 0; i < 36; ++i) inv[i] = 0.0;
  for (int i = 0; i < 6; ++i) inv[i * 6 + i] = 1.0;
  for (int c = 0; c < 6; ++c) {
    int p = c;
    for (int r = c + 1; r < 6; ++r)
      if (std::fabs(m[c * 6 + r]) > std::fabs(m[c * 6 + p])) p = r;
    if (m[c * 6 + p] == 0.0) return false;
    if (p != c)
      for (int k = 0; k < 6; ++k) { std::swap(m[k * 6 + p], m[k * 6 + c]); std::swap(inv[k * 6 + p], inv[k * 6 + c]); }
    double d = m[c * 6 + c];
    for (int k = 0; k < 6; ++k) { m[k * 6 + c] /= d; inv[k * 6 + c] /= d; }
    for (int r = 0; r < 6; ++r) {
      if (r == c) continue;
      double f = m[c * 6 + r];
      if (f == 0.0) continue;
      for (int k = 0; k < 6; ++k) { m[k * 6 + r] -= f * m[k * 6 + c]; inv[k * 6 + r] -= f * inv[k * 6 + c]; }
    }
  }
  std::memcpy(Ai, inv, sizeof(inv));
  return true;
}

/* singular-value ratio s_max / s_min of a symmetric 3x3 block */
double sv_ratio3(const double *O6, int off) {
  double a[9], ev[3], V[9];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) a[c * 3 + r] = O6[(off + c) * 6 + off + r];
  jacobi_eigen(3, a, ev, V);
  double s[3] = {std::fabs(ev[0]), std::fabs(ev[1]), std::fabs(ev[2])};
  std::sort(s, s + 3);
  return s[2] / s[0];
}

}  // namespace

extern "C" {

/* A1: N/imageutils.cpp:85-99 */
void orc_depth_u16_to_f32(const uint16_t *raw, int n, float scale, float *out) {
  for (int i = 0; i < n; ++i) out[i] = raw[i] ? scale * (float)raw[i] : 0.0f;
}

/* DepthImage_scale, N/imageutils.cpp:6-37 */
void orc_depth_scale(const float *src, int rows, int cols, int step, float max_cov, float *dst) {
  const int orows = rows / step, ocols = cols / step;
  for (int r = 0; r < orows; ++r)
    for (int c = 0; c < ocols; ++c) {
      float acc = 0.0f, acc2 = 0.0f;
      int np = 0;
      const int sr = r * step, sc = c * step;
      for (int i = 0; i < step; ++i)
        for (int j = 0; j < step; ++j)
          if (sr + i < rows && sc + j < cols) {
            const float f = src[(size_t)(sr + i) * cols + sc + j];
            acc += f;
            acc2 += f * f;
            np += f > 0;
          }
      float out = 0.0f;
      if (np) {
        float mu = acc / np;
        float sigma = acc2 / np - mu * mu;
        if (!(sigma > max_cov)) out = mu;
      }
      dst[(size_t)r * ocols + c] = out;
    }
}

/* A2: Eigen's 3x3 cofactor inverse (compute_inverse<...,3>), called at
 * N/pinholepointprojector.cpp:19-20 */
void orc_inverse3(const float *K, float *iK) {
  auto m = [&](int r, int c) { return K[c * 3 + r]; };
  auto cof = [&](int i, int j) {
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return m(i1, j1) * m(i2, j2) - m(i1, j2) * m(i2, j1);
  };
  float c0[3] = {cof(0, 0), cof(1, 0), cof(2, 0)};
  float det = red3(c0[0] * m(0, 0), c0[1] * m(1, 0), c0[2] * m(2, 0));
  float invdet = 1.0f / det;
  float r[9];
  auto at = [&](int rr, int cc) -> float & { return r[cc * 3 + rr]; };
  at(0, 0) = c0[0] * invdet; at(0, 1) = c0[1] * invdet; at(0, 2) = c0[2] * invdet;
  at(1, 0) = cof(0, 1) * invdet; at(1, 1) = cof(1, 1) * invdet; at(1, 2) = cof(2, 1) * invdet;
  at(2, 0) = cof(0, 2) * invdet; at(2, 1) = cof(1, 2) * invdet; at(2, 2) = cof(2, 2) * invdet;
  std::memcpy(iK, r, sizeof(r));
}

/* A2: PinholePointProjector::_updateMatrices, N/pinholepointprojector.cpp:16-31 */
void orc_projector_matrices(const float *K, const float *transform, float *KRt, float *iKRt) {
  float t[16];
  iso_inverse(transform, t);
  float iK[9];
  orc_inverse3(K, iK);
  set_identity4(KRt);
  set_identity4(iKRt);
  auto k = [&](int r, int c) { return K[c * 3 + r]; };
  auto ik = [&](int r, int c) { return iK[c * 3 + r]; };
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      mr(KRt, i, j) = red3(k(i, 0) * mc(t, 0, j), k(i, 1) * mc(t, 1, j), k(i, 2) * mc(t, 2, j));
      mr(iKRt, i, j) = red3(mc(transform, i, 0) * ik(0, j), mc(transform, i, 1) * ik(1, j),
                            mc(transform, i, 2) * ik(2, j));
    }
    mr(KRt, i, 3) = red3(k(i, 0) * mc(t, 0, 3), k(i, 1) * mc(t, 1, 3), k(i, 2) * mc(t, 2, 3));
    mr(iKRt, i, 3) = mc(transform, i, 3);
  }
}

/* A3: PinholePointProjector::unProject, N/pinholepointprojector.cpp:93-134 and
 * _unProject, N/pinholepointprojector.h:275-280 (x = column, y = row). */
int orc_unproject(const float *depth, int rows, int cols, const float *iKRt, float min_d, float max_d,
                  float *points, int *index_image) {
  int count = 0;
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c) {
      float d = depth[r * cols + c];
      if (d == 0.0f || d < min_d || d > max_d) {
        index_image[r * cols + c] = -1;
        continue;
      }
      float v[4] = {(float)c * d, (float)r * d, d, 1.0f};
      mul4_packet(iKRt, v, points + 4 * count);
      index_image[r * cols + c] = count;
      ++count;
    }
  return count;
}

/* A4: projectIntervals, N/pinholepointprojector.cpp:136-148 + _projectInterval (.h:293-302) */
void orc_project_intervals(const float *depth, int rows, int cols, const float *K, float min_d, float max_d,
                           float world_radius, int *intervals) {
  auto k = [&](int r, int c) { return K[c * 3 + r]; };
  float p0 = red3(k(0, 0) * world_radius, k(0, 1) * world_radius, k(0, 2) * 0.0f);
  float p1 = red3(k(1, 0) * world_radius, k(1, 1) * world_radius, k(1, 2) * 0.0f);
  for (int i = 0; i < rows * cols; ++i) {
    float d = depth[i];
    if (d < min_d || d > max_d) {
      intervals[2 * i] = intervals[2 * i + 1] = -1;
      continue;
    }
    float inv = 1.0f / d;
    float a = p0 * inv, b = p1 * inv;
    float ms = std::max(a, b);
    intervals[2 * i] = intervals[2 * i + 1] = (int)ms;
  }
}

/* A6: PointIntegralImage::compute, N/pointintegralimage.cpp:7-44: place (p, p p^T),
 * prefix along x within each image row, then along y within each column. */
void orc_integral_image(const int *index_image, int rows, int cols, const float *points, float *ii) {
  const size_t P = (size_t)rows * cols;
  for (size_t i = 0; i < P; ++i) {
    float *o = ii + 10 * i;
    int idx = index_image[i];
    if (idx < 0) {
      for (int k = 0; k < 10; ++k) o[k] = 0.0f;
      continue;
    }
    const float *p = points + 4 * (size_t)idx;
    o[0] = p[3] * p[3];
    o[1] = p[0]; o[2] = p[1]; o[3] = p[2];
    o[4] = p[0] * p[0]; o[5] = p[0] * p[1]; o[6] = p[0] * p[2];
    o[7] = p[1] * p[1]; o[8] = p[1] * p[2]; o[9] = p[2] * p[2];
  }
#pragma omp parallel for schedule(static)
  for (int r = 0; r < rows; ++r)
    for (int c = 1; c < cols; ++c)
      for (int k = 0; k < 10; ++k)
        ii[10 * ((size_t)r * cols + c) + k] = ii[10 * ((size_t)r * cols + c) + k] + ii[10 * ((size_t)r * cols + c - 1) + k];
#pragma omp parallel for schedule(static)
  for (int c = 0; c < cols; ++c)
    for (int r = 1; r < rows; ++r)
      for (int k = 0; k < 10; ++k)
        ii[10 * ((size_t)r * cols + c) + k] = ii[10 * ((size_t)r * cols + c) + k] + ii[10 * ((size_t)(r - 1) * cols + c) + k];
}

/* A7: PointIntegralImage::getRegion, N/pointintegralimage.cpp:53-66 (x = column) */
void orc_get_region(const float *ii, int rows, int cols, int xmin, int xmax, int ymin, int ymax, float *out) {
  auto clampi = [](int v, int lo, int hi) { v = v < lo ? lo : v; v = v > hi ? hi : v; return v; };
  xmin = clampi(xmin - 1, 0, cols - 1);
  xmax = clampi(xmax - 1, 0, cols - 1);
  ymin = clampi(ymin - 1, 0, rows - 1);
  ymax = clampi(ymax - 1, 0, rows - 1);
  const float *A = ii + 10 * ((size_t)ymax * cols + xmax);
  const float *D = ii + 10 * ((size_t)ymin * cols + xmin);
  const float *B = ii + 10 * ((size_t)ymax * cols + xmin);
  const float *C = ii + 10 * ((size_t)ymin * cols + xmax);
  for (int k = 0; k < 10; ++k) out[k] = ((A[k] + D[k]) - B[k]) - C[k];
}

void orc_eigen3_direct(const float *cov, float *evals, float *evecs) { eig3_direct(cov, evals, evecs); }

/* A5: DepthImageConverterIntegralImage::compute, N/depthimageconverterintegralimage.cpp:18-64,
 * with A8 (N/pointaccumulator.h:66-110), A9 (N/statscalculatorintegralimage.cpp:20-82),
 * A10 (N/stats.h:93-98), A11 (N/informationmatrixcalculator.cpp:9-57). Sensor offset identity. */
/* ---------------- SphericalPointProjector ----------------
 * N/sphericalpointprojector.cpp:19-27 (_updateParameters) at the image size of the call,
 * N/sphericalpointprojector.h:252-303 (_project, _unProject, _projectInterval). The float
 * overloads of sin, cos, tanf, atan2 and sqrt are the reference's calls. Quirks kept: the field
 * of view test of _project is fabs(theta > fov), so angles below -fov pass; pixels are
 * truncated toward zero. */
struct SphRes {
  float hres, ihres, vres, ivres;
};
static SphRes sph_res(const orc_camera *cam, int rows, int cols) {
  SphRes s;
  s.hres = (float)cols / (2.0f * cam->horizontal_fov);
  s.ihres = 1.0f / s.hres;
  s.vres = (float)rows / (2.0f * cam->vertical_fov);
  s.ivres = 1.0f / s.vres;
  return s;
}

/* _project: p seen through iT (P3), returns false where the reference does */
static bool sph_project1(const orc_camera *cam, const SphRes &s, const float *iT, const float *p, int &x, int &y,
                         float &d) {
  float cp[4];
  iso_mul4(iT, p, cp);
  const float theta = std::atan2(cp[0], cp[2]);
  d = std::sqrt(cp[0] * cp[0] + cp[2] * cp[2]);
  const float phi = std::atan2(cp[1], d);
  if (d > cam->max_distance || d < cam->min_distance) return false;
  if (std::fabs((float)(theta > cam->horizontal_fov))) return false;
  if (std::fabs((float)(phi > cam->vertical_fov))) return false;
  x = (int)(s.hres * theta + cam->horizontal_center);
  y = (int)(s.vres * phi + cam->vertical_center);
  return true;
}

int orc_spherical_unproject(const float *depth, int rows, int cols, const orc_camera *cam, const float *T,
                            float *points, int *index_image) {
  const SphRes s = sph_res(cam, rows, cols);
  int count = 0;
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c) {
      const float d = depth[r * cols + c];
      index_image[r * cols + c] = -1;
      if (d < cam->min_distance || d > cam->max_distance) continue;
      const float theta = s.ihres * (c - cam->horizontal_center);
      const float phi = s.ivres * (r - cam->vertical_center);
      if (std::fabs(theta) > cam->horizontal_fov) continue;
      if (std::fabs(phi) > cam->vertical_fov) continue;
      const float x = std::sin(theta) * d;
      const float z = std::cos(theta) * d;
      const float y = d * tanf(phi);
      float *o = points + 4 * (size_t)count;
      for (int i = 0; i < 3; ++i) o[i] = mc(T, i, 3) + red3(mc(T, i, 0) * x, mc(T, i, 1) * y, mc(T, i, 2) * z);
      o[3] = 1.0f;
      index_image[r * cols + c] = count++;
    }
  return count;
}

void orc_spherical_intervals(const float *depth, int rows, int cols, const orc_camera *cam, float world_radius,
                             int *intervals) {
  const SphRes s = sph_res(cam, rows, cols);
  for (int i = 0; i < rows * cols; ++i) {
    const float d = depth[i];
    if (d > cam->max_distance || d < cam->min_distance) {
      intervals[2 * i] = intervals[2 * i + 1] = -1;
      continue;
    }
    const float theta = std::atan2(world_radius, d);
    const float phi = std::atan2(world_radius, d);
    intervals[2 * i] = (int)(s.hres * theta);
    intervals[2 * i + 1] = (int)(s.vres * phi);
  }
}

void orc_spherical_project(const float *points, int n, const float *transform, const orc_camera *cam, int rows,
                           int cols, int *index_image, float *depth_image) {
  const SphRes s = sph_res(cam, rows, cols);
  float t[16], iT[16];
  std::memcpy(t, transform, sizeof(t));
  mr(t, 3, 0) = 0.0f; mr(t, 3, 1) = 0.0f; mr(t, 3, 2) = 0.0f; mr(t, 3, 3) = 1.0f;
  iso_inverse(t, iT);
  const size_t P = (size_t)rows * cols;
  std::vector<float> dimg(P, std::numeric_limits<float>::max());
  for (size_t i = 0; i < P; ++i) index_image[i] = -1;
  for (int i = 0; i < n; ++i) {
    int x, y;
    float d;
    if (!sph_project1(cam, s, iT, points + 4 * (size_t)i, x, y, d) || d < cam->min_distance ||
        d > cam->max_distance || x < 0 || x >= cols || y < 0 || y >= rows)
      continue;
    float &other = dimg[(size_t)y * cols + x];
    if (!other || other > d) {
      other = d;
      index_image[(size_t)y * cols + x] = i;
    }
  }
  if (depth_image) std::memcpy(depth_image, dimg.data(), P * sizeof(float));
}

/* the projection of a cloud from `transform` by either projector */
static void project_cam(const orc_camera *cam, const float *transform, const float *points, int n, int rows, int cols,
                        int *index_image, float *depth_image) {
  if (cam->model == 1) {
    orc_spherical_project(points, n, transform, cam, rows, cols, index_image, depth_image);
    return;
  }
  float KRt[16], iKRt[16];
  orc_projector_matrices(cam->K, transform, KRt, iKRt);
  orc_project(points, n, KRt, rows, cols, cam->min_distance, cam->max_distance, index_image, depth_image);
}

int orc_frame_from_depth(const uint16_t *raw, int rows, int cols, float depth_scale, const orc_camera *cam,
                         const orc_stats_params *sp, int *index_image, float *points, float *normals,
                         float *curvature, float *eigvecs, float *eigvals, float *means, int *npts,
                         float *omega_p, float *omega_n) {
  const size_t P = (size_t)rows * cols;
  std::vector<float> depth(P), pts(P * 4);
  std::vector<int> idx(P), iv(2 * P);
  orc_depth_u16_to_f32(raw, (int)P, depth_scale, depth.data());
  float I4[16], KRt[16], iKRt[16];
  set_identity4(I4);
  int n;
  /* the spherical projector suppresses the image-radius limits of the stats
   * (N/depthimageconverterintegralimage.cpp:48-52) */
  const bool suppress = cam->model == 1;
  if (suppress) {
    n = orc_spherical_unproject(depth.data(), rows, cols, cam, I4, pts.data(), idx.data());
    orc_spherical_intervals(depth.data(), rows, cols, cam, sp->world_radius, iv.data());
  } else {
    orc_projector_matrices(cam->K, I4, KRt, iKRt);
    n = orc_unproject(depth.data(), rows, cols, iKRt, cam->min_distance, cam->max_distance, pts.data(), idx.data());
    orc_project_intervals(depth.data(), rows, cols, cam->K, cam->min_distance, cam->max_distance, sp->world_radius,
                          iv.data());
  }
  std::vector<float> ii(P * 10);
  orc_integral_image(idx.data(), rows, cols, pts.data(), ii.data());

  /* per-point outputs, initialised as Stats() / zero normals */
  std::vector<float> nrm((size_t)n * 4, 0.0f), ev((size_t)n * 3, 0.0f), evec((size_t)n * 9, 0.0f),
      mean((size_t)n * 4, 0.0f), curv((size_t)n, 0.0f);
  std::vector<int> nn((size_t)n, 0);
  for (int i = 0; i < n; ++i) {
    evec[9 * i + 0] = evec[9 * i + 4] = evec[9 * i + 8] = 1.0f;
    mean[4 * i + 3] = 1.0f;
  }
#pragma omp parallel for schedule(static)
  for (int r = 0; r < rows; ++r) {
    for (int c = 0; c < cols; ++c) {
      size_t pix = (size_t)r * cols + c;
      int index = idx[pix];
      int R0 = iv[2 * pix], R1 = iv[2 * pix + 1];
      if (index < 0 || R0 < 0 || R1 < 0) continue;
      if (!suppress) {
        if (R0 < sp->min_image_radius || R1 < sp->min_image_radius) continue;
        if (R0 > sp->max_image_radius) R0 = sp->max_image_radius;
        if (R1 > sp->max_image_radius) R1 = sp->max_image_radius;
      }
      float acc[10];
      orc_get_region(ii.data(), rows, cols, c - R0, c + R0, r - R1, r + R1, acc);
      int accn = (int)acc[0];
      if (accn < sp->min_points) continue;
      /* covariance: S*d - mean*mean^T, d = 1/n (N/pointaccumulator.h:78-86) */
      float d = 1.0f / acc[0];
      float m[3] = {acc[1] * d, acc[2] * d, acc[3] * d};
      float S[9] = {acc[4], acc[5], acc[6], acc[5], acc[7], acc[8], acc[6], acc[8], acc[9]};
      float cov[9];
      for (int cc = 0; cc < 3; ++cc)
        for (int rr = 0; rr < 3; ++rr) cov[cc * 3 + rr] = S[cc * 3 + rr] * d - m[rr] * m[cc];
      float e[3], V[9];
      eig3_direct(cov, e, V);
      if (e[0] < 0.0f) e[0] = 0.0f;
      float *vo = &evec[9 * (size_t)index];
      std::memcpy(vo, V, sizeof(V));
      float *eo = &ev[3 * (size_t)index];
      std::memcpy(eo, e, sizeof(e));
      float invn = 1.0f / acc[0];
      float *mo = &mean[4 * (size_t)index];
      /* PointAccumulator::mean() returns a Point, whose constructor sets w = 1
       * (N/pointaccumulator.h:66-70, N/homogeneousvector4f.h:19) */
      mo[0] = acc[1] * invn; mo[1] = acc[2] * invn; mo[2] = acc[3] * invn; mo[3] = 1.0f;
      nn[index] = accn;
      float k = stats_curvature(e);
      curv[index] = k;
      float nv[4] = {V[0], V[1], V[2], 0.0f};
      if (k < sp->curvature_threshold) {
        if (dot4(nv, &pts[4 * (size_t)index]) > 0.0f) { nv[0] = -nv[0]; nv[1] = -nv[1]; nv[2] = -nv[2]; nv[3] = -nv[3]; }
      } else {
        nv[0] = nv[1] = nv[2] = nv[3] = 0.0f;
      }
      std::memcpy(&nrm[4 * (size_t)index], nv, sizeof(nv));
    }
  }
  /* curvature of untouched Stats(): eigenvalues 0 -> 0/(0+1e-9) = 0 */
  if (index_image) std::memcpy(index_image, idx.data(), P * sizeof(int));
  if (points) std::memcpy(points, pts.data(), (size_t)n * 4 * sizeof(float));
  if (normals) std::memcpy(normals, nrm.data(), (size_t)n * 4 * sizeof(float));
  if (curvature) std::memcpy(curvature, curv.data(), (size_t)n * sizeof(float));
  if (eigvecs) std::memcpy(eigvecs, evec.data(), (size_t)n * 9 * sizeof(float));
  if (eigvals) std::memcpy(eigvals, ev.data(), (size_t)n * 3 * sizeof(float));
  if (means) std::memcpy(means, mean.data(), (size_t)n * 4 * sizeof(float));
  if (npts) std::memcpy(npts, nn.data(), (size_t)n * sizeof(int));
  /* A11: Omega = U diag(d) U^T with the association order of Eigen's lazy products */
  for (int which = 0; which < 2; ++which) {
    float *out = which == 0 ? omega_p : omega_n;
    if (!out) continue;
    const float *flat = which == 0 ? sp->point_flat : sp->normal_flat;
    const float *nonflat = which == 0 ? sp->point_nonflat : sp->normal_nonflat;
    float thr = which == 0 ? sp->point_curvature_threshold : sp->normal_curvature_threshold;
    for (int i = 0; i < n; ++i) {
      float *O = out + 16 * (size_t)i;
      std::memset(O, 0, 16 * sizeof(float));
      const float *nv = &nrm[4 * (size_t)i];
      if (!(dot4(nv, nv) > 0.0f)) continue;
      const float *dg = curv[i] < thr ? flat : nonflat;
      const float *U = &evec[9 * (size_t)i];
      for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
          float a0 = (U[0 * 3 + r] * dg[0]) * U[0 * 3 + c];
          float a1 = (U[1 * 3 + r] * dg[1]) * U[1 * 3 + c];
          float a2 = (U[2 * 3 + r] * dg[2]) * U[2 * 3 + c];
          O[c * 4 + r] = (a0 + a1) + a2;
        }
    }
  }
  return n;
}

/* FP64 companion of A9: the same window, its covariance in double through a double
 * integral image (no FP32 cancellation), eigenvectors by Jacobi. */
void orc_stats_f64_bruteforce(const uint16_t *raw, int rows, int cols, float depth_scale, const orc_camera *cam,
                              const orc_stats_params *sp, const int *index_image, const float *points, int n,
                              double *normals64, double *curvature64, int *valid) {
  const size_t P = (size_t)rows * cols;
  std::vector<double> ii(P * 10, 0.0);
  for (size_t i = 0; i < P; ++i) {
    int idx = index_image[i];
    if (idx < 0) continue;
    double x = points[4 * idx], y = points[4 * idx + 1], z = points[4 * idx + 2];
    double v[10] = {1, x, y, z, x * x, x * y, x * z, y * y, y * z, z * z};
    for (int k = 0; k < 10; ++k) ii[10 * i + k] = v[k];
  }
  for (int r = 0; r < rows; ++r)
    for (int c = 1; c < cols; ++c)
      for (int k = 0; k < 10; ++k) ii[10 * ((size_t)r * cols + c) + k] += ii[10 * ((size_t)r * cols + c - 1) + k];
  for (int r = 1; r < rows; ++r)
    for (int c = 0; c < cols; ++c)
      for (int k = 0; k < 10; ++k) ii[10 * ((size_t)r * cols + c) + k] += ii[10 * ((size_t)(r - 1) * cols + c) + k];
  std::vector<float> depth(P);
  std::vector<int> iv(2 * P);
  orc_depth_u16_to_f32(raw, (int)P, depth_scale, depth.data());
  const bool suppress = cam->model == 1;
  if (suppress)
    orc_spherical_intervals(depth.data(), rows, cols, cam, sp->world_radius, iv.data());
  else
    orc_project_intervals(depth.data(), rows, cols, cam->K, cam->min_distance, cam->max_distance, sp->world_radius,
                          iv.data());
  for (int i = 0; i < n; ++i) {
    valid[i] = 0;
    curvature64[i] = 0.0;
    normals64[3 * i] = normals64[3 * i + 1] = normals64[3 * i + 2] = 0.0;
  }
  auto clampi = [](int v, int lo, int hi) { v = v < lo ? lo : v; v = v > hi ? hi : v; return v; };
#pragma omp parallel for schedule(static)
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c) {
      size_t pix = (size_t)r * cols + c;
      int index = index_image[pix];
      int R0 = iv[2 * pix], R1 = iv[2 * pix + 1];
      if (index < 0 || R0 < 0 || R1 < 0) continue;
      if (!suppress) {
        if (R0 < sp->min_image_radius || R1 < sp->min_image_radius) continue;
        if (R0 > sp->max_image_radius) R0 = sp->max_image_radius;
        if (R1 > sp->max_image_radius) R1 = sp->max_image_radius;
      }
      int xmin = clampi(c - R0 - 1, 0, cols - 1), xmax = clampi(c + R0 - 1, 0, cols - 1);
      int ymin = clampi(r - R1 - 1, 0, rows - 1), ymax = clampi(r + R1 - 1, 0, rows - 1);
      double acc[10];
      for (int k = 0; k < 10; ++k)
        acc[k] = ((ii[10 * ((size_t)ymax * cols + xmax) + k] + ii[10 * ((size_t)ymin * cols + xmin) + k]) -
                  ii[10 * ((size_t)ymax * cols + xmin) + k]) - ii[10 * ((size_t)ymin * cols + xmax) + k];
      if ((int)std::lround(acc[0]) < sp->min_points) continue;
      double nn = acc[0];
      double m[3] = {acc[1] / nn, acc[2] / nn, acc[3] / nn};
      double S[9] = {acc[4], acc[5], acc[6], acc[5], acc[7], acc[8], acc[6], acc[8], acc[9]};
      double a[9];
      for (int cc = 0; cc < 3; ++cc)
        for (int rr = 0; rr < 3; ++rr) a[cc * 3 + rr] = S[cc * 3 + rr] / nn - m[rr] * m[cc];
      double e[3], V[9];
      jacobi_eigen(3, a, e, V);
      if (e[0] < 0.0) e[0] = 0.0;
      double k = e[0] / (e[0] + e[1] + e[2] + 1e-9);
      curvature64[index] = k;
      double nv[3] = {V[0], V[1], V[2]};
      const float *p = points + 4 * index;
      if (nv[0] * p[0] + nv[1] * p[1] + nv[2] * p[2] > 0.0) { nv[0] = -nv[0]; nv[1] = -nv[1]; nv[2] = -nv[2]; }
      normals64[3 * index] = nv[0]; normals64[3 * index + 1] = nv[1]; normals64[3 * index + 2] = nv[2];
      valid[index] = 1;
    }
}

/* A14: PinholePointProjector::project, N/pinholepointprojector.cpp:33-66 + _project (.h:253-262) */
void orc_project(const float *points, int n, const float *KRt, int rows, int cols, float min_d, float max_d,
                 int *index_image, float *depth_image) {
  const size_t P = (size_t)rows * cols;
  std::vector<float> dimg(P, std::numeric_limits<float>::max());
  for (size_t i = 0; i < P; ++i) index_image[i] = -1;
  for (int i = 0; i < n; ++i) {
    float ip[4];
    mul4_packet(KRt, points + 4 * (size_t)i, ip);
    float d = ip[2];
    if (d < min_d || d > max_d) continue;
    float inv = 1.0f / d;
    float u = ip[0] * inv, v = ip[1] * inv;
    int x = (int)std::round(u), y = (int)std::round(v);
    if (x < 0 || x >= cols || y < 0 || y >= rows) continue;
    float &other = dimg[(size_t)y * cols + x];
    if (!other || other > d) {
      other = d;
      index_image[(size_t)y * cols + x] = i;
    }
  }
  if (depth_image) std::memcpy(depth_image, dimg.data(), P * sizeof(float));
}

/* A15: CorrespondenceFinderProjective::_computeSingle, N/correspondencefinderprojective.cpp:23-115;
 * A16: _computeMulti, :119-257. The OpenMP row partition and ordered compaction are
 * reproduced for num_threads threads (P10). */
int orc_correspondences(const orc_cloud *ref, const orc_cloud *cur, const int *ref_index, const int *cur_index,
                        int rows, int cols, const float *Tin, const orc_align_params *ap, int num_threads,
                        int *pairs) {
  float T[16];
  std::memcpy(T, Tin, sizeof(T));
  mr(T, 3, 0) = 0.0f; mr(T, 3, 1) = 0.0f; mr(T, 3, 2) = 0.0f; mr(T, 3, 3) = 1.0f;
  const float sq_thr = ap->inlier_distance_threshold * ap->inlier_distance_threshold;
  const float min_ratio = 1.0f / ap->inlier_curvature_ratio_threshold;
  const float max_ratio = ap->inlier_curvature_ratio_threshold;
  const float cos_thr = ap->inlier_normal_angular_threshold;
  const float flat = ap->flat_curvature_threshold;
  const bool multi = ap->row_search_region != 0 || ap->col_search_region != 0;
  const int nt = num_threads;
  const int rows_per_thread = rows / nt;
  const int iters_per_thread = (rows * cols) / nt;
  std::vector<int> local_count(nt, 0);
  std::vector<int> tmp((size_t)rows * cols * 2 + 2);
#pragma omp parallel for num_threads(nt) schedule(static, 1)
  for (int t = 0; t < nt; ++t) {
    int rmin = t * rows_per_thread, rmax = std::min(rmin + rows_per_thread, rows);
    int k = t * iters_per_thread;
    for (int r = rmin; r < rmax; ++r) {
      for (int c = 0; c < cols; ++c) {
        const int ri = ref_index[(size_t)r * cols + c];
        if (!multi) {
          const int ci = cur_index[(size_t)r * cols + c];
          if (ri < 0 || ci < 0) continue;
          const float *cn = cur->normals + 4 * (size_t)ci;
          const float *rn0 = ref->normals + 4 * (size_t)ri;
          const float *cp = cur->points + 4 * (size_t)ci;
          const float *rp0 = ref->points + 4 * (size_t)ri;
          if (!ap->demoted_to_gicp && (dot4(cn, cn) == 0.0f || dot4(rn0, rn0) == 0.0f)) continue;
          float rp[4], rn[4];
          iso_mul4(T, rp0, rp);
          iso_mul4(T, rn0, rn);
          float dv[4] = {cp[0] - rp[0], cp[1] - rp[1], cp[2] - rp[2], cp[3] - rp[3]};
          if (dot4(dv, dv) > sq_thr) continue;
          if (!ap->demoted_to_gicp) {
            if (dot4(cn, rn) < cos_thr) continue;
            float rc = ref->curvature[ri], cc = cur->curvature[ci];
            if (rc < flat) rc = flat;
            if (cc < flat) cc = flat;
            float ratio = (float)(((double)rc + 1e-5) / ((double)cc + 1e-5));
            if (ratio < min_ratio || ratio > max_ratio) continue;
          }
          tmp[2 * k] = ri; tmp[2 * k + 1] = ci; ++k;
        } else {
          if (ri < 0) continue;
          const float *rp0 = ref->points + 4 * (size_t)ri;
          const float *rn0 = ref->normals + 4 * (size_t)ri;
          if (dot4(rn0, rn0) == 0.0f) continue;
          float rp[4], rn[4];
          iso_mul4(T, rp0, rp);
          iso_mul4(T, rn0, rn);
          float rc = ref->curvature[ri];
          if (rc < flat) rc = flat;
          int r0 = std::max(r - ap->row_search_region, 0), r1 = std::min(r + ap->row_search_region + 1, rows);
          int c0 = std::max(c - ap->col_search_region, 0), c1 = std::min(c + ap->col_search_region + 1, cols);
          int best = -1;
          float best_nd = 0.0f, best_pd = 1e20f;
          for (int r2 = r0; r2 < r1; ++r2)
            for (int c2 = c0; c2 < c1; ++c2) {
              int ci = cur_index[(size_t)r2 * cols + c2];
              if (ci < 0) continue;
              const float *cn = cur->normals + 4 * (size_t)ci;
              if (dot4(cn, cn) == 0.0f) continue;
              const float *cp = cur->points + 4 * (size_t)ci;
              float nd = dot4(cn, rn);
              if (nd < cos_thr) continue;
              float dv[4] = {cp[0] - rp[0], cp[1] - rp[1], cp[2] - rp[2], cp[3] - rp[3]};
              float pd = dot4(dv, dv);
              if (pd > sq_thr) continue;
              float cc = cur->curvature[ci];
              if (cc < flat) cc = flat;
              float ratio = (float)(((double)rc + 1e-5) / ((double)cc + 1e-5));
              if (ratio < min_ratio || ratio > max_ratio) continue;
              if (best == -1 || (pd < best_pd && nd > best_nd)) { best = ci; best_nd = nd; best_pd = pd; }
            }
          if (best > -1) { tmp[2 * k] = ri; tmp[2 * k + 1] = best; ++k; }
        }
      }
    }
    local_count[t] = k;
  }
  int out = 0;
  for (int t = 0; t < nt; ++t)
    for (int i = t * iters_per_thread; i < local_count[t]; ++i) {
      pairs[2 * out] = tmp[2 * i];
      pairs[2 * out + 1] = tmp[2 * i + 1];
      ++out;
    }
  return out;
}

/* A17/A18: Linearizer::update, N/linearizer.cpp:21-134, in FP32 with the 4x4 product orders
 * of Eigen's lazy products (parity unpinned at the last bit). */
void orc_linearize(const orc_cloud *ref, const orc_cloud *cur, const int *pairs, int num_pairs, const float *T,
                   const orc_align_params *ap, int num_threads, float *H, float *b, float *error, int *inliers) {
  const int nt = num_threads;
  const int ipt = num_pairs / nt;
  std::vector<float> Htt(16 * nt), Htr(16 * nt), Hrr(16 * nt), bt(4 * nt), br(4 * nt), err(nt);
  std::vector<int> inl(nt);
#pragma omp parallel for num_threads(nt) schedule(static, 1)
  for (int t = 0; t < nt; ++t) {
    float htt[16] = {0}, htr[16] = {0}, hrr[16] = {0}, vbt[4] = {0}, vbr[4] = {0};
    float e = 0.0f;
    int in = 0;
    int imin = ipt * t, imax = std::min(imin + ipt, num_pairs);
    for (int i = imin; i < imax; ++i) {
      int ri = pairs[2 * i], ci = pairs[2 * i + 1];
      float rp[4], rn[4];
      iso_mul4(T, ref->points + 4 * (size_t)ri, rp);
      iso_mul4(T, ref->normals + 4 * (size_t)ri, rn);
      const float *cp = cur->points + 4 * (size_t)ci;
      const float *cn = cur->normals + 4 * (size_t)ci;
      float Op[16], On[16];
      std::memcpy(Op, cur->omega_p + 16 * (size_t)ci, sizeof(Op));
      for (int k = 0; k < 16; ++k) On[k] = ap->normal_omega_scale * cur->omega_n[16 * (size_t)ci + k];
      if (ap->demoted_to_gicp) std::memset(On, 0, sizeof(On));
      float pe[4] = {rp[0] - cp[0], rp[1] - cp[1], rp[2] - cp[2], rp[3] - cp[3]};
      float ne[4] = {rn[0] - cn[0], rn[1] - cn[1], rn[2] - cn[2], rn[3] - cn[3]};
      float ep[4], en[4];
      mul4_packet(Op, pe, ep);
      mul4_packet(On, ne, en);
      float le = dot4(pe, ep) + dot4(ne, en);
      float ks = 1.0f;
      if (le > ap->inlier_max_chi2) {
        if (ap->robust_kernel) ks = ap->inlier_max_chi2 / le;
        else continue;
      }
      ++in;
      e += ks * le;
      float Sp[16], Sn[16];
      skew4(rp, Sp);
      skew4(rn, Sn);
      for (int k = 0; k < 16; ++k) htt[k] += Op[k] * ks;
      float OS[16];
      mul44_packet(Op, Sp, OS);
      for (int k = 0; k < 16; ++k) htr[k] += OS[k] * ks;
      float a1[16], a2[16], b1[16], b2[16];
      mul44_tn(Sp, Op, a1);
      mul44_packet(a1, Sp, a2);
      mul44_tn(Sn, On, b1);
      mul44_packet(b1, Sn, b2);
      for (int k = 0; k < 16; ++k) hrr[k] += (a2[k] + b2[k]) * ks;
      for (int k = 0; k < 4; ++k) vbt[k] += ks * ep[k];
      float s1[4], s2[4];
      mul4_tn(Sp, ep, s1);
      mul4_tn(Sn, en, s2);
      for (int k = 0; k < 4; ++k) vbr[k] += ks * (s1[k] + s2[k]);
    }
    std::memcpy(&Htt[16 * t], htt, sizeof(htt));
    std::memcpy(&Htr[16 * t], htr, sizeof(htr));
    std::memcpy(&Hrr[16 * t], hrr, sizeof(hrr));
    std::memcpy(&bt[4 * t], vbt, sizeof(vbt));
    std::memcpy(&br[4 * t], vbr, sizeof(vbr));
    err[t] = e;
    inl[t] = in;
  }
  float htt[16] = {0}, htr[16] = {0}, hrr[16] = {0}, vbt[4] = {0}, vbr[4] = {0};
  float e = 0.0f;
  int in = 0;
  for (int t = 0; t < nt; ++t) {
    for (int k = 0; k < 16; ++k) { htt[k] += Htt[16 * t + k]; htr[k] += Htr[16 * t + k]; hrr[k] += Hrr[16 * t + k]; }
    for (int k = 0; k < 4; ++k) { vbt[k] += bt[4 * t + k]; vbr[k] += br[4 * t + k]; }
    in += inl[t];
    e += err[t];
  }
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) {
      H[c * 6 + r] = htt[c * 4 + r];
      H[(c + 3) * 6 + r] = htr[c * 4 + r];
      H[(c + 3) * 6 + r + 3] = hrr[c * 4 + r];
      H[c * 6 + r + 3] = htr[r * 4 + c];
    }
  for (int k = 0; k < 3; ++k) { b[k] = vbt[k]; b[k + 3] = vbr[k]; }
  *error = e;
  *inliers = in;
}

/* FP64 twin of A17 on identical float inputs: closed-form blocks J^T Omega J. */
void orc_linearize_f64(const orc_cloud *ref, const orc_cloud *cur, const int *pairs, int num_pairs, const float *Tf,
                       const orc_align_params *ap, double *H, double *b, double *error, int *inliers) {
  double T[16];
  for (int i = 0; i < 16; ++i) T[i] = Tf[i];
  double h[36] = {0}, g[6] = {0}, e = 0.0;
  int in = 0;
  for (int i = 0; i < num_pairs; ++i) {
    int ri = pairs[2 * i], ci = pairs[2 * i + 1];
    const float *p = ref->points + 4 * (size_t)ri, *n = ref->normals + 4 * (size_t)ri;
    double rp[3], rn[3];
    for (int k = 0; k < 3; ++k) {
      rp[k] = T[k] * p[0] + T[4 + k] * p[1] + T[8 + k] * p[2] + T[12 + k];
      rn[k] = T[k] * n[0] + T[4 + k] * n[1] + T[8 + k] * n[2];
    }
    const float *cp = cur->points + 4 * (size_t)ci, *cn = cur->normals + 4 * (size_t)ci;
    double Op[9], On[9];
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r) {
        Op[c * 3 + r] = cur->omega_p[16 * (size_t)ci + c * 4 + r];
        On[c * 3 + r] = ap->demoted_to_gicp ? 0.0 : (double)ap->normal_omega_scale * cur->omega_n[16 * (size_t)ci + c * 4 + r];
      }
    double pe[3] = {rp[0] - cp[0], rp[1] - cp[1], rp[2] - cp[2]};
    double ne[3] = {rn[0] - cn[0], rn[1] - cn[1], rn[2] - cn[2]};
    double ep[3], en[3];
    for (int r = 0; r < 3; ++r) {
      ep[r] = Op[r] * pe[0] + Op[3 + r] * pe[1] + Op[6 + r] * pe[2];
      en[r] = On[r] * ne[0] + On[3 + r] * ne[1] + On[6 + r] * ne[2];
    }
    double le = pe[0] * ep[0] + pe[1] * ep[1] + pe[2] * ep[2] + ne[0] * en[0] + ne[1] * en[1] + ne[2] * en[2];
    double ks = 1.0;
    if (le > ap->inlier_max_chi2) {
      if (ap->robust_kernel) ks = ap->inlier_max_chi2 / le;
      else continue;
    }
    ++in;
    e += ks * le;
    /* J_p = [I | S(rp)], J_n = [0 | S(rn)], S(v) = -2[v]x */
    auto skew = [](const double *v, double *S) {
      S[0] = 0; S[3] = 2 * v[2]; S[6] = -2 * v[1];
      S[1] = -2 * v[2]; S[4] = 0; S[7] = 2 * v[0];
      S[2] = 2 * v[1]; S[5] = -2 * v[0]; S[8] = 0;
    };
    double Sp[9], Sn[9];
    skew(rp, Sp);
    skew(rn, Sn);
    double Jp[18], Jn[18];  // 3x6 col-major
    for (int c = 0; c < 6; ++c)
      for (int r = 0; r < 3; ++r) {
        Jp[c * 3 + r] = c < 3 ? (r == c ? 1.0 : 0.0) : Sp[(c - 3) * 3 + r];
        Jn[c * 3 + r] = c < 3 ? 0.0 : Sn[(c - 3) * 3 + r];
      }
    for (int a = 0; a < 6; ++a) {
      double OJp[3], OJn[3];
      for (int r = 0; r < 3; ++r) {
        OJp[r] = Op[r] * Jp[a * 3] + Op[3 + r] * Jp[a * 3 + 1] + Op[6 + r] * Jp[a * 3 + 2];
        OJn[r] = On[r] * Jn[a * 3] + On[3 + r] * Jn[a * 3 + 1] + On[6 + r] * Jn[a * 3 + 2];
      }
      for (int c = 0; c < 6; ++c) {
        double v = 0.0;
        for (int r = 0; r < 3; ++r) v += Jp[c * 3 + r] * OJp[r] + Jn[c * 3 + r] * OJn[r];
        h[a * 6 + c] += ks * v;
      }
      double gb = 0.0;
      for (int r = 0; r < 3; ++r) gb += Jp[a * 3 + r] * ep[r] + Jn[a * 3 + r] * en[r];
      g[a] += ks * gb;
    }
  }
  std::memcpy(H, h, sizeof(h));
  std::memcpy(b, g, sizeof(g));
  *error = e;
  *inliers = in;
}

/* OpenMP threads of the conversion's parallel loops (the aligner takes its count per call) */
void orc_set_num_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_max_threads(void) { return omp_get_max_threads(); }

void orc_v2t(const float *x, float *T) { v2t_t<float>(x, T); }
void orc_t2v(const float *T, float *x) { t2v_t<float>(T, x); }
void orc_isometry_mul(const float *A, const float *B, float *C) { iso_mul(A, B, C); }
void orc_isometry_inverse(const float *A, float *Ai) { iso_inverse(A, Ai); }
int orc_ldlt_solve_f32(const float *H, const float *rhs, float *x) { return ldlt_solve<float>(H, rhs, x); }
int orc_ldlt_solve_f64(const double *H, const double *rhs, double *x) { return ldlt_solve<double>(H, rhs, x); }

/* A20: Aligner::_computeStatistics, N/aligner.cpp:82-135 + N/unscented.h:11-65, in FP64.
 * Zero information (no inliers) or a failed Cholesky is reported invalid. */
void orc_compute_statistics(const double *H, const float *Tf, const orc_align_params *ap, orc_align_result *res) {
  double Sigma[36];
  pinv6_sym(H, Sigma);
  const int dim = 6;
  const double alpha = 1e-3, beta = 2.0, lambda = alpha * alpha * dim;
  const double wi = 1.0 / (2.0 * (dim + lambda));
  double w0i = lambda / (dim + lambda), w0p = lambda / (dim + lambda) + (1.0 - alpha * alpha + beta);
  double C[36], L[36];
  for (int i = 0; i < 36; ++i) C[i] = Sigma[i] * (dim + lambda);
  bool ok = llt6(C, L);
  double samples[13][6], wI[13], wP[13];
  for (int k = 0; k < 6; ++k) samples[0][k] = 0.0;
  wI[0] = w0i; wP[0] = w0p;
  int s = 1;
  for (int i = 0; i < dim; ++i) {
    for (int k = 0; k < 6; ++k) { samples[s][k] = L[i * 6 + k]; samples[s + 1][k] = -L[i * 6 + k]; }
    wI[s] = wP[s] = wI[s + 1] = wP[s + 1] = wi;
    s += 2;
  }
  double T[16];
  for (int i = 0; i < 16; ++i) T[i] = Tf[i];
  for (int i = 0; i < 13; ++i) {
    double X[16], Xi[16], Y[16];
    v2t_t<double>(samples[i], X);
    /* inverse of an isometry */
    for (int a = 0; a < 16; ++a) Xi[a] = 0.0;
    Xi[15] = 1.0;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) Xi[c * 4 + r] = X[r * 4 + c];
    for (int r = 0; r < 3; ++r) Xi[12 + r] = -(Xi[r] * X[12] + Xi[4 + r] * X[13] + Xi[8 + r] * X[14]);
    for (int a = 0; a < 16; ++a) Y[a] = 0.0;
    Y[15] = 1.0;
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) Y[c * 4 + r] = T[r] * Xi[c * 4] + T[4 + r] * Xi[c * 4 + 1] + T[8 + r] * Xi[c * 4 + 2];
      Y[12 + r] = T[r] * Xi[12] + T[4 + r] * Xi[13] + T[8 + r] * Xi[14] + T[12 + r];
    }
    t2v_t<double>(Y, samples[i]);
  }
  double mean[6] = {0}, cov[36] = {0};
  for (int i = 0; i < 13; ++i)
    for (int k = 0; k < 6; ++k) mean[k] += wI[i] * samples[i][k];
  for (int i = 0; i < 13; ++i) {
    double d[6];
    for (int k = 0; k < 6; ++k) d[k] = samples[i][k] - mean[k];
    for (int c = 0; c < 6; ++c)
      for (int r = 0; r < 6; ++r) cov[c * 6 + r] += wP[i] * d[r] * d[c];
  }
  double Om[36];
  bool inv_ok = inverse6(cov, Om);
  double O[36];
  for (int c = 0; c < 6; ++c)
    for (int r = 0; r < 6; ++r) O[c * 6 + r] = 0.5 * (Om[c * 6 + r] + Om[r * 6 + c]);
  std::memcpy(res->mean, mean, sizeof(mean));
  std::memcpy(res->omega, O, sizeof(O));
  double tr = sv_ratio3(O, 0), rr = sv_ratio3(O, 3);
  res->translational_ratio = tr;
  res->rotational_ratio = rr;
  bool finite = std::isfinite(tr) && std::isfinite(rr);
  bool zeroH = true;
  for (int i = 0; i < 36; ++i) if (H[i] != 0.0) zeroH = false;
  res->solution_valid = ok && inv_ok && finite && !zeroH && !(rr > ap->rotational_min_eigen_ratio) &&
                        !(tr > ap->translational_min_eigen_ratio);
}

static void gn_step(const orc_cloud *ref, const orc_cloud *cur, const int *pairs, int nc, float *invT,
                    const orc_align_params *ap, int num_threads, int mode, float *H, float *b, float *err, int *inl);

/* A22: CorrespondenceFinderNN, N/correspondencefindernn.cpp:13-97. FLANN's randomized kd-tree
 * search (SearchParams(16), approximate) is replaced by exact brute force; the distance is
 * FLANN's L2 functor order for 6 dimensions: (((d0^2 + d1^2) + d2^2) + d3^2) then + d4^2, + d5^2;
 * FLANN's radius search keeps distances strictly below the radius (the squared threshold). */
int orc_nn_correspondences(const orc_cloud *ref, const orc_cloud *cur, const float *Tin, const orc_align_params *ap,
                           float ns, int *pairs, int *nn_out) {
  float T[16];
  std::memcpy(T, Tin, sizeof(T));
  const float sq_thr = ap->inlier_distance_threshold * ap->inlier_distance_threshold;
  const int nr = ref->n, nc = cur->n;
  /* _model2linear of the current cloud: (p, ns * n) */
  std::vector<float> cl((size_t)nc * 6);
  for (int i = 0; i < nc; ++i) {
    const float *p = cur->points + 4 * (size_t)i, *n = cur->normals + 4 * (size_t)i;
    cl[6 * i] = p[0]; cl[6 * i + 1] = p[1]; cl[6 * i + 2] = p[2];
    cl[6 * i + 3] = ns * n[0]; cl[6 * i + 4] = ns * n[1]; cl[6 * i + 5] = ns * n[2];
  }
  /* sR = T.linear() * ns (elementwise) */
  float sR[9];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) sR[c * 3 + r] = mc(T, r, c) * ns;
  std::vector<int> nn(nr, -1);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < nr; ++i) {
    const float *p = ref->points + 4 * (size_t)i, *n = ref->normals + 4 * (size_t)i;
    float q[6], pt[4];
    iso_mul4(T, p, pt);
    q[0] = pt[0]; q[1] = pt[1]; q[2] = pt[2];
    for (int r = 0; r < 3; ++r) q[3 + r] = red3(sR[r] * n[0], sR[3 + r] * n[1], sR[6 + r] * n[2]);
    float best = sq_thr;
    int bi = -1;
    for (int j = 0; j < nc; ++j) {
      const float *c = &cl[6 * (size_t)j];
      float d0 = q[0] - c[0], d1 = q[1] - c[1], d2 = q[2] - c[2], d3 = q[3] - c[3], d4 = q[4] - c[4], d5 = q[5] - c[5];
      float d = 0.0f + (d0 * d0 + d1 * d1 + d2 * d2 + d3 * d3);
      d += d4 * d4;
      d += d5 * d5;
      if (d < best) { best = d; bi = j; }
    }
    nn[i] = bi;
  }
  if (nn_out) std::memcpy(nn_out, nn.data(), (size_t)nr * sizeof(int));
  int k = 0;
  for (int i = 0; i < nr; ++i) {
    int ci = nn[i];
    if (ci < 0) continue;
    const float *p = ref->points + 4 * (size_t)i, *n = ref->normals + 4 * (size_t)i;
    float rn[3], rp[3];
    for (int r = 0; r < 3; ++r) {
      rn[r] = red3(mc(T, r, 0) * n[0], mc(T, r, 1) * n[1], mc(T, r, 2) * n[2]);
      rp[r] = red3(mc(T, r, 0) * p[0], mc(T, r, 1) * p[1], mc(T, r, 2) * p[2]) + mc(T, r, 3);
    }
    const float *cn = cur->normals + 4 * (size_t)ci, *cp = cur->points + 4 * (size_t)ci;
    if (!ap->demoted_to_gicp && (red3(rn[0] * rn[0], rn[1] * rn[1], rn[2] * rn[2]) == 0.0f ||
                                 red3(cn[0] * cn[0], cn[1] * cn[1], cn[2] * cn[2]) == 0.0f))
      continue;
    float dv[3] = {rp[0] - cp[0], rp[1] - cp[1], rp[2] - cp[2]};
    if (red3(dv[0] * dv[0], dv[1] * dv[1], dv[2] * dv[2]) > sq_thr) continue;
    if (std::isnan(cn[0]) || std::isnan(cn[1]) || std::isnan(cn[2]) || std::isnan(rn[0]) || std::isnan(rn[1]) ||
        std::isnan(rn[2]))
      continue;
    if (!ap->demoted_to_gicp && red3(rn[0] * cn[0], rn[1] * cn[1], rn[2] * cn[2]) < ap->inlier_normal_angular_threshold)
      continue;
    pairs[2 * k] = i;
    pairs[2 * k + 1] = ci;
    ++k;
  }
  return k;
}

/* A22: AlignerNN::align, N/alignernn.cpp:21-118 */
void orc_align_nn(const orc_cloud *ref, const orc_cloud *cur, const float *init_guess, const orc_align_params *ap,
                  float ns, float *T_out, orc_align_result *res) {
  orc_align_nn_mode(ref, cur, init_guess, ap, ns, 0, T_out, res);
}

void orc_align_nn_mode(const orc_cloud *ref, const orc_cloud *cur, const float *init_guess, const orc_align_params *ap,
                       float ns, int mode, float *T_out, orc_align_result *res) {
  std::vector<int> pairs(2 * (size_t)std::max(ref->n, 1));
  float T[16];
  std::memcpy(T, init_guess, sizeof(T));
  float H[36] = {0}, b[6] = {0}, err = 0.0f;
  int inl = 0, nc = 0;
  for (int it = 0; it < ap->outer_iterations; ++it) {
    mr(T, 3, 0) = 0.0f; mr(T, 3, 1) = 0.0f; mr(T, 3, 2) = 0.0f; mr(T, 3, 3) = 1.0f;
    float Ti[16];
    iso_inverse(T, Ti);
    nc = orc_nn_correspondences(ref, cur, Ti, ap, ns, pairs.data(), nullptr);
    float invT[16];
    iso_inverse(T, invT);
    for (int k = 0; k < ap->inner_iterations; ++k) {
      mr(invT, 3, 0) = 0.0f; mr(invT, 3, 1) = 0.0f; mr(invT, 3, 2) = 0.0f; mr(invT, 3, 3) = 1.0f;
      gn_step(ref, cur, pairs.data(), nc, invT, ap, 1, mode, H, b, &err, &inl);
    }
    iso_inverse(invT, T);
    float x[6];
    t2v_t<float>(T, x);
    v2t_t<float>(x, T);
  }
  std::memcpy(T_out, T, sizeof(T));
  res->error = err;
  res->inliers = inl;
  res->num_correspondences = nc;
  float invT[16];
  iso_inverse(T, invT);
  float Hs[36], bs[6], es;
  int is;
  orc_linearize(ref, cur, pairs.data(), nc, invT, ap, 1, Hs, bs, &es, &is);
  std::memcpy(res->H, Hs, sizeof(Hs));
  std::memcpy(res->b, bs, sizeof(bs));
  double Hd[36];
  for (int i = 0; i < 36; ++i) Hd[i] = Hs[i];
  orc_compute_statistics(Hd, T, ap, res);
}

/* A13: AlignerProjective::align, N/alignerprojective.cpp:21-136 (sensor offsets identity, no priors) */
/* one Gauss-Newton step of the aligner: mode 0 = the reference's FP32 (Linearizer FP32,
 * LDLT<Matrix6f>); mode 1 = H and b in FP64 and an FP64 LDLT, the update still in FP32 */
static void gn_step(const orc_cloud *ref, const orc_cloud *cur, const int *pairs, int nc, float *invT,
                    const orc_align_params *ap, int num_threads, int mode, float *H, float *b, float *err, int *inl) {
  float dx[6];
  if (mode == 0) {
    orc_linearize(ref, cur, pairs, nc, invT, ap, num_threads, H, b, err, inl);
    for (int d = 0; d < 6; ++d) H[d * 6 + d] = H[d * 6 + d] + ap->lambda;
    float mb[6];
    for (int d = 0; d < 6; ++d) mb[d] = -b[d];
    ldlt_solve<float>(H, mb, dx);
  } else {
    double Hd[36], bd[6], ed;
    orc_linearize_f64(ref, cur, pairs, nc, invT, ap, Hd, bd, &ed, inl);
    for (int i = 0; i < 36; ++i) H[i] = (float)Hd[i];
    for (int i = 0; i < 6; ++i) b[i] = (float)bd[i];
    *err = (float)ed;
    for (int d = 0; d < 6; ++d) Hd[d * 6 + d] += (double)ap->lambda;
    double mb[6], x[6];
    for (int d = 0; d < 6; ++d) mb[d] = -bd[d];
    ldlt_solve<double>(Hd, mb, x);
    for (int d = 0; d < 6; ++d) dx[d] = (float)x[d];
  }
  float dT[16];
  v2t_t<float>(dx, dT);
  iso_mul(dT, invT, invT);
}

void orc_align(const orc_cloud *ref, const orc_cloud *cur, int rows, int cols, const orc_camera *cam,
               const float *init_guess, const orc_align_params *ap, int num_threads, float *T_out,
               orc_align_result *res) {
  orc_align_mode(ref, cur, rows, cols, cam, init_guess, ap, num_threads, 0, T_out, res);
}

void orc_align_mode(const orc_cloud *ref, const orc_cloud *cur, int rows, int cols, const orc_camera *cam,
                    const float *init_guess, const orc_align_params *ap, int num_threads, int mode, float *T_out,
                    orc_align_result *res) {
  orc_align_offsets(ref, cur, rows, cols, cam, init_guess, nullptr, nullptr, ap, num_threads, mode, T_out, res);
}

/* A13 with the aligner's sensor offsets (N/aligner.h:157-196): the current cloud is projected from
 * the current offset (N/alignerprojective.cpp:36-39), the reference from T * reference offset
 * (N/alignerprojective.cpp:51); each offset's last row is set to 0 0 0 1 (N/aligner.h:177,194);
 * NULL is the identity. */
void orc_align_offsets(const orc_cloud *ref, const orc_cloud *cur, int rows, int cols, const orc_camera *cam,
                       const float *init_guess, const float *ref_offset, const float *cur_offset,
                       const orc_align_params *ap, int num_threads, int mode, float *T_out, orc_align_result *res) {
  const size_t P = (size_t)rows * cols;
  std::vector<int> cur_index(P), ref_index(P), pairs(2 * P);
  float RO[16], CO[16];
  set_identity4(RO);
  set_identity4(CO);
  if (ref_offset) std::memcpy(RO, ref_offset, sizeof(RO));
  if (cur_offset) std::memcpy(CO, cur_offset, sizeof(CO));
  mr(RO, 3, 0) = 0.0f; mr(RO, 3, 1) = 0.0f; mr(RO, 3, 2) = 0.0f; mr(RO, 3, 3) = 1.0f;
  mr(CO, 3, 0) = 0.0f; mr(CO, 3, 1) = 0.0f; mr(CO, 3, 2) = 0.0f; mr(CO, 3, 3) = 1.0f;
  project_cam(cam, CO, cur->points, cur->n, rows, cols, cur_index.data(), nullptr);
  float T[16];
  std::memcpy(T, init_guess, sizeof(T));
  float H[36] = {0}, b[6] = {0}, err = 0.0f;
  int inl = 0, nc = 0;
  for (int it = 0; it < ap->outer_iterations; ++it) {
    mr(T, 3, 0) = 0.0f; mr(T, 3, 1) = 0.0f; mr(T, 3, 2) = 0.0f; mr(T, 3, 3) = 1.0f;
    float TR[16];
    iso_mul(T, RO, TR);
    project_cam(cam, TR, ref->points, ref->n, rows, cols, ref_index.data(), nullptr);
    float Ti[16];
    iso_inverse(T, Ti);
    nc = orc_correspondences(ref, cur, ref_index.data(), cur_index.data(), rows, cols, Ti, ap, num_threads, pairs.data());
    float invT[16];
    iso_inverse(T, invT);
    for (int k = 0; k < ap->inner_iterations; ++k) {
      mr(invT, 3, 0) = 0.0f; mr(invT, 3, 1) = 0.0f; mr(invT, 3, 2) = 0.0f; mr(invT, 3, 3) = 1.0f;
      gn_step(ref, cur, pairs.data(), nc, invT, ap, num_threads, mode, H, b, &err, &inl);
    }
    iso_inverse(invT, T);
    float x[6];
    t2v_t<float>(T, x);
    v2t_t<float>(x, T);
  }
  std::memcpy(T_out, T, sizeof(T));
  res->error = err;
  res->inliers = inl;
  res->num_correspondences = nc;
  /* statistics: one more linearisation at the final T over the last correspondences */
  float invT[16];
  iso_inverse(T, invT);
  float Hs[36], bs[6], es;
  int is;
  orc_linearize(ref, cur, pairs.data(), nc, invT, ap, num_threads, Hs, bs, &es, &is);
  std::memcpy(res->H, Hs, sizeof(Hs));
  std::memcpy(res->b, bs, sizeof(bs));
  double Hd[36];
  for (int i = 0; i < 36; ++i) Hd[i] = Hs[i];
  orc_compute_statistics(Hd, T, ap, res);
}


/* ---------------- Gaussian3f, Cloud::transformInPlace, Merger::merge ----------------
 * Gaussian3f keeps a moments form (mean, covariance) and an information form (matrix,
 * vector), each computed from the other only when read (N/gaussian.h:104-122); the flags
 * of orc_gaussian record which form is current. Parity: the covariance of the
 * unprojection is pinned by its closed form in FP64, the fusion of merge by the
 * information-form product of Gaussians in FP64, the transforms by R S R^T in FP64, the
 * merge decisions by a brute force in tests/test_merge.py. The association orders of the
 * 3x3 products (R3) are the ones of the rest of this file. */

namespace {

inline void mat3_vec(const float *M, const float *v, float *o) {
  float r[3];
  for (int i = 0; i < 3; ++i) r[i] = red3(M[i] * v[0], M[3 + i] * v[1], M[6 + i] * v[2]);
  std::memcpy(o, r, sizeof(r));
}

/* Gaussian::_updateInfo (N/gaussian.h:116-122) */
void g_update_info(orc_gaussian *g) {
  if (g->flags & 2) return;
  orc_inverse3(g->covariance, g->information);
  mat3_vec(g->information, g->mean, g->information_vector);
  g->flags |= 2;
}

/* Gaussian::_updateMoments (N/gaussian.h:108-114) */
void g_update_moments(orc_gaussian *g) {
  if (g->flags & 1) return;
  orc_inverse3(g->information, g->covariance);
  mat3_vec(g->covariance, g->information_vector, g->mean);
  g->flags |= 1;
}

/* Gaussian::addInformation (N/gaussian.h:50-56) */
void g_add_information(orc_gaussian *t, orc_gaussian *s) {
  g_update_info(t);
  g_update_info(s);
  for (int k = 0; k < 9; ++k) t->information[k] += s->information[k];
  for (int k = 0; k < 3; ++k) t->information_vector[k] += s->information_vector[k];
  t->flags = 2;
}

/* 4x4 column-major product with P1 per column */
void mat4_mul(const float *A, const float *B, float *C) {
  float o[16];
  for (int j = 0; j < 4; ++j) mul4_packet(A, B + 4 * j, o + 4 * j);
  std::memcpy(C, o, sizeof(o));
}

}  // namespace

/* PinholePointProjector::unProject with Gaussians, N/pinholepointprojector.cpp:93-134 */
int orc_unproject_gaussians(const float *depth, int rows, int cols, const float *K, const float *iKRt, float min_d,
                            float max_d, float baseline, float alpha, float *points, orc_gaussian *gaussians,
                            int *index_image) {
  const int n = orc_unproject(depth, rows, cols, iKRt, min_d, max_d, points, index_image);
  float iK[9];
  orc_inverse3(K, iK);
  const float fB = baseline * K[0];
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c) {
      const int i = index_image[r * cols + c];
      if (i < 0) continue;
      const float z = depth[r * cols + c];
      const float zVariation = (alpha * z * z) / (fB + z * alpha);
      /* J << z, 0, c, 0, z, r, 0, 0, 1 (row by row), then J = iK * J */
      float J0[3][3] = {{z, 0.0f, (float)c}, {0.0f, z, (float)r}, {0.0f, 0.0f, 1.0f}};
      float J[3][3];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          J[a][b] = red3(iK[0 * 3 + a] * J0[0][b], iK[1 * 3 + a] * J0[1][b], iK[2 * 3 + a] * J0[2][b]);
      const float D[3] = {3.0f, 3.0f, zVariation};
      orc_gaussian &g = gaussians[i];
      std::memset(&g, 0, sizeof(g));
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          g.covariance[b * 3 + a] =
              red3((J[a][0] * D[0]) * J[b][0], (J[a][1] * D[1]) * J[b][1], (J[a][2] * D[2]) * J[b][2]);
      for (int k = 0; k < 3; ++k) g.mean[k] = points[4 * (size_t)i + k];
      g.flags = 1;
    }
  return n;
}

/* Cloud::transformInPlace, N/cloud.cpp:193-205, with the transformable vectors of
 * N/transformable_vector.h:31-37 (points, normals), N/stats.h:121-126, N/gaussian3.h:59-67,
 * N/informationmatrix.h:111-121. Any array may be NULL. */
void orc_cloud_transform(int n, const float *T, float *points, float *normals, float *stats16, float *omega_p,
                         float *omega_n, orc_gaussian *g) {
  float m[16];
  std::memcpy(m, T, sizeof(m));
  mr(m, 3, 0) = 0.0f; mr(m, 3, 1) = 0.0f; mr(m, 3, 2) = 0.0f; mr(m, 3, 3) = 1.0f;
  /* the vectors are left alone when m is the identity (N/cloud.cpp:196) */
  float I[16];
  set_identity4(I);
  bool identity = true;
  for (int k = 0; k < 16; ++k) identity = identity && m[k] == I[k];
  if (identity) return;
  float Tz[16], Tt[16];
  std::memcpy(Tz, m, sizeof(Tz));
  for (int k = 0; k < 4; ++k) { mr(Tz, 3, k) = 0.0f; mr(Tz, k, 3) = 0.0f; }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) mr(Tt, r, c) = mc(Tz, c, r);
  for (int i = 0; i < n; ++i) {
    if (points) {
      float *p = points + 4 * (size_t)i;
      mul4_packet(m, p, p);
      p[3] = 1.0f;
    }
    if (normals) {
      float *q = normals + 4 * (size_t)i;
      mul4_packet(m, q, q);
      q[3] = 0.0f;
    }
    if (stats16) mat4_mul(m, stats16 + 16 * (size_t)i, stats16 + 16 * (size_t)i);
    for (float *O : {omega_p, omega_n}) {
      if (!O) continue;
      float *W = O + 16 * (size_t)i, A[16];
      mat4_mul(Tz, W, A);
      mat4_mul(A, Tt, W);
    }
    if (g) {
      orc_gaussian &G = g[i];
      g_update_moments(&G);
      float mean[3], RC[9], cov[9];
      for (int a = 0; a < 3; ++a)
        mean[a] = red3(mc(m, a, 0) * G.mean[0], mc(m, a, 1) * G.mean[1], mc(m, a, 2) * G.mean[2]) + mc(m, a, 3);
      for (int a = 0; a < 3; ++a)
        for (int k = 0; k < 3; ++k)
          RC[k * 3 + a] = red3(mc(m, a, 0) * G.covariance[k * 3], mc(m, a, 1) * G.covariance[k * 3 + 1],
                               mc(m, a, 2) * G.covariance[k * 3 + 2]);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          cov[b * 3 + a] = red3(RC[a] * mc(m, b, 0), RC[3 + a] * mc(m, b, 1), RC[6 + a] * mc(m, b, 2));
      std::memcpy(G.mean, mean, sizeof(mean));
      std::memcpy(G.covariance, cov, sizeof(cov));
      G.flags = 1;
    }
  }
}

/* Merger::merge, N/merger.cpp:15-109, in place on the given arrays (any of normals,
 * curvature, stats16, omega_p, omega_n may be NULL). Returns the new size; collapsed
 * (optional, n ints) receives the per-point decision. */
int orc_merge(int n, float *points, float *normals, float *curvature, float *stats16, float *omega_p, float *omega_n,
              orc_gaussian *g, const float *transform, const orc_camera *cam, int rows, int cols,
              float distance_threshold, float max_point_depth, int *collapsed_out) {
  float T[16], KRt[16], iKRt[16];
  std::memcpy(T, transform, sizeof(T));
  mr(T, 3, 0) = 0.0f; mr(T, 3, 1) = 0.0f; mr(T, 3, 2) = 0.0f; mr(T, 3, 3) = 1.0f;
  orc_projector_matrices(cam->K, T, KRt, iKRt);
  std::vector<int> idx((size_t)rows * cols);
  std::vector<float> dep((size_t)rows * cols);
  project_cam(cam, T, points, n, rows, cols, idx.data(), dep.data());
  const SphRes sr = sph_res(cam, rows, cols);
  float iT[16];
  iso_inverse(T, iT);
  std::vector<int> collapsed((size_t)n, -1);
  for (int i = 0; i < n; ++i) {
    float depth;
    int r = -1, c = -1;
    if (cam->model == 1) {
      int x, y;
      if (sph_project1(cam, sr, iT, points + 4 * (size_t)i, x, y, depth)) {
        c = x;
        r = y;
      }
    } else {
      float ip[4];
      mul4_packet(KRt, points + 4 * (size_t)i, ip);
      depth = ip[2];
      if (!(depth < cam->min_distance || depth > cam->max_distance)) {
        const float inv = 1.0f / depth;
        c = (int)std::round(ip[0] * inv);
        r = (int)std::round(ip[1] * inv);
      }
    }
    if (depth < 0 || depth > max_point_depth || r < 0 || r >= rows || c < 0 || c >= cols) continue;
    const float targetZ = dep[(size_t)r * cols + c];
    const int target = idx[(size_t)r * cols + c];
    if (target < 0) continue;
    if (target == i) {
      collapsed[i] = i;
    } else if (std::fabs(depth - targetZ) < distance_threshold) {
      g_add_information(&g[target], &g[i]);
      collapsed[i] = target;
    }
  }
  int k = 0;
  for (int i = 0; i < n; ++i) {
    const int ci = collapsed[i];
    if (ci == i) {
      g_update_moments(&g[i]);
      for (int e = 0; e < 3; ++e) points[4 * (size_t)i + e] = g[i].mean[e];
    }
    if (ci < 0 || ci == i) {
      std::memmove(points + 4 * (size_t)k, points + 4 * (size_t)i, 4 * sizeof(float));
      if (normals) std::memmove(normals + 4 * (size_t)k, normals + 4 * (size_t)i, 4 * sizeof(float));
      if (curvature) curvature[k] = curvature[i];
      if (stats16) std::memmove(stats16 + 16 * (size_t)k, stats16 + 16 * (size_t)i, 16 * sizeof(float));
      if (omega_p) std::memmove(omega_p + 16 * (size_t)k, omega_p + 16 * (size_t)i, 16 * sizeof(float));
      if (omega_n) std::memmove(omega_n + 16 * (size_t)k, omega_n + 16 * (size_t)i, 16 * sizeof(float));
      g[k] = g[i];
      ++k;
    }
  }
  if (collapsed_out) std::memcpy(collapsed_out, collapsed.data(), (size_t)n * sizeof(int));
  return k;
}

}  // extern "C"
