// Small fixed-size math of the NICP registration loop, shared by the HIP
// kernels and the host side of the C-ABI. Every function states the
// association order it reproduces so that device results are bit-identical
// to the reference's FP32 arithmetic where the reference fixes it:
//
//   P1  Matrix4f * Vector4f (Eigen lazy packet product):
//       ((a0 v0 + a1 v1) + a2 v2) + a3 v3
//   P3  Isometry3f * Vector4f (3x4 affine block, scalar redux):
//       (a0 v0 + a1 v1) + (a2 v2 + a3 v3), w copied from the input
//   P4  Vector4f dot / squaredNorm (SSE predux):
//       (x0 y0 + x2 y2) + (x1 y1 + x3 y3)
//   R3  3-term redux (3x3 products, trace): x0 + (x1 + x2)
//
// The library is compiled with -ffp-contract=off, so no product is fused
// into a following add, on the host or on the device.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define NICP_HD __host__ __device__ inline

namespace nicp_math {

// 4x4 column-major
NICP_HD float mc(const float *m, int r, int c) { return m[c * 4 + r]; }
NICP_HD float &mr(float *m, int r, int c) { return m[c * 4 + r]; }

NICP_HD float red3(float x0, float x1, float x2) { return x0 + (x1 + x2); }

// P1 for a homogeneous input with the given components
NICP_HD void mul4_packet(const float *A, float v0, float v1, float v2, float v3, float *o) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float r = mc(A, i, 0) * v0;
    r = mc(A, i, 1) * v1 + r;
    r = mc(A, i, 2) * v2 + r;
    r = mc(A, i, 3) * v3 + r;
    o[i] = r;
  }
}

// Isometry3f * Vector4f under Eigen's packet order (a0 v0 + a1 v1) + (a2 v2 + a3 v3); only the
// first three components (w is the input's)
NICP_HD void iso_mul3(const float *A, float v0, float v1, float v2, float v3, float *o) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
    o[i] = (mc(A, i, 0) * v0 + mc(A, i, 1) * v1) + (mc(A, i, 2) * v2 + mc(A, i, 3) * v3);
}

// iso_mul3 of a direction (w = 0): with a finite translation a3 * 0 is a signed zero, and adding
// a signed zero changes no value (at most the sign of a zero result), so the term is dropped;
// every comparison and sum downstream sees the same values
NICP_HD void iso_rot3(const float *A, float v0, float v1, float v2, float *o) {
#pragma unroll
  for (int i = 0; i < 3; ++i) o[i] = (mc(A, i, 0) * v0 + mc(A, i, 1) * v1) + mc(A, i, 2) * v2;
}

// P4
NICP_HD float dot4(float a0, float a1, float a2, float a3, float b0, float b1, float b2, float b3) {
  return (a0 * b0 + a2 * b2) + (a1 * b1 + a3 * b3);
}

NICP_HD void set_identity4(float *m) {
  for (int i = 0; i < 16; ++i) m[i] = 0.0f;
  m[0] = m[5] = m[10] = m[15] = 1.0f;
}

// Transform<float,3,Isometry>::inverse(): [R^T | -(R^T t)]
NICP_HD void iso_inverse(const float *A, float *B) {
  float o[16];
  set_identity4(o);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) mr(o, i, j) = mc(A, j, i);
  for (int i = 0; i < 3; ++i)
    mr(o, i, 3) = -red3(mc(o, i, 0) * mc(A, 0, 3), mc(o, i, 1) * mc(A, 1, 3), mc(o, i, 2) * mc(A, 2, 3));
  for (int i = 0; i < 16; ++i) B[i] = o[i];
}

// Isometry * Isometry: linear = A.R B.R, translation = A.R B.t + A.t
NICP_HD void iso_mul(const float *A, const float *B, float *C) {
  float o[16];
  set_identity4(o);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      mr(o, i, j) = red3(mc(A, i, 0) * mc(B, 0, j), mc(A, i, 1) * mc(B, 1, j), mc(A, i, 2) * mc(B, 2, j));
  for (int i = 0; i < 3; ++i)
    mr(o, i, 3) = red3(mc(A, i, 0) * mc(B, 0, 3), mc(A, i, 1) * mc(B, 1, 3), mc(A, i, 2) * mc(B, 2, 3)) + mc(A, i, 3);
  for (int i = 0; i < 16; ++i) C[i] = o[i];
}

NICP_HD void set_row3(float *m) { m[3] = 0.0f; m[7] = 0.0f; m[11] = 0.0f; m[15] = 1.0f; }

// Eigen's 3x3 cofactor inverse (K is column-major 3x3)
NICP_HD void inverse3(const float *K, float *iK) {
#define M3(r, c) K[(c)*3 + (r)]
  float cof[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      cof[i][j] = M3(i1, j1) * M3(i2, j2) - M3(i1, j2) * M3(i2, j1);
    }
  float det = red3(cof[0][0] * M3(0, 0), cof[1][0] * M3(1, 0), cof[2][0] * M3(2, 0));
  float invdet = 1.0f / det;
  // result(i, j) = cof(j, i) * invdet
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) iK[j * 3 + i] = cof[j][i] * invdet;
#undef M3
}

// PinholePointProjector::_updateMatrices (nicp/nicp/pinholepointprojector.cpp:16-31)
NICP_HD void projector_matrices(const float *K, const float *transform, float *KRt, float *iKRt) {
  float t[16];
  iso_inverse(transform, t);
  float iK[9];
  inverse3(K, iK);
  set_identity4(KRt);
  set_identity4(iKRt);
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      mr(KRt, i, j) = red3(K[i] * mc(t, 0, j), K[3 + i] * mc(t, 1, j), K[6 + i] * mc(t, 2, j));
      mr(iKRt, i, j) = red3(mc(transform, i, 0) * iK[j * 3 + 0], mc(transform, i, 1) * iK[j * 3 + 1],
                            mc(transform, i, 2) * iK[j * 3 + 2]);
    }
    mr(KRt, i, 3) = red3(K[i] * mc(t, 0, 3), K[3 + i] * mc(t, 1, 3), K[6 + i] * mc(t, 2, 3));
    mr(iKRt, i, 3) = mc(transform, i, 3);
  }
}

// The projection state of a projector seen from `transform`: PinholePointProjector's KRt, or
// SphericalPointProjector's _iT = transform^-1 (sphericalpointprojector.cpp:19-21).
NICP_HD void projector_state(int model, const float *K, const float *transform, float *out) {
  if (model == 1) {
    float t[16];
    for (int i = 0; i < 16; ++i) t[i] = transform[i];
    set_row3(t);
    iso_inverse(t, out);
  } else {
    float iKRt[16];
    projector_matrices(K, transform, out, iKRt);
  }
}

// quat2mat / mat2quat / v2t / t2v (nicp/nicp/bm_se3.h:10-51); mat2quat restates
// Eigen's Quaternion(const Matrix3&) (Shepperd's method), normalize, w >= 0.
template <typename S>
NICP_HD void quat2mat(S qx, S qy, S qz, S *R) {
  S qw = sqrt(S(1) - (qx * qx + (qy * qy + qz * qz)));
  R[0] = qw * qw + qx * qx - qy * qy - qz * qz;
  R[3] = S(2) * (qx * qy - qw * qz);
  R[6] = S(2) * (qx * qz + qw * qy);
  R[1] = S(2) * (qx * qy + qz * qw);
  R[4] = qw * qw - qx * qx + qy * qy - qz * qz;
  R[7] = S(2) * (qy * qz - qx * qw);
  R[2] = S(2) * (qx * qz - qy * qw);
  R[5] = S(2) * (qy * qz + qx * qw);
  R[8] = qw * qw - qx * qx - qy * qy + qz * qz;
}

template <typename S>
NICP_HD void mat2quat(const S *R, S *rq) {
#define RM(r, c) R[(c)*3 + (r)]
  S q[4];
  S t = RM(0, 0) + (RM(1, 1) + RM(2, 2));
  if (t > S(0)) {
    t = sqrt(t + S(1));
    q[3] = S(0.5) * t;
    t = S(0.5) / t;
    q[0] = (RM(2, 1) - RM(1, 2)) * t;
    q[1] = (RM(0, 2) - RM(2, 0)) * t;
    q[2] = (RM(1, 0) - RM(0, 1)) * t;
  } else {
    int i = 0;
    if (RM(1, 1) > RM(0, 0)) i = 1;
    if (RM(2, 2) > RM(i, i)) i = 2;
    int j = (i + 1) % 3, k = (j + 1) % 3;
    t = sqrt(RM(i, i) - RM(j, j) - RM(k, k) + S(1));
    q[i] = S(0.5) * t;
    t = S(0.5) / t;
    q[3] = (RM(k, j) - RM(j, k)) * t;
    q[j] = (RM(j, i) + RM(i, j)) * t;
    q[k] = (RM(k, i) + RM(i, k)) * t;
  }
#undef RM
  S z = (q[0] * q[0] + q[2] * q[2]) + (q[1] * q[1] + q[3] * q[3]);
  if (z > S(0)) {
    S s = sqrt(z);
    for (int i = 0; i < 4; ++i) q[i] = q[i] / s;
  }
  rq[0] = q[0]; rq[1] = q[1]; rq[2] = q[2];
  if (q[3] < S(0)) { rq[0] = -rq[0]; rq[1] = -rq[1]; rq[2] = -rq[2]; }
}

template <typename S>
NICP_HD void v2t(const S *x, S *T) {
  S R[9];
  quat2mat(x[3], x[4], x[5], R);
  for (int i = 0; i < 16; ++i) T[i] = S(0);
  T[15] = S(1);
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) T[c * 4 + r] = R[c * 3 + r];
  T[12] = x[0]; T[13] = x[1]; T[14] = x[2];
}

template <typename S>
NICP_HD void t2v(const S *T, S *x) {
  S R[9];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) R[c * 3 + r] = T[c * 4 + r];
  x[0] = T[12]; x[1] = T[13]; x[2] = T[14];
  mat2quat(R, x + 3);
}

// LDLT with diagonal pivoting and the pseudo-inverse-of-D solve (Eigen's
// LDLT<Matrix6> algorithm, nicp/nicp/alignerprojective.cpp:92), for one 6x6
// column-major system. Returns 0 when the whole diagonal is zero. Every loop is unrolled and
// every swap with the pivot row is written as a swap with each candidate row under the
// condition that it is the pivot, so all array indices are constants (a GPU thread keeps
// the matrix in registers); the arithmetic and its order are those of the plain loops.
template <typename S>
NICP_HD int ldlt_solve6(const S *Hin, const S *rhs, S *x) {
  const int n = 6;
  S m[36];
#pragma unroll
  for (int i = 0; i < 36; ++i) m[i] = Hin[i];
#define A(r, c) m[(c)*6 + (r)]
  int tr[6];
  S temp[6];
  int ok = 1;
  bool stop = false;
#pragma unroll
  for (int k = 0; k < n; ++k) {
    if (stop) break;
    int idx = k;
    S best = fabs(A(k, k));
#pragma unroll
    for (int i = k + 1; i < n; ++i)
      if (fabs(A(i, i)) > best) { best = fabs(A(i, i)); idx = i; }
    tr[k] = idx;
#pragma unroll
    for (int c = k + 1; c < n; ++c) {
      if (c == idx) {
#pragma unroll
        for (int j = 0; j < k; ++j) { S t = A(k, j); A(k, j) = A(c, j); A(c, j) = t; }
#pragma unroll
        for (int i = c + 1; i < n; ++i) { S t = A(i, k); A(i, k) = A(i, c); A(i, c) = t; }
        { S t = A(k, k); A(k, k) = A(c, c); A(c, c) = t; }
#pragma unroll
        for (int i = k + 1; i < c; ++i) { S t = A(i, k); A(i, k) = A(c, i); A(c, i) = t; }
      }
    }
    if (k > 0) {
#pragma unroll
      for (int j = 0; j < k; ++j) temp[j] = A(j, j) * A(k, j);
      S s = A(k, 0) * temp[0];
#pragma unroll
      for (int j = 1; j < k; ++j) s = s + A(k, j) * temp[j];
      A(k, k) -= s;
#pragma unroll
      for (int i = k + 1; i < n; ++i) {
        S t = A(i, 0) * temp[0];
#pragma unroll
        for (int j = 1; j < k; ++j) t = t + A(i, j) * temp[j];
        A(i, k) -= t;
      }
    }
    S akk = A(k, k);
    bool valid = fabs(akk) > S(0);
    if (k == 0 && !valid) {
#pragma unroll
      for (int j = 0; j < n; ++j) tr[j] = j;
      ok = 0;
      stop = true;
      continue;
    }
    if (k + 1 < n && valid) {
#pragma unroll
      for (int i = k + 1; i < n; ++i) A(i, k) /= akk;
    }
  }
  S d[6];
#pragma unroll
  for (int i = 0; i < n; ++i) d[i] = rhs[i];
  // d[i] <-> d[tr[i]] for i ascending (tr[i] >= i)
#pragma unroll
  for (int i = 0; i < n; ++i) {
#pragma unroll
    for (int c = i + 1; c < n; ++c)
      if (c == tr[i]) { S t = d[i]; d[i] = d[c]; d[c] = t; }
  }
#pragma unroll
  for (int i = 0; i < n; ++i)
#pragma unroll
    for (int j = 0; j < i; ++j) d[i] -= A(i, j) * d[j];
  // Eigen uses numeric_limits<S>::min() as the zero-pivot tolerance
  const S tol = sizeof(S) == 8 ? S(2.2250738585072014e-308) : S(1.17549435e-38f);
#pragma unroll
  for (int i = 0; i < n; ++i) d[i] = fabs(A(i, i)) > tol ? d[i] / A(i, i) : S(0);
#pragma unroll
  for (int i = n - 1; i >= 0; --i)
#pragma unroll
    for (int j = i + 1; j < n; ++j) d[i] -= A(j, i) * d[j];
  // d[i] <-> d[tr[i]] for i descending
#pragma unroll
  for (int i = n - 1; i >= 0; --i) {
#pragma unroll
    for (int c = i + 1; c < n; ++c)
      if (c == tr[i]) { S t = d[i]; d[i] = d[c]; d[c] = t; }
  }
#undef A
#pragma unroll
  for (int i = 0; i < n; ++i) x[i] = d[i];
  return ok;
}

}  // namespace nicp_math
