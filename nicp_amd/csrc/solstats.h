// Solution statistics of Aligner::_computeStatistics (nicp/nicp/aligner.cpp:82-135,
// nicp/nicp/unscented.h:11-65) in FP64, for one pair, on host or device: the pseudo-inverse
// of the final linearisation's H (JacobiSVD's default threshold for a Matrix6f), 13 unscented
// sigma points remapped by t2v(T v2t(s)^-1), the reconstructed Gaussian, Omega = Sigma^-1
// symmetrised and the singular-value ratios of its translational and rotational 3x3 blocks.
// Every loop has constant bounds, so a device thread keeps its arrays in registers or its
// private stack; the device runs the operations of the host in the same order, so both give the
// same bits (FP64 add, multiply, divide and square root are correctly rounded on both, and the
// library is compiled without contraction).
#pragma once
#include <float.h>
#include <math.h>

#include "nicp_math.h"

namespace nicp_stats {

using nicp_math::t2v;
using nicp_math::v2t;

// doubles per pair of the statistics output: Omega (36, column-major), mean (6),
// translational ratio, rotational ratio, valid (0/1), 3 spare
constexpr int kSolStats = 48;
constexpr int kSolOmega = 0, kSolMean = 36, kSolTrans = 42, kSolRot = 43, kSolValid = 44;

// symmetric eigen-decomposition of an n x n column-major matrix (n <= 6) by cyclic Jacobi
// rotations; eigenvalues ascending, eigenvector k in V[k*n .. k*n + n)
template <int N>
NICP_HD void sym_eigen(const double *Ain, double *w, double *V) {
  double a[N * N];
  for (int i = 0; i < N * N; ++i) a[i] = Ain[i];
  for (int i = 0; i < N * N; ++i) V[i] = 0.0;
  for (int i = 0; i < N; ++i) V[i * N + i] = 1.0;
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int c = 0; c < N; ++c)
      for (int r = 0; r < N; ++r) {
        const double x = a[c * N + r] * a[c * N + r];
        if (r == c) diag += x;
        else off += x;
      }
    if (off <= 1e-32 * diag || off == 0.0) break;
    for (int p = 0; p < N - 1; ++p)
      for (int q = p + 1; q < N; ++q) {
        const double apq = a[q * N + p];
        if (apq == 0.0) continue;
        const double th = (a[q * N + q] - a[p * N + p]) / (2.0 * apq);
        const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
        for (int k = 0; k < N; ++k) {
          const double x = a[p * N + k], y = a[q * N + k];
          a[p * N + k] = c * x - sn * y;
          a[q * N + k] = sn * x + c * y;
        }
        for (int k = 0; k < N; ++k) {
          const double x = a[k * N + p], y = a[k * N + q];
          a[k * N + p] = c * x - sn * y;
          a[k * N + q] = sn * x + c * y;
        }
        for (int k = 0; k < N; ++k) {
          const double x = V[p * N + k], y = V[q * N + k];
          V[p * N + k] = c * x - sn * y;
          V[q * N + k] = sn * x + c * y;
        }
      }
  }
  for (int i = 0; i < N; ++i) w[i] = a[i * N + i];
  // selection sort, ascending
  for (int i = 0; i < N; ++i) {
    int m = i;
    for (int j = i + 1; j < N; ++j)
      if (w[j] < w[m]) m = j;
    if (m != i) {
      const double tw = w[i];
      w[i] = w[m];
      w[m] = tw;
      for (int k = 0; k < N; ++k) {
        const double tv = V[i * N + k];
        V[i * N + k] = V[m * N + k];
        V[m * N + k] = tv;
      }
    }
  }
}

// Cholesky factor L (column-major, lower) of a 6x6; false when not positive definite
NICP_HD bool cholesky6(const double *A, double *L) {
  for (int i = 0; i < 36; ++i) L[i] = 0.0;
  for (int j = 0; j < 6; ++j) {
    double d = A[j * 6 + j];
    for (int k = 0; k < j; ++k) d -= L[k * 6 + j] * L[k * 6 + j];
    if (!(d > 0.0)) return false;
    d = sqrt(d);
    L[j * 6 + j] = d;
    for (int i = j + 1; i < 6; ++i) {
      double v = A[j * 6 + i];
      for (int k = 0; k < j; ++k) v -= L[k * 6 + i] * L[k * 6 + j];
      L[j * 6 + i] = v / d;
    }
  }
  return true;
}

// inverse of a symmetric positive definite 6x6 through its eigen-decomposition
NICP_HD bool spd_inverse6(const double *A, double *Ai) {
  double w[6], V[36];
  sym_eigen<6>(A, w, V);
  if (!(w[0] > 0.0)) return false;
  for (int c = 0; c < 6; ++c)
    for (int r = 0; r < 6; ++r) {
      double s = 0.0;
      for (int k = 0; k < 6; ++k) s += V[k * 6 + r] * V[k * 6 + c] / w[k];
      Ai[c * 6 + r] = s;
    }
  return true;
}

// largest over smallest singular value of the symmetric 3x3 block at (off, off)
NICP_HD double block_ratio(const double *O, int off) {
  double a[9], w[3], V[9];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) a[c * 3 + r] = O[(off + c) * 6 + off + r];
  sym_eigen<3>(a, w, V);
  double s0 = fabs(w[0]), s1 = fabs(w[1]), s2 = fabs(w[2]);
  const double lo = fmin(s0, fmin(s1, s2)), hi = fmax(s0, fmax(s1, s2));
  return hi / lo;
}

// H: column-major 6x6 of the statistics linearisation; T: the final transform (column-major
// float 4x4); out: kSolStats doubles (layout above)
NICP_HD void solution_statistics(const double *H, const float *Tf, float trans_max_ratio, float rot_max_ratio,
                                 double *out) {
  // localSigma = pinv(H) with JacobiSVD's default threshold for Matrix6f (6 * FLT_EPSILON * s_max)
  double w[6], V[36], Sigma[36];
  sym_eigen<6>(H, w, V);
  double smax = 0.0;
  for (int i = 0; i < 6; ++i) smax = fmax(smax, fabs(w[i]));
  const double thr = fmax(smax * 6.0 * (double)FLT_EPSILON, 2.2250738585072014e-308);
  for (int i = 0; i < 36; ++i) Sigma[i] = 0.0;
  bool zero = true;
  for (int i = 0; i < 36; ++i) zero = zero && H[i] == 0.0;
  for (int k = 0; k < 6; ++k) {
    if (fabs(w[k]) < thr) continue;
    for (int c = 0; c < 6; ++c)
      for (int r = 0; r < 6; ++r) Sigma[c * 6 + r] += V[k * 6 + r] * V[k * 6 + c] / w[k];
  }
  // unscented transform: 2 dim + 1 sigma points (alpha 1e-3, beta 2)
  const int dim = 6;
  const double alpha = 1e-3, beta = 2.0, lam = alpha * alpha * dim;
  double C[36], L[36];
  for (int i = 0; i < 36; ++i) C[i] = Sigma[i] * (dim + lam);
  const bool chol = cholesky6(C, L);
  double smp[13][6], wi[13], wp[13];
  for (int k = 0; k < 6; ++k) smp[0][k] = 0.0;
  wi[0] = lam / (dim + lam);
  wp[0] = lam / (dim + lam) + (1.0 - alpha * alpha + beta);
  for (int i = 0; i < dim; ++i)
    for (int sgn = 0; sgn < 2; ++sgn) {
      const int j = 1 + 2 * i + sgn;
      for (int k = 0; k < 6; ++k) smp[j][k] = sgn ? -L[i * 6 + k] : L[i * 6 + k];
      wi[j] = wp[j] = 1.0 / (2.0 * (dim + lam));
    }
  double T[16];
  for (int i = 0; i < 16; ++i) T[i] = Tf[i];
  for (int j = 0; j < 13; ++j) {
    double X[16], Y[16];
    v2t<double>(smp[j], X);
    // Y = T * X^-1
    double Ri[9], ti[3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) Ri[c * 3 + r] = X[r * 4 + c];
    for (int r = 0; r < 3; ++r) ti[r] = -(Ri[r] * X[12] + Ri[3 + r] * X[13] + Ri[6 + r] * X[14]);
    for (int i = 0; i < 16; ++i) Y[i] = 0.0;
    Y[15] = 1.0;
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) Y[c * 4 + r] = T[r] * Ri[c * 3] + T[4 + r] * Ri[c * 3 + 1] + T[8 + r] * Ri[c * 3 + 2];
      Y[12 + r] = T[r] * ti[0] + T[4 + r] * ti[1] + T[8 + r] * ti[2] + T[12 + r];
    }
    t2v<double>(Y, smp[j]);
  }
  double mean[6] = {0, 0, 0, 0, 0, 0}, cov[36];
  for (int j = 0; j < 13; ++j)
    for (int k = 0; k < 6; ++k) mean[k] += wi[j] * smp[j][k];
  for (int i = 0; i < 36; ++i) cov[i] = 0.0;
  for (int j = 0; j < 13; ++j) {
    double d[6];
    for (int k = 0; k < 6; ++k) d[k] = smp[j][k] - mean[k];
    for (int c = 0; c < 6; ++c)
      for (int r = 0; r < 6; ++r) cov[c * 6 + r] += wp[j] * d[r] * d[c];
  }
  double Om[36];
  const bool inv = spd_inverse6(cov, Om);
  double *omega = out + kSolOmega;
  for (int c = 0; c < 6; ++c)
    for (int r = 0; r < 6; ++r) omega[c * 6 + r] = inv ? 0.5 * (Om[c * 6 + r] + Om[r * 6 + c]) : 0.0;
  for (int k = 0; k < 6; ++k) out[kSolMean + k] = mean[k];
  const double tr = inv ? block_ratio(omega, 0) : INFINITY;
  const double rr = inv ? block_ratio(omega, 3) : INFINITY;
  out[kSolTrans] = tr;
  out[kSolRot] = rr;
  const bool finite = isfinite(tr) && isfinite(rr);
  out[kSolValid] = (!zero && chol && inv && finite && !(rr > (double)rot_max_ratio) && !(tr > (double)trans_max_ratio))
                       ? 1.0
                       : 0.0;
  out[45] = out[46] = out[47] = 0.0;
}

}  // namespace nicp_stats
