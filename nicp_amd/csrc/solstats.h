// Solution statistics of Aligner::_computeStatistics (nicp/nicp/aligner.cpp:82-135,
// nicp/nicp/unscented.h:11-65) in FP64, one wavefront per pair:
//  1. localSigma = pseudo-inverse of the final linearisation's H with JacobiSVD's default
//     threshold for a Matrix6f (singular values below 6 FLT_EPSILON s_max are dropped), from an
//     eigen-decomposition by parallel cyclic Jacobi: lane c*6 + r holds H(r, c) and V(r, c);
//     every round applies three disjoint rotations of a round-robin ordering at once;
//  2. the 13 unscented sigma points (alpha 1e-3, beta 2) from the Cholesky factor of
//     (dim + lambda) Sigma, lane j remapping point j by t2v(T v2t(s)^-1);
//  3. the reconstructed Gaussian (mean, covariance, in the point order of reconstructGaussian),
//     Omega = covariance^-1 symmetrised, and the largest over smallest singular value of its
//     translational and rotational 3x3 blocks.
// Steps 2-3 run the same arithmetic in every lane (values broadcast from the lanes that hold
// them), so every array index is a constant and stays in registers.
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

// value of x in lane `src` (the same lane for the whole wavefront)
__device__ __forceinline__ double bcast(double x, int src) {
  const unsigned long long b = __double_as_longlong(x);
  const unsigned int lo = __builtin_amdgcn_readlane((unsigned int)b, src);
  const unsigned int hi = __builtin_amdgcn_readlane((unsigned int)(b >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// the Jacobi rotation (c, s) that zeroes a_pq of a symmetric matrix; (1, 0) when a_pq is zero
__device__ __forceinline__ void jacobi_cs(double app, double aqq, double apq, double &c, double &s) {
  const bool rot = apq != 0.0;
  const double th = (aqq - app) / (2.0 * (rot ? apq : 1.0));
  const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
  const double c0 = 1.0 / sqrt(t * t + 1.0);
  c = rot ? c0 : 1.0;
  s = rot ? t * c0 : 0.0;
}

// the partner of index i (0..5) in round k (0..4) of the round-robin ordering: positions
// (0, 1 + (k mod 5), ..., 1 + (k + 4 mod 5)), paired first with last, second with fifth, third
// with fourth; the five rounds pair every index with every other once
__device__ __forceinline__ int rr_partner(int i, int k) {
  const int m = i == 0 ? 0 : ((i - 1 - k + 10) % 5) + 1;
  const int pm = 5 - m;
  return pm == 0 ? 0 : 1 + ((pm - 1 + k) % 5);
}

// serial 3x3 symmetric eigenvalues by cyclic Jacobi (constant indices)
__device__ __forceinline__ void sym_eigvals3(const double *Ain, double *w) {
  double a[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) a[i] = Ain[i];
  for (int sweep = 0; sweep < 64; ++sweep) {
    const double off = a[3] * a[3] + a[6] * a[6] + a[7] * a[7];
    const double diag = a[0] * a[0] + a[4] * a[4] + a[8] * a[8];
    if (off <= 1e-32 * diag || off == 0.0) break;
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int q = p + 1; q < 3; ++q) {
        double c, s;
        jacobi_cs(a[p * 4], a[q * 4], a[q * 3 + p], c, s);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double x = a[p * 3 + k], y = a[q * 3 + k];
          a[p * 3 + k] = c * x - s * y;
          a[q * 3 + k] = s * x + c * y;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double x = a[k * 3 + p], y = a[k * 3 + q];
          a[k * 3 + p] = c * x - s * y;
          a[k * 3 + q] = s * x + c * y;
        }
      }
  }
  w[0] = a[0];
  w[1] = a[4];
  w[2] = a[8];
}

// Cholesky factor L (column-major, lower) of a symmetric 6x6; false when not positive definite
__device__ __forceinline__ bool cholesky6(const double *A, double *L) {
#pragma unroll
  for (int i = 0; i < 36; ++i) L[i] = 0.0;
  bool ok = true;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    double d = A[j * 6 + j];
#pragma unroll
    for (int k = 0; k < j; ++k) d -= L[k * 6 + j] * L[k * 6 + j];
    ok = ok && d > 0.0;
    d = sqrt(ok ? d : 1.0);
    L[j * 6 + j] = d;
#pragma unroll
    for (int i = j + 1; i < 6; ++i) {
      double v = A[j * 6 + i];
#pragma unroll
      for (int k = 0; k < j; ++k) v -= L[k * 6 + i] * L[k * 6 + j];
      L[j * 6 + i] = v / d;
    }
  }
  return ok;
}

// inverse of a symmetric positive definite 6x6 as L^-T L^-1 from its Cholesky factor
__device__ __forceinline__ bool spd_inverse6(const double *A, double *Ai) {
  double L[36], M[36];
  const bool ok = cholesky6(A, L);
  // M = L^-1 (lower), column by column
#pragma unroll
  for (int c = 0; c < 6; ++c)
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      if (r < c) {
        M[c * 6 + r] = 0.0;
        continue;
      }
      double v = r == c ? 1.0 : 0.0;
#pragma unroll
      for (int k = c; k < r; ++k) v -= L[k * 6 + r] * M[c * 6 + k];
      M[c * 6 + r] = v / L[r * 6 + r];
    }
  // (L^-T L^-1)(r, c) = sum over k of M(k, r) M(k, c)
#pragma unroll
  for (int c = 0; c < 6; ++c)
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 6; ++k) s += M[r * 6 + k] * M[c * 6 + k];
      Ai[c * 6 + r] = s;
    }
  return ok;
}

// largest over smallest singular value of the symmetric 3x3 block at (off, off)
__device__ __forceinline__ double block_ratio(const double *O, int off) {
  double a[9], w[3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) a[c * 3 + r] = O[(off + c) * 6 + off + r];
  sym_eigvals3(a, w);
  const double s0 = fabs(w[0]), s1 = fabs(w[1]), s2 = fabs(w[2]);
  return fmax(s0, fmax(s1, s2)) / fmin(s0, fmin(s1, s2));
}

// One wavefront (64 lanes, all active) per pair. H: column-major 6x6 of the statistics
// linearisation; Tf: the final transform (column-major float 4x4); out: kSolStats doubles
// (layout above), written by lane 0.
__device__ void solution_statistics_wave(const double *H, const float *Tf, float trans_max_ratio, float rot_max_ratio,
                                         double *out) {
  const int lane = threadIdx.x & 63;
  const bool in = lane < 36;
  const int r = in ? lane % 6 : 0, c = in ? lane / 6 : 0;
  double a = in ? H[lane] : 0.0;
  const bool zero = __all(a == 0.0);
  double v = in && r == c ? 1.0 : 0.0;
  // 1. eigen-decomposition of H
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = in && r != c ? a * a : 0.0, diag = in && r == c ? a * a : 0.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      off += __shfl_xor(off, o, 64);
      diag += __shfl_xor(diag, o, 64);
    }
    // lane 0's sums decide for the whole wavefront
    off = bcast(off, 0);
    diag = bcast(diag, 0);
    if (off <= 1e-32 * diag || off == 0.0) break;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const int pr = rr_partner(r, k), pc = rr_partner(c, k);
      double cr, sr, cc, sc;
      {
        const int p = min(r, pr), q = max(r, pr);
        jacobi_cs(__shfl(a, p * 7, 64), __shfl(a, q * 7, 64), __shfl(a, q * 6 + p, 64), cr, sr);
      }
      {
        const int p = min(c, pc), q = max(c, pc);
        jacobi_cs(__shfl(a, p * 7, 64), __shfl(a, q * 7, 64), __shfl(a, q * 6 + p, 64), cc, sc);
      }
      // columns c and pc: column p' = c col_p - s col_q, column q' = s col_p + c col_q
      const double ao = __shfl(a, pc * 6 + r, 64), vo = __shfl(v, pc * 6 + r, 64);
      const double a1 = c < pc ? cc * a - sc * ao : sc * ao + cc * a;
      v = c < pc ? cc * v - sc * vo : sc * vo + cc * v;
      // rows r and pr, likewise
      const double a1o = __shfl(a1, c * 6 + pr, 64);
      a = r < pr ? cr * a1 - sr * a1o : sr * a1o + cr * a1;
    }
  }
  // pseudo-inverse: eigenvalue k in lane 7k, eigenvector k in lanes 6k .. 6k + 5
  double w[6], smax = 0.0;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    w[k] = bcast(a, 7 * k);
    smax = fmax(smax, fabs(w[k]));
  }
  const double thr = fmax(smax * 6.0 * (double)FLT_EPSILON, 2.2250738585072014e-308);
  double sig = 0.0;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const double vr = __shfl(v, k * 6 + r, 64), vc = __shfl(v, k * 6 + c, 64);
    if (!(fabs(w[k]) < thr)) sig += vr * vc / w[k];
  }
  // 2. sigma points
  const int dim = 6;
  const double alpha = 1e-3, beta = 2.0, lam = alpha * alpha * dim;
  double C[36], L[36];
#pragma unroll
  for (int i = 0; i < 36; ++i) C[i] = bcast(sig, i) * (dim + lam);
  const bool chol = cholesky6(C, L);
  // lane j < 13 takes point j: 0 the mean (zero), 2i + 1 and 2i + 2 the mean +/- column i of L
  double smp[6];
  const int j = lane < 13 ? lane : 0;
  const int ci = j > 0 ? (j - 1) >> 1 : 0;
  const double sg = j == 0 ? 0.0 : (((j - 1) & 1) ? -1.0 : 1.0);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double col = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) col = ci == i ? L[i * 6 + k] : col;
    smp[k] = sg * col;
  }
  {
    double T[16], X[16], Y[16], Ri[9], ti[3];
#pragma unroll
    for (int i = 0; i < 16; ++i) T[i] = Tf[i];
    v2t<double>(smp, X);
    // Y = T * X^-1
#pragma unroll
    for (int rr = 0; rr < 3; ++rr)
#pragma unroll
      for (int k = 0; k < 3; ++k) Ri[k * 3 + rr] = X[rr * 4 + k];
#pragma unroll
    for (int rr = 0; rr < 3; ++rr) ti[rr] = -(Ri[rr] * X[12] + Ri[3 + rr] * X[13] + Ri[6 + rr] * X[14]);
#pragma unroll
    for (int i = 0; i < 16; ++i) Y[i] = 0.0;
    Y[15] = 1.0;
#pragma unroll
    for (int rr = 0; rr < 3; ++rr) {
#pragma unroll
      for (int k = 0; k < 3; ++k) Y[k * 4 + rr] = T[rr] * Ri[k * 3] + T[4 + rr] * Ri[k * 3 + 1] + T[8 + rr] * Ri[k * 3 + 2];
      Y[12 + rr] = T[rr] * ti[0] + T[4 + rr] * ti[1] + T[8 + rr] * ti[2] + T[12 + rr];
    }
    t2v<double>(Y, smp);
  }
  // 3. reconstructGaussian in point order, every lane on the broadcast points
  const double w0i = lam / (dim + lam), w0p = lam / (dim + lam) + (1.0 - alpha * alpha + beta);
  const double wj = 1.0 / (2.0 * (dim + lam));
  double mean[6] = {0, 0, 0, 0, 0, 0};
  for (int p = 0; p < 13; ++p) {
    const double wi = p == 0 ? w0i : wj;
#pragma unroll
    for (int k = 0; k < 6; ++k) mean[k] += wi * bcast(smp[k], p);
  }
  double cov[36];
#pragma unroll
  for (int i = 0; i < 36; ++i) cov[i] = 0.0;
  for (int p = 0; p < 13; ++p) {
    const double wp = p == 0 ? w0p : wj;
    double d[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) d[k] = bcast(smp[k], p) - mean[k];
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
      for (int rr = 0; rr < 6; ++rr) cov[k * 6 + rr] += wp * d[rr] * d[k];
  }
  double Om[36], O[36];
  const bool inv = spd_inverse6(cov, Om);
#pragma unroll
  for (int k = 0; k < 6; ++k)
#pragma unroll
    for (int rr = 0; rr < 6; ++rr) O[k * 6 + rr] = inv ? 0.5 * (Om[k * 6 + rr] + Om[rr * 6 + k]) : 0.0;
  const double tr = inv ? block_ratio(O, 0) : INFINITY;
  const double rt = inv ? block_ratio(O, 3) : INFINITY;
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 36; ++i) out[kSolOmega + i] = O[i];
#pragma unroll
    for (int k = 0; k < 6; ++k) out[kSolMean + k] = mean[k];
    out[kSolTrans] = tr;
    out[kSolRot] = rt;
    const bool finite = isfinite(tr) && isfinite(rt);
    out[kSolValid] =
        (!zero && chol && inv && finite && !(rt > (double)rot_max_ratio) && !(tr > (double)trans_max_ratio)) ? 1.0
                                                                                                              : 0.0;
    out[45] = out[46] = out[47] = 0.0;
  }
}

}  // namespace nicp_stats
