// Gaussian3f records, Cloud::transformInPlace / Cloud::add and Merger::merge on the device.
// Included at the end of nicp_capi.hip: it uses the context, cloud, buffer and error
// handling defined there.
//
// A Gaussian record is the nicp_gaussian struct, 7 float4 per point:
//   [0..2] mean, [3] flags (bit 0 moments current, bit 1 information form current),
//   [4..12] covariance, [13..21] information matrix, [22..24] information vector, [25..27] pad.
// The flags restate the lazy evaluation of nicp/nicp/gaussian.h:104-122: a Gaussian made from
// moments computes its information form the first time it is added to or added from, and a
// Gaussian that received information computes its moments when they are read. Which form is
// current changes later results (an information form accumulated by earlier merges is used as
// is, not recomputed from the covariance), so the records carry both forms and the flags.

namespace mergek {

constexpr int kGF4 = 7;
constexpr int gMean = 0, gFlags = 3, gCov = 4, gInfo = 13, gIvec = 22;
constexpr int kMoments = 1, kInfoForm = 2;

__device__ __forceinline__ void g_load(const float4 *__restrict__ g, float *r) {
#pragma unroll
  for (int k = 0; k < kGF4; ++k) {
    const float4 v = g[k];
    r[4 * k] = v.x;
    r[4 * k + 1] = v.y;
    r[4 * k + 2] = v.z;
    r[4 * k + 3] = v.w;
  }
}
__device__ __forceinline__ void g_store(float4 *__restrict__ g, const float *r) {
#pragma unroll
  for (int k = 0; k < kGF4; ++k) g[k] = make_float4(r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3]);
}
__device__ __forceinline__ int g_flags(const float *r) { return __float_as_int(r[gFlags]); }
__device__ __forceinline__ void g_set_flags(float *r, int f) { r[gFlags] = __int_as_float(f); }

// Matrix3f * Vector3f, Eigen's coefficient-based product: o_i = M(i,0) v0 + (M(i,1) v1 + M(i,2) v2)
__device__ __forceinline__ void mat3_vec(const float *M, const float *v, float *o) {
#pragma unroll
  for (int i = 0; i < 3; ++i) o[i] = red3(M[i] * v[0], M[3 + i] * v[1], M[6 + i] * v[2]);
}

// Gaussian::_updateInfo (gaussian.h:116-122): information = covariance^-1, vector = information * mean
__device__ __forceinline__ void ensure_info(float *r) {
  if (g_flags(r) & kInfoForm) return;
  inverse3(r + gCov, r + gInfo);
  mat3_vec(r + gInfo, r + gMean, r + gIvec);
  g_set_flags(r, g_flags(r) | kInfoForm);
}

// Gaussian::_updateMoments (gaussian.h:108-114): covariance = information^-1, mean = covariance * vector
__device__ __forceinline__ void ensure_moments(float *r) {
  if (g_flags(r) & kMoments) return;
  inverse3(r + gInfo, r + gCov);
  mat3_vec(r + gCov, r + gIvec, r + gMean);
  g_set_flags(r, g_flags(r) | kMoments);
}

struct GaussCam {
  float iK[9];    // PinholePointProjector::_iK (column-major)
  float fB;       // _baseline * _cameraMatrix(0, 0)
  float alpha;
  int spherical;  // SphericalPointProjector: identity covariance (sphericalpointprojector.cpp:112-113)
};

struct Xform {
  float m[16];   // Cloud::transformInPlace's matrix with its last row 0 0 0 1
  float mz[16];  // the same with the last row and column zeroed (information matrices)
};

// Omega' = (Tz Omega) Tz^T with 4x4 lazy packet products (P1 per column) on the symmetric
// matrix given by its upper triangle u (00 01 02 11 12 22); returns the upper triangle.
__device__ __forceinline__ void omega_transform(const float *Tz, const float *u, float *o6) {
  const float W[16] = {u[0], u[1], u[2], 0.0f, u[1], u[3], u[4], 0.0f,
                       u[2], u[4], u[5], 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  float A[16], C[16];
#pragma unroll
  for (int j = 0; j < 4; ++j) mul4_packet(Tz, W[4 * j], W[4 * j + 1], W[4 * j + 2], W[4 * j + 3], A + 4 * j);
#pragma unroll
  for (int j = 0; j < 4; ++j) mul4_packet(A, mc(Tz, j, 0), mc(Tz, j, 1), mc(Tz, j, 2), mc(Tz, j, 3), C + 4 * j);
  o6[0] = mc(C, 0, 0);
  o6[1] = mc(C, 0, 1);
  o6[2] = mc(C, 0, 2);
  o6[3] = mc(C, 1, 1);
  o6[4] = mc(C, 1, 2);
  o6[5] = mc(C, 2, 2);
}

}  // namespace mergek

using namespace mergek;

// The Gaussian half of PinholePointProjector::unProject (pinholepointprojector.cpp:103-131),
// one thread per pixel of a converted frame.
extern "C" __global__ __launch_bounds__(kThreads) void k_gauss_unproject(const FrameDev *__restrict__ F, GaussCam gc,
                                                                        float4 *__restrict__ gauss) {
  const FrameDev &fd = F[0];
  const int P = fd.rows * fd.cols;
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= P) return;
  const int idx = fd.index[p];
  if (idx < 0) return;
  float rec[kGF4 * 4];
#pragma unroll
  for (int k = 0; k < kGF4 * 4; ++k) rec[k] = 0.0f;
  const float4 pt = fd.pts[idx];
  rec[gMean] = pt.x;
  rec[gMean + 1] = pt.y;
  rec[gMean + 2] = pt.z;
  g_set_flags(rec, kMoments);
  if (gc.spherical) {
    rec[gCov] = rec[gCov + 4] = rec[gCov + 8] = 1.0f;
    g_store(gauss + (size_t)kGF4 * idx, rec);
    return;
  }
  const float z = frame_depth(fd, p);
  const int r = p / fd.cols, c = p - r * fd.cols;
  const float zvar = (gc.alpha * z * z) / (gc.fB + z * gc.alpha);
  // J0 = [[z, 0, c], [0, z, r], [0, 0, 1]] (column-major), J = iK * J0
  const float J0[9] = {z, 0.0f, 0.0f, 0.0f, z, 0.0f, (float)c, (float)r, 1.0f};
  float J[9];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      J[j * 3 + i] = red3(gc.iK[i] * J0[j * 3], gc.iK[3 + i] * J0[j * 3 + 1], gc.iK[6 + i] * J0[j * 3 + 2]);
  const float D[3] = {3.0f, 3.0f, zvar};
  // covariance = (J diag(D)) J^T
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      rec[gCov + j * 3 + i] = red3((J[i] * D[0]) * J[j], (J[3 + i] * D[1]) * J[3 + j], (J[6 + i] * D[2]) * J[6 + j]);
  g_store(gauss + (size_t)kGF4 * idx, rec);
}

// Cloud::transformInPlace (cloud.cpp:193-205), one thread per point.
extern "C" __global__ __launch_bounds__(kThreads) void k_cloud_transform(int n, Xform X, float4 *__restrict__ pts,
                                                                        float4 *__restrict__ nrm,
                                                                        float4 *__restrict__ omg,
                                                                        float4 *__restrict__ stats,
                                                                        float4 *__restrict__ gauss) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float *m = X.m;
  float o[4];
  // PointVector / NormalVector: *t = m * (*t), then w = 1 / 0 (homogeneousvector4f.h);
  // the normal's w slot holds the curvature here and is left alone
  const float4 p = pts[i];
  mul4_packet(m, p.x, p.y, p.z, p.w, o);
  pts[i] = make_float4(o[0], o[1], o[2], 1.0f);
  const float4 q = nrm[i];
  mul4_packet(m, q.x, q.y, q.z, 0.0f, o);
  nrm[i] = make_float4(o[0], o[1], o[2], q.w);
  // InformationMatrixVector::transformInPlace (informationmatrix.h:111-121)
  {
    float4 *w = omg + 3 * (size_t)i;
    const float4 a = w[0], b = w[1], c = w[2];
    const float up[6] = {a.x, a.y, a.z, a.w, b.x, b.y}, un[6] = {b.z, b.w, c.x, c.y, c.z, c.w};
    float tp[6], tn[6];
    omega_transform(X.mz, up, tp);
    omega_transform(X.mz, un, tn);
    w[0] = make_float4(tp[0], tp[1], tp[2], tp[3]);
    w[1] = make_float4(tp[4], tp[5], tn[0], tn[1]);
    w[2] = make_float4(tn[2], tn[3], tn[4], tn[5]);
  }
  // StatsVector::transformInPlace (stats.h:121-126): the Stats matrix [eigenvectors | mean]
  // (row 3 = 0 0 0 1) becomes m * S; eigenvalues, n and curvature stay
  if (stats) {
    float4 *s4 = stats + 4 * (size_t)i;
    float s[16];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float4 v = s4[k];
      s[4 * k] = v.x;
      s[4 * k + 1] = v.y;
      s[4 * k + 2] = v.z;
      s[4 * k + 3] = v.w;
    }
#pragma unroll
    for (int col = 0; col < 3; ++col) {
      mul4_packet(m, s[col * 3], s[col * 3 + 1], s[col * 3 + 2], 0.0f, o);
      s[col * 3] = o[0];
      s[col * 3 + 1] = o[1];
      s[col * 3 + 2] = o[2];
    }
    mul4_packet(m, s[12], s[13], s[14], 1.0f, o);
    s[12] = o[0];
    s[13] = o[1];
    s[14] = o[2];
#pragma unroll
    for (int k = 0; k < 4; ++k) s4[k] = make_float4(s[4 * k], s[4 * k + 1], s[4 * k + 2], s[4 * k + 3]);
  }
  // Gaussian3fVector::transformInPlace (gaussian3.h:59-67): moments (R mean + t, R cov R^T)
  if (gauss) {
    float r[kGF4 * 4];
    float4 *g = gauss + (size_t)kGF4 * i;
    g_load(g, r);
    ensure_moments(r);
    float mean[3], RC[9];
#pragma unroll
    for (int a = 0; a < 3; ++a)
      mean[a] = red3(mc(m, a, 0) * r[gMean], mc(m, a, 1) * r[gMean + 1], mc(m, a, 2) * r[gMean + 2]) + mc(m, a, 3);
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int a = 0; a < 3; ++a)
        RC[k * 3 + a] = red3(mc(m, a, 0) * r[gCov + k * 3], mc(m, a, 1) * r[gCov + k * 3 + 1],
                             mc(m, a, 2) * r[gCov + k * 3 + 2]);
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int a = 0; a < 3; ++a)
        r[gCov + b * 3 + a] = red3(RC[a] * mc(m, b, 0), RC[3 + a] * mc(m, b, 1), RC[6 + a] * mc(m, b, 2));
    r[gMean] = mean[0];
    r[gMean + 1] = mean[1];
    r[gMean + 2] = mean[2];
    g_set_flags(r, kMoments);
    g_store(g, r);
  }
}

// ---- Merger::merge (merger.cpp:15-109) ----

extern "C" __global__ __launch_bounds__(kThreads) void k_merge_init(int n, int *__restrict__ count) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) count[i] = 0;
}

// merger.cpp:32-72: the decision of every point. collapsed[i] = -1 (kept as is), i (it owns its
// pixel), or the pixel's owner whose Gaussian takes its information; slot[i] numbers the points
// folded into the same owner (in arbitrary order; the owner sorts them).
extern "C" __global__ __launch_bounds__(kThreads) void k_merge_classify(
    int n, Xform X, int model, SphDev sph, float min_d, float max_d, int rows, int cols, float dist_thr, float max_depth,
    const float4 *__restrict__ pts, const int *__restrict__ index_image, const float *__restrict__ depth_image,
    int *__restrict__ collapsed, int *__restrict__ count, int *__restrict__ slot) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  // PinholePointProjector::_project (pinholepointprojector.h:253-262) with the KRt in X.m:
  // depth is set first, the pixel only when the depth passes the projector's range
  const float4 p = pts[i];
  float d;
  int r = -1, c = -1;
  if (model == kSpherical) {
    int x, y;
    if (sph_project(sph, X.m, p, min_d, max_d, x, y, d)) {
      c = x;
      r = y;
    }
  } else {
    float ip[4];
    mul4_packet(X.m, p.x, p.y, p.z, p.w, ip);
    d = ip[2];
    if (!(d < min_d || d > max_d)) {
      const float inv = 1.0f / d;
      c = (int)roundf(ip[0] * inv);
      r = (int)roundf(ip[1] * inv);
    }
  }
  int out = -1;
  if (!(d < 0.0f || d > max_depth || r < 0 || r >= rows || c < 0 || c >= cols)) {
    const int pix = r * cols + c;
    const int t = index_image[pix];
    if (t == i) {
      out = i;
    } else if (t >= 0 && fabsf(d - depth_image[pix]) < dist_thr) {
      out = t;
      slot[i] = atomicAdd(count + t, 1);
    }
  }
  collapsed[i] = out;
}

// the start of every owner's list of folded points
extern "C" __global__ __launch_bounds__(kThreads) void k_merge_offsets(int n, const int *__restrict__ count,
                                                                      int *__restrict__ off, int *__restrict__ total) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const int c = count[i];
  if (c > 0) off[i] = atomicAdd(total, c);
}

extern "C" __global__ __launch_bounds__(kThreads) void k_merge_bucket(int n, const int *__restrict__ collapsed,
                                                                     const int *__restrict__ slot,
                                                                     const int *__restrict__ off,
                                                                     int *__restrict__ bucket) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const int t = collapsed[i];
  if (t >= 0 && t != i) bucket[off[t] + slot[i]] = i;
}

// merger.cpp:65-79: every owner adds the information of the points folded into it in
// increasing point order (Gaussian::addInformation, gaussian.h:50-56), and a point that owns
// its pixel moves to its Gaussian's mean. A folded point is never an owner (it does not win
// its own pixel), so the Gaussians read here are not written by any thread.
extern "C" __global__ __launch_bounds__(kThreads) void k_merge_accumulate(
    int n, const int *__restrict__ collapsed, const int *__restrict__ count, const int *__restrict__ off,
    const int *__restrict__ bucket, float4 *__restrict__ gauss, float4 *__restrict__ pts) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const int cnt = count[i];
  const bool own = collapsed[i] == i;
  if (cnt == 0 && !own) return;
  float r[kGF4 * 4];
  g_load(gauss + (size_t)kGF4 * i, r);
  if (cnt > 0) {
    ensure_info(r);
    const int *lst = bucket + off[i];
    int last = -1;
    for (int k = 0; k < cnt; ++k) {
      // the next folded point in point order
      int s = 0x7fffffff;
      for (int j = 0; j < cnt; ++j) {
        const int v = lst[j];
        if (v > last && v < s) s = v;
      }
      last = s;
      float g[kGF4 * 4];
      g_load(gauss + (size_t)kGF4 * s, g);
      ensure_info(g);
#pragma unroll
      for (int e = 0; e < 9; ++e) r[gInfo + e] = r[gInfo + e] + g[gInfo + e];
#pragma unroll
      for (int e = 0; e < 3; ++e) r[gIvec + e] = r[gIvec + e] + g[gIvec + e];
    }
    g_set_flags(r, kInfoForm);
  }
  if (own) {
    ensure_moments(r);
    const float4 p = pts[i];
    pts[i] = make_float4(r[gMean], r[gMean + 1], r[gMean + 2], p.w);
  }
  g_store(gauss + (size_t)kGF4 * i, r);
}

// kept points per 1024-point tile: collapsed < 0 or collapsed == own index
extern "C" __global__ __launch_bounds__(kThreads) void k_keep_count(int n, const int *__restrict__ collapsed,
                                                                   int *__restrict__ tile_count) {
  const int base = blockIdx.x * kTile;
  int cnt = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int i = base + k * kThreads + threadIdx.x;
    if (i < n) {
      const int c = collapsed[i];
      cnt += c < 0 || c == i;
    }
  }
  __shared__ int ws[kThreads / 64];
  cnt = wave_sum_i(cnt);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) tile_count[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// merger.cpp:74-95: order-preserving compaction into new arrays. Each thread owns four
// consecutive points so one scan per block numbers them in order.
extern "C" __global__ __launch_bounds__(kThreads) void k_keep_scatter(
    int n, const int *__restrict__ collapsed, const int *__restrict__ tile_off, const float4 *__restrict__ pts,
    const float4 *__restrict__ nrm, const float4 *__restrict__ omg, const float4 *__restrict__ stats,
    const float4 *__restrict__ gauss, float4 *__restrict__ opts, float4 *__restrict__ onrm,
    float4 *__restrict__ oomg, float4 *__restrict__ ostats, float4 *__restrict__ ogauss) {
  const int i0 = blockIdx.x * kTile + threadIdx.x * 4;
  int keep[4], cnt = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int i = i0 + k;
    keep[k] = 0;
    if (i < n) {
      const int c = collapsed[i];
      keep[k] = c < 0 || c == i;
    }
    cnt += keep[k];
  }
  __shared__ int wsum[kThreads / 64];
  const int incl = wave_incl_scan(cnt);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wave; ++w) before += wsum[w];
  int o = tile_off[blockIdx.x] + before + incl - cnt;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (!keep[k]) continue;
    const size_t i = (size_t)(i0 + k);
    opts[o] = pts[i];
    onrm[o] = nrm[i];
#pragma unroll
    for (int e = 0; e < 3; ++e) oomg[3 * (size_t)o + e] = omg[3 * i + e];
    if (ostats)
#pragma unroll
      for (int e = 0; e < 4; ++e) ostats[4 * (size_t)o + e] = stats[4 * i + e];
    if (ogauss)
#pragma unroll
      for (int e = 0; e < kGF4; ++e) ogauss[(size_t)kGF4 * o + e] = gauss[(size_t)kGF4 * i + e];
    ++o;
  }
}

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------
namespace {

// point-only storage of a cloud (after add / merge): points, normals, Omega, optional Stats and
// Gaussians, the device point count
struct PointStore {
  void *mem = nullptr;
  float4 *pts = nullptr, *nrm = nullptr, *omg = nullptr, *stats = nullptr, *gauss = nullptr;
  int *npoints = nullptr;
};

int alloc_points(size_t cap, bool with_stats, bool with_gauss, PointStore &ps) {
  cap = std::max<size_t>(cap, 1);
  Carver cv;
  const size_t o_pts = cv.take<float4>(cap), o_nrm = cv.take<float4>(cap), o_omg = cv.take<float4>(cap * 3);
  const size_t o_st = with_stats ? cv.take<float4>(cap * 4) : 0;
  const size_t o_g = with_gauss ? cv.take<float4>(cap * kGF4) : 0;
  const size_t o_np = cv.take<int>(1);
  void *mem = nullptr;
  hipError_t e = hipMalloc(&mem, cv.off);
  if (e != hipSuccess) return set_err(NICP_ERR_OOM, "hipMalloc(%zu): %s", cv.off, hipGetErrorString(e));
  char *b = (char *)mem;
  ps.mem = mem;
  ps.pts = (float4 *)(b + o_pts);
  ps.nrm = (float4 *)(b + o_nrm);
  ps.omg = (float4 *)(b + o_omg);
  ps.stats = with_stats ? (float4 *)(b + o_st) : nullptr;
  ps.gauss = with_gauss ? (float4 *)(b + o_g) : nullptr;
  ps.npoints = (int *)(b + o_np);
  return NICP_OK;
}

// makes `ps` the cloud's storage holding n points: frees what the cloud owned and drops the
// index / depth images of a converted cloud
int adopt_store(nicp_cloud *cl, PointStore &ps, int n, int cap) {
  cl->fd.omega_closed = 0;  // the new arrays' information matrices are added or merged ones
  hipStream_t s = cl->ctx->stream;
  HIPCHK(hipMemcpyAsync(ps.npoints, &n, sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  if (cl->alloc) (void)hipFree(cl->alloc);
  if (cl->gauss_alloc) (void)hipFree(cl->gauss_alloc);
  cl->alloc = ps.mem;
  cl->alloc_bytes = 0;  // not from the pool's sizes: freed, not pooled, on destroy
  cl->gauss_alloc = nullptr;
  cl->cap = std::max(cap, 1);
  cl->n = n;
  FrameDev &f = cl->fd;
  f.pts = ps.pts;
  f.nrm = ps.nrm;
  f.omg = ps.omg;
  f.stats = ps.stats;
  f.npoints = ps.npoints;
  f.depth = nullptr;
  f.depthf = nullptr;
  f.index = nullptr;
  f.ii = nullptr;
  f.tile_count = nullptr;
  f.tile_off = nullptr;
  cl->gauss = ps.gauss;
  cl->has_index = 0;
  cl->has_stats = ps.stats != nullptr;
  return NICP_OK;
}

inline unsigned grid_of(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

int ensure_gauss_storage(nicp_cloud *cl) {
  if (cl->gauss) return NICP_OK;
  void *mem = nullptr;
  const size_t bytes = (size_t)std::max(cl->cap, 1) * kGF4 * sizeof(float4);
  hipError_t e = hipMalloc(&mem, bytes);
  if (e != hipSuccess) return set_err(NICP_ERR_OOM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  cl->gauss_alloc = mem;
  cl->gauss = (float4 *)mem;
  return NICP_OK;
}

bool is_identity(const float *m) {
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r)
      if (m[c * 4 + r] != (r == c ? 1.0f : 0.0f)) return false;
  return true;
}

int transform_cloud(nicp_cloud *cl, const float T[16]) {
  Xform X;
  memcpy(X.m, T, sizeof(X.m));
  set_row3(X.m);
  if (is_identity(X.m) || cl->n == 0) return NICP_OK;  // cloud.cpp:196
  cl->fd.omega_closed = 0;  // the rotated information matrices are used as stored
  memcpy(X.mz, X.m, sizeof(X.mz));
  for (int k = 0; k < 4; ++k) {
    X.mz[k * 4 + 3] = 0.0f;
    X.mz[12 + k] = 0.0f;
  }
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_cloud_transform, dim3(grid_of(cl->n)), dim3(kThreads), 0, cl->ctx->stream, cl->n, X,
                     cl->fd.pts, cl->fd.nrm, cl->fd.omg, cl->has_stats ? cl->fd.stats : nullptr, cl->gauss);
  HIPCHK(hipGetLastError());
  return NICP_OK;
}

}  // namespace

extern "C" int nicp_cloud_gaussians(nicp_cloud *cl, float baseline, float alpha) {
  if (!cl) return set_err(NICP_ERR_INVALID, "cloud is null");
  if (!cl->has_index || !cl->fd.index || (!cl->fd.depth && !cl->fd.depthf))
    return set_err(NICP_ERR_STATE, "the cloud has no index and depth images (it was not made from depth)");
  HIPCHK(hipSetDevice(cl->ctx->device));
  CHK(ensure_gauss_storage(cl));
  nicp_ctx *ctx = cl->ctx;
  std::vector<FrameDev> frames(1, cl->fd);
  FrameDev *d_f;
  CHK(upload_frames(ctx, frames, &d_f));
  GaussCam gc;
  inverse3(cl->fd.K, gc.iK);
  gc.fB = baseline * cl->fd.K[0];
  gc.alpha = alpha;
  gc.spherical = cl->fd.model == kSpherical;
  const long long P = (long long)cl->fd.rows * cl->fd.cols;
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_gauss_unproject, dim3(grid_of(P)), dim3(kThreads), 0, ctx->stream, d_f, gc, cl->gauss);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return NICP_OK;
}

extern "C" int nicp_cloud_set_gaussians(nicp_cloud *cl, const nicp_gaussian *g) {
  static_assert(sizeof(nicp_gaussian) == kGF4 * sizeof(float4), "nicp_gaussian is 7 float4");
  if (!cl || (cl->n > 0 && !g)) return set_err(NICP_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(cl->ctx->device));
  CHK(ensure_gauss_storage(cl));
  if (cl->n)
    HIPCHK(hipMemcpyAsync(cl->gauss, g, (size_t)cl->n * sizeof(nicp_gaussian), hipMemcpyHostToDevice, cl->ctx->stream));
  HIPCHK(hipStreamSynchronize(cl->ctx->stream));
  return NICP_OK;
}

extern "C" int nicp_cloud_download_gaussians(nicp_cloud *cl, nicp_gaussian *g) {
  if (!cl || (cl->n > 0 && !g)) return set_err(NICP_ERR_INVALID, "null argument");
  if (!cl->gauss) return set_err(NICP_ERR_STATE, "the cloud has no Gaussians");
  HIPCHK(hipSetDevice(cl->ctx->device));
  if (cl->n)
    HIPCHK(hipMemcpyAsync(g, cl->gauss, (size_t)cl->n * sizeof(nicp_gaussian), hipMemcpyDeviceToHost, cl->ctx->stream));
  HIPCHK(hipStreamSynchronize(cl->ctx->stream));
  return NICP_OK;
}

extern "C" int nicp_cloud_transform(nicp_cloud *cl, const float T[16]) {
  if (!cl || !T) return set_err(NICP_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(cl->ctx->device));
  CHK(transform_cloud(cl, T));
  HIPCHK(hipStreamSynchronize(cl->ctx->stream));
  return NICP_OK;
}

extern "C" int nicp_cloud_add(nicp_cloud *dst, nicp_cloud *src, const float T[16]) {
  if (!dst || !src || !T) return set_err(NICP_ERR_INVALID, "null argument");
  if (dst == src) return set_err(NICP_ERR_INVALID, "a cloud cannot be added to itself");
  if (dst->ctx != src->ctx) return set_err(NICP_ERR_INVALID, "the clouds belong to different contexts");
  if (dst->n > 0 && dst->gauss && !src->gauss)
    return set_err(NICP_ERR_STATE, "the destination has Gaussians and the added cloud has none");
  nicp_ctx *ctx = dst->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  CHK(transform_cloud(src, T));  // cloud.cpp:164
  const int n0 = dst->n, n1 = src->n;
  const long long total = (long long)n0 + n1;
  if (total > (1ll << 30)) return set_err(NICP_ERR_INVALID, "cloud too large");
  const bool with_stats = n0 == 0 ? src->has_stats != 0 : (dst->has_stats && src->has_stats);
  const bool with_gauss = n0 == 0 ? src->gauss != nullptr : (dst->gauss && src->gauss);
  // grow by half again so that a map built by repeated adds copies O(n) points per add
  const int cap = (int)std::max<long long>(total, std::min<long long>((long long)n0 + n0 / 2, 1ll << 30));
  PointStore ps;
  CHK(alloc_points((size_t)cap, with_stats, with_gauss, ps));
  hipError_t e = hipSuccess;
  auto copy = [&](void *d, const void *sp, size_t bytes) {
    if (e == hipSuccess && bytes) e = hipMemcpyAsync(d, sp, bytes, hipMemcpyDeviceToDevice, s);
  };
  copy(ps.pts, dst->fd.pts, (size_t)n0 * sizeof(float4));
  copy(ps.nrm, dst->fd.nrm, (size_t)n0 * sizeof(float4));
  copy(ps.omg, dst->fd.omg, (size_t)n0 * 3 * sizeof(float4));
  if (with_stats) copy(ps.stats, dst->fd.stats, (size_t)n0 * 4 * sizeof(float4));
  if (with_gauss) copy(ps.gauss, dst->gauss, (size_t)n0 * kGF4 * sizeof(float4));
  copy(ps.pts + n0, src->fd.pts, (size_t)n1 * sizeof(float4));
  copy(ps.nrm + n0, src->fd.nrm, (size_t)n1 * sizeof(float4));
  copy(ps.omg + 3 * (size_t)n0, src->fd.omg, (size_t)n1 * 3 * sizeof(float4));
  if (with_stats) copy(ps.stats + 4 * (size_t)n0, src->fd.stats, (size_t)n1 * 4 * sizeof(float4));
  if (with_gauss) copy(ps.gauss + (size_t)kGF4 * n0, src->gauss, (size_t)n1 * kGF4 * sizeof(float4));
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(s);
    (void)hipFree(ps.mem);
    return set_err(NICP_ERR_HIP, "add: %s", hipGetErrorString(e));
  }
  return adopt_store(dst, ps, (int)total, cap);
}

extern "C" int nicp_merge(nicp_cloud *cl, const float transform[16], const nicp_camera *cam, int rows, int cols,
                          const nicp_merger_params *mp, int *num_suppressed, int *collapsed_out) {
  if (!cl || !transform || !mp) return set_err(NICP_ERR_INVALID, "null argument");
  if (rows <= 0 || cols <= 0 || (long long)rows * cols > (1ll << 28)) return set_err(NICP_ERR_INVALID, "bad image size");
  CHK(check_camera(cam));
  if (cl->n > 0 && !cl->gauss) return set_err(NICP_ERR_STATE, "the cloud has no Gaussians");
  nicp_ctx *ctx = cl->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int n = cl->n;
  if (n == 0) {
    if (num_suppressed) *num_suppressed = 0;
    return NICP_OK;
  }
  cl->fd.omega_closed = 0;  // merging adds information matrices to the kept points
  // z-buffer projection of the cloud from `transform` (merger.cpp:24-27)
  CHK(project_device(ctx, cl, transform, cam, rows, cols));
  // per-point work arrays: collapsed, slot, count, offsets, bucket; tile counts / offsets; total
  const int tiles = (n + kTile - 1) / kTile;
  CHK(ctx->scratch_in.reserve(((size_t)5 * n + 2 * (size_t)tiles + 64) * sizeof(int)));
  int *w = ctx->scratch_in.as<int>();
  int *collapsed = w, *slot = w + n, *count = w + 2 * (size_t)n, *off = w + 3 * (size_t)n, *bucket = w + 4 * (size_t)n;
  int *tile_count = w + 5 * (size_t)n, *tile_off = tile_count + tiles, *total = tile_off + tiles;
  int *kept_dev = total + 16;
  Xform X;
  {
    float T[16];
    memcpy(T, transform, sizeof(T));
    set_row3(T);
    projector_state(cam->model, cam->K, T, X.m);
    memset(X.mz, 0, sizeof(X.mz));
  }
  const SphDev sph = sph_params(cam, rows, cols);
  const int zero = 0;
  HIPCHK(hipMemcpyAsync(total, &zero, sizeof(int), hipMemcpyHostToDevice, s));
  const dim3 tb(kThreads), g(grid_of(n));
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_merge_init, g, tb, 0, s, n, count);
  hipLaunchKernelGGL(k_merge_classify, g, tb, 0, s, n, X, cam->model, sph, cam->min_distance, cam->max_distance, rows, cols,
                     mp->distance_threshold, mp->max_point_depth, (const float4 *)cl->fd.pts,
                     (const int *)ctx->cur_index.as<int>(), (const float *)ctx->depth_tmp.as<float>(), collapsed,
                     count, slot);
  hipLaunchKernelGGL(k_merge_offsets, g, tb, 0, s, n, (const int *)count, off, total);
  hipLaunchKernelGGL(k_merge_bucket, g, tb, 0, s, n, (const int *)collapsed, (const int *)slot, (const int *)off,
                     bucket);
  hipLaunchKernelGGL(k_merge_accumulate, g, tb, 0, s, n, (const int *)collapsed, (const int *)count,
                     (const int *)off, (const int *)bucket, cl->gauss, cl->fd.pts);
  hipLaunchKernelGGL(k_keep_count, dim3(tiles), tb, 0, s, n, (const int *)collapsed, tile_count);
  HIPCHK(hipGetLastError());
  // exclusive scan of the tile counts (the conversion's k_scan over a frame of `tiles` tiles)
  {
    std::vector<FrameDev> fr(1);
    memset(&fr[0], 0, sizeof(FrameDev));
    fr[0].n_tiles = tiles;
    fr[0].tile_count = tile_count;
    fr[0].tile_off = tile_off;
    fr[0].npoints = kept_dev;
    FrameDev *d_f;
    CHK(upload_frames(ctx, fr, &d_f));
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, (const FrameDev *)d_f);
    HIPCHK(hipGetLastError());
  }
  int kept = 0;
  HIPCHK(hipMemcpyAsync(&kept, kept_dev, sizeof(int), hipMemcpyDeviceToHost, s));
  if (collapsed_out) HIPCHK(hipMemcpyAsync(collapsed_out, collapsed, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  PointStore ps;
  const bool with_stats = cl->has_stats && cl->fd.stats;
  CHK(alloc_points((size_t)kept, with_stats, true, ps));
  hipLaunchKernelGGL(k_keep_scatter, dim3(tiles), tb, 0, s, n, (const int *)collapsed, (const int *)tile_off,
                     (const float4 *)cl->fd.pts, (const float4 *)cl->fd.nrm, (const float4 *)cl->fd.omg,
                     (const float4 *)(with_stats ? cl->fd.stats : nullptr), (const float4 *)cl->gauss, ps.pts, ps.nrm,
                     ps.omg, ps.stats, ps.gauss);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(s);
    (void)hipFree(ps.mem);
    return set_err(NICP_ERR_HIP, "k_keep_scatter: %s", hipGetErrorString(e));
  }
  if (num_suppressed) *num_suppressed = n - kept;
  return adopt_store(cl, ps, kept, kept);
}
