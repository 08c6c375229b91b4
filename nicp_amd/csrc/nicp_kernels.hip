// CDNA4 (gfx950) kernels of the NICP registration loop.
//
// Conversion of a depth frame into a cloud (DepthImageConverterIntegralImage::compute,
// nicp/nicp/depthimageconverterintegralimage.cpp:18-64):
//   k_count      u16 depth -> valid-pixel count per 1024-pixel tile
//   k_scan       per frame: exclusive scan of tile counts (row-major point numbering)
//   k_unproject  points + index image (PinholePointProjector::unProject, pinholepointprojector.cpp:93-134)
//   k_integral_x x-prefix of (n, p, p p^T) per image row    (pointintegralimage.cpp:16-35)
//   k_integral_y y-prefix per image column                  (pointintegralimage.cpp:38-43)
//   k_stats      box query, covariance, 3x3 eigensolve, normal, curvature and
//                both information matrices per pixel      (statscalculatorintegralimage.cpp:38-81,
//                                                            informationmatrixcalculator.cpp:9-57)
// One registration iteration (AlignerProjective::align, alignerprojective.cpp:45-101):
//   k_zproject   z-buffer projection with a 64-bit atomicMin on (depth bits << 32 | index)
//                (PinholePointProjector::project, pinholepointprojector.cpp:33-66)
//   k_corr_lin   projective correspondence gates fused with the Linearizer's H/b terms
//                (correspondencefinderprojective.cpp:23-257, linearizer.cpp:21-134):
//                compensated FP32 per thread, FP64 wave/block partials
//   k_solve      FP64 reduction of the partials, H + lambda I, LDLT, SE(3) update
//
// All kernels read frame / pair descriptors from device tables and map each
// block to its frame or pair through a block map, so frames and pairs of
// uneven size share one launch per stage.
#include <float.h>

#include "nicp_kernels.h"
#include "nicp_math.h"
#include "solstats.h"

using namespace nicp_dev;
using namespace nicp_math;

namespace {

// global-address-space view of a device pointer held in a frame table (flat loads would
// also count as LDS operations and wait with them)
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gview(const T *p) {
  return (const __attribute__((address_space(1))) T *)p;
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gmut(T *p) {
  return (__attribute__((address_space(1))) T *)p;
}
// 16-byte global load of a float4 (through a clang vector type: HIP's float4 class has no
// address-space-qualified copy)
typedef float gvec4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 gload4(const float4 *p) {
  const gvec4 v = *gview((const gvec4 *)p);
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gstore4(float4 *p, float4 v) {
  *gmut((gvec4 *)p) = gvec4{v.x, v.y, v.z, v.w};
}
typedef float gvec2 __attribute__((ext_vector_type(2)));
typedef unsigned int gvec2u __attribute__((ext_vector_type(2)));
typedef int gint4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float depth_of(uint16_t raw, float scale) { return raw ? scale * (float)raw : 0.0f; }
__device__ __forceinline__ bool depth_valid(float d, float mn, float mx) { return !(d == 0.0f || d < mn || d > mx); }
// the depth of pixel p of a frame: float input, or uint16 times the scale (imageutils.cpp:85-99)
__device__ __forceinline__ float frame_depth(const FrameDev &fd, int p) {
  return fd.depthf ? *gview(fd.depthf + p) : depth_of(*gview(fd.depth + p), fd.depth_scale);
}
// the depths of N pixels (p[k] < 0: none, depth 0), every load issued before the first is used:
// frame_depth's choice of the input type, taken per pixel, makes each load wait for its value
template <int N>
__device__ __forceinline__ void frame_depths(const FrameDev &fd, const int *p, float *d) {
  const __attribute__((address_space(1))) float *f = gview(fd.depthf);
  if (f) {
#pragma unroll
    for (int k = 0; k < N; ++k) d[k] = p[k] >= 0 ? f[p[k]] : 0.0f;
  } else {
    const __attribute__((address_space(1))) uint16_t *u = gview(fd.depth);
    const float sc = fd.depth_scale;
    uint16_t raw[N];
#pragma unroll
    for (int k = 0; k < N; ++k) raw[k] = p[k] >= 0 ? u[p[k]] : (uint16_t)0;
#pragma unroll
    for (int k = 0; k < N; ++k) d[k] = depth_of(raw[k], sc);
  }
}

// PinholePointProjector::_unProject (pinholepointprojector.h:275-280): x = column, y = row
__device__ __forceinline__ float4 unproject_point(const float *iKRt, int c, int r, float d) {
  float o[4];
  mul4_packet(iKRt, (float)c * d, (float)r * d, d, 1.0f, o);
  return make_float4(o[0], o[1], o[2], o[3]);
}

// SphericalPointProjector::_unProject (sphericalpointprojector.h:275-289): the depth range and
// the field of view decide validity; x = sin(theta) d, z = cos(theta) d, y = d tan(phi), then
// the projector transform T (Isometry3f * Vector3f: translation + linear * v). The trigonometric
// functions are evaluated in double and rounded to float: the correctly rounded value of the
// float functions the reference calls.
__device__ __forceinline__ bool sph_valid(const SphDev &s, int c, int r, float d, float mn, float mx) {
  if (d < mn || d > mx) return false;
  const float theta = s.ihres * ((float)c - s.hc);
  const float phi = s.ivres * ((float)r - s.vc);
  return !(fabsf(theta) > s.hfov) && !(fabsf(phi) > s.vfov);
}
__device__ __forceinline__ float4 sph_point(const SphDev &s, const float *T, int c, int r, float d) {
  const float theta = s.ihres * ((float)c - s.hc);
  const float phi = s.ivres * ((float)r - s.vc);
  const float x = (float)sin((double)theta) * d;
  const float z = (float)cos((double)theta) * d;
  const float y = d * (float)tan((double)phi);
  float o[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) o[i] = mc(T, i, 3) + red3(mc(T, i, 0) * x, mc(T, i, 1) * y, mc(T, i, 2) * z);
  return make_float4(o[0], o[1], o[2], 1.0f);
}
// SphericalPointProjector::_project (sphericalpointprojector.h:252-265) with S = _iT: P3 product,
// azimuth and elevation, the cylinder distance as depth; the reference tests fabs(theta > fov),
// so angles below -fov pass; pixels are truncated, not rounded.
__device__ __forceinline__ bool sph_project(const SphDev &s, const float *S, float4 p, float mn, float mx, int &x,
                                            int &y, float &d) {
  float cp[3];
  iso_mul3(S, p.x, p.y, p.z, p.w, cp);
  const float theta = (float)atan2((double)cp[0], (double)cp[2]);
  d = sqrtf(cp[0] * cp[0] + cp[2] * cp[2]);
  const float phi = (float)atan2((double)cp[1], (double)d);
  if (d > mx || d < mn) return false;
  if (theta > s.hfov) return false;
  if (phi > s.vfov) return false;
  x = (int)(s.hres * theta + s.hc);
  y = (int)(s.vres * phi + s.vc);
  return true;
}

__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// inclusive scan within a wavefront
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// Conversion
// ---------------------------------------------------------------------------------------

template <int M>
__device__ __forceinline__ bool pixel_valid(const FrameDev &fd, int p, float d) {
  if (M == kSpherical) {
    const int r = p / fd.cols;
    return sph_valid(fd.sph, p - r * fd.cols, r, d, fd.min_d, fd.max_d);
  }
  return depth_valid(d, fd.min_d, fd.max_d);
}

// Block g counts the valid pixels of the kCountTiles tiles kCountTiles g + j (each tile's frame
// from the tile map), all their depth loads in flight before the counts.
template <int M>
__device__ __forceinline__ void count_body(const FrameDev *__restrict__ F, const int *__restrict__ blk_frame,
                                           int total_tiles) {
  // frame_depths' loads for all tiles first (uint16 bits, or float bits for float depth), then the
  // conversions, so that no tile's loads wait for another's
  int px[kCountTiles][4];
  unsigned int raw[kCountTiles][4];
  float dv[kCountTiles][4];
#pragma unroll
  for (int j = 0; j < kCountTiles; ++j) {
    const int g = kCountTiles * blockIdx.x + j;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      px[j][k] = -1;
      raw[j][k] = 0u;
    }
    if (g >= total_tiles) continue;
    const FrameDev &fd = F[blk_frame[g]];
    const int P = fd.rows * fd.cols;
    const int base = (g - fd.tile_base) * kTile;
    const __attribute__((address_space(1))) float *df = gview(fd.depthf);
    const __attribute__((address_space(1))) uint16_t *du = gview(fd.depth);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int p = base + k * kThreads + threadIdx.x;
      px[j][k] = p < P ? p : -1;
      if (p < P) raw[j][k] = df ? __float_as_uint(df[p]) : (unsigned int)du[p];
    }
  }
#pragma unroll
  for (int j = 0; j < kCountTiles; ++j) {
    const int g = kCountTiles * blockIdx.x + j;
    if (g >= total_tiles) continue;
    const FrameDev &fd = F[blk_frame[g]];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      dv[j][k] = fd.depthf ? __uint_as_float(raw[j][k]) : depth_of((uint16_t)raw[j][k], fd.depth_scale);
  }
  __shared__ int ws[kCountTiles][kThreads / 64];
#pragma unroll
  for (int j = 0; j < kCountTiles; ++j) {
    const int g = kCountTiles * blockIdx.x + j;
    int cnt = 0;
    if (g < total_tiles) {
      const FrameDev &fd = F[blk_frame[g]];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (px[j][k] >= 0) cnt += pixel_valid<M>(fd, px[j][k], dv[j][k]);
    }
    cnt = wave_sum_i(cnt);
    if ((threadIdx.x & 63) == 0) ws[j][threadIdx.x >> 6] = cnt;
  }
  __syncthreads();
  if (threadIdx.x < kCountTiles) {
    const int g = kCountTiles * blockIdx.x + threadIdx.x;
    if (g < total_tiles) {
      const FrameDev &fd = F[blk_frame[g]];
      *gmut(fd.tile_count + (g - fd.tile_base)) =
          ws[threadIdx.x][0] + ws[threadIdx.x][1] + ws[threadIdx.x][2] + ws[threadIdx.x][3];
    }
  }
}
extern "C" __global__ __launch_bounds__(kThreads) void k_count(const FrameDev *__restrict__ F,
                                                              const int *__restrict__ blk_frame, int total_tiles) {
  count_body<kPinhole>(F, blk_frame, total_tiles);
}
extern "C" __global__ __launch_bounds__(kThreads) void k_count_sph(const FrameDev *__restrict__ F,
                                                                  const int *__restrict__ blk_frame,
                                                                  int total_tiles) {
  count_body<kSpherical>(F, blk_frame, total_tiles);
}

// one block of 1024 threads per frame
extern "C" __global__ __launch_bounds__(1024) void k_scan(const FrameDev *__restrict__ F) {
  const FrameDev &fd = F[blockIdx.x];
  const int n = fd.n_tiles;
  const int chunk = (n + 1023) / 1024;
  const int start = threadIdx.x * chunk;
  int local = 0;
  for (int i = 0; i < chunk; ++i)
    if (start + i < n) local += fd.tile_count[start + i];
  __shared__ int wsum[16];
  int incl = wave_incl_scan(local);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < 16; ++w) { int t = wsum[w]; wsum[w] = s; s += t; }
    *fd.npoints = s;
  }
  __syncthreads();
  int run = wsum[wave] + incl - local;
  for (int i = 0; i < chunk; ++i)
    if (start + i < n) {
      int c = fd.tile_count[start + i];
      fd.tile_off[start + i] = run;
      run += c;
    }
}

// Compaction of the valid pixels of a tile into the cloud, numbered in raster order.
template <int M>
__device__ __forceinline__ void unproject_body(const FrameDev *__restrict__ F, const int *__restrict__ blk_frame) {
  const int g = blockIdx.x;
  const FrameDev &fd = F[blk_frame[g]];
  const int P = fd.rows * fd.cols;
  const int t = g - fd.tile_base;
  const int base = t * kTile;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // pixel base + k * kThreads + threadIdx.x: each load and store instruction of a wavefront
  // covers 64 consecutive pixels, and the valid ones take consecutive cloud indices
  float dv[4];
  int px[4];
  unsigned long long bal[4];
  __shared__ int wsum[4][kThreads / 64];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int p = base + k * kThreads + threadIdx.x;
    px[k] = p < P ? p : -1;
  }
  frame_depths<4>(fd, px, dv);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int p = base + k * kThreads + threadIdx.x;
    bal[k] = __ballot(p < P && pixel_valid<M>(fd, p, dv[k]));
    if (lane == 0) wsum[k][wave] = __popcll(bal[k]);
  }
  __syncthreads();
  // the cloud index of a valid pixel is its rank in raster order: the tile's offset, the
  // valid pixels of the earlier rounds k, of the earlier wavefronts of round k, and of the
  // lower lanes of its own wavefront
  const unsigned long long below = lane ? ~0ull >> (64 - lane) : 0ull;
  int before = *gview(fd.tile_off + t);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int p = base + k * kThreads + threadIdx.x;
    int at = before, round = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) {
      if (w < wave) at += wsum[k][w];
      round += wsum[k][w];
    }
    before += round;
    if (p >= P) continue;
    if ((bal[k] >> lane) & 1ull) {
      const int idx = at + __popcll(bal[k] & below);
      const int r = p / fd.cols, c = p - r * fd.cols;
      gstore4(fd.pts + idx, M == kSpherical ? sph_point(fd.sph, fd.iKRt, c, r, dv[k]) : unproject_point(fd.iKRt, c, r, dv[k]));
      *gmut(fd.index + p) = idx;
    } else {
      *gmut(fd.index + p) = -1;
    }
  }
}
extern "C" __global__ __launch_bounds__(kThreads) void k_unproject(const FrameDev *__restrict__ F,
                                                                  const int *__restrict__ blk_frame) {
  unproject_body<kPinhole>(F, blk_frame);
}
extern "C" __global__ __launch_bounds__(kThreads) void k_unproject_sph(const FrameDev *__restrict__ F,
                                                                      const int *__restrict__ blk_frame) {
  unproject_body<kSpherical>(F, blk_frame);
}

// The depth range of the valid pixels of every kSTH x kSTW source tile of a pinhole frame (for
// the tile-binned projection): one wavefront per tile, lane l covering its pixels l, l + 64, ...
// in row-major order; an empty tile stores (+inf, -inf). Where the frame has packed source-tile
// pixels (FrameDev::spix), the same wavefront writes its tile's valid pixels (cloud index >= 0 in
// the index image), compacted in row-major order: depth, cloud index and offset in the tile.
// Block t of a frame (its 1024-pixel tile map) takes source tiles 4 t + wave, 4 t + wave +
// 4 n_tiles, ...
extern "C" __global__ __launch_bounds__(kThreads) void k_src_tiles(const FrameDev *__restrict__ F,
                                                                  const int *__restrict__ blk_frame) {
  const int g = blockIdx.x;
  const FrameDev &fd = F[blk_frame[g]];
  if (!fd.stile_d) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rows = fd.rows, cols = fd.cols;
  const float mind = fd.min_d, maxd = fd.max_d;
  const int sx = stiles_x(cols), ns = stiles(rows, cols);
  for (int s = 4 * (g - fd.tile_base) + wave; s < ns; s += 4 * fd.n_tiles) {
    const int ty = s / sx, tx = s - ty * sx;
    float lo = INFINITY, hi = -INFINITY;
    constexpr int kN = (kSTH * kSTW + 63) / 64;
    int px[kN];
    float d[kN];
#pragma unroll
    for (int i = 0; i < kN; ++i) {
      const int w = i * 64 + lane;
      const int r = ty * kSTH + w / kSTW, c = tx * kSTW + w % kSTW;
      px[i] = w < kSTH * kSTW && r < rows && c < cols ? r * cols + c : -1;
    }
    int idx[kN];
    if (fd.spix) {
#pragma unroll
      for (int i = 0; i < kN; ++i) idx[i] = px[i] >= 0 ? *gview(fd.index + px[i]) : -1;
    }
    frame_depths<kN>(fd, px, d);
#pragma unroll
    for (int i = 0; i < kN; ++i)
      if (px[i] >= 0 && depth_valid(d[i], mind, maxd)) {
        lo = fminf(lo, d[i]);
        hi = fmaxf(hi, d[i]);
      }
    if (fd.spix) {
      const unsigned long long below = lane ? ~0ull >> (64 - lane) : 0ull;
      uint2 *out = fd.spix + (size_t)s * (kSTH * kSTW);
      int at = 0;
#pragma unroll
      for (int i = 0; i < kN; ++i) {
        const unsigned long long m = __ballot(idx[i] >= 0);
        if (idx[i] >= 0)
          *gmut((gvec2u *)(out + at + __popcll(m & below))) =
              gvec2u{__float_as_uint(d[i]), (unsigned int)idx[i] | ((unsigned int)(i * 64 + lane) << kSPosShift)};
        at += __popcll(m);
      }
      if (lane == 0) *gmut(fd.stile_n + s) = at;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      lo = fminf(lo, __shfl_xor(lo, o, 64));
      hi = fmaxf(hi, __shfl_xor(hi, o, 64));
    }
    if (lane == 0) *gmut((gvec2 *)(fd.stile_d + s)) = gvec2{lo, hi};
  }
}

// x-prefix: one thread per (row, channel), sequential along the row as the
// reference's pass over image columns. Channel ch = p[i] * p[j] with p = (x, y, z, 1).
__constant__ int c_ch_i[10] = {3, 0, 1, 2, 0, 0, 0, 1, 1, 2};
__constant__ int c_ch_j[10] = {3, 3, 3, 3, 0, 1, 2, 1, 2, 2};

extern "C" __global__ __launch_bounds__(kThreads) void k_integral_x(const FrameDev *__restrict__ F,
                                                                   const int *__restrict__ xblk_frame) {
  const int g = blockIdx.x;
  const FrameDev &fd = F[xblk_frame[g]];
  const int lt = (g - fd.xblk_base) * kThreads + threadIdx.x;
  const int rows = fd.rows, cols = fd.cols;
  if (lt >= rows * 10) return;
  const int r = lt / 10, ch = lt - r * 10;
  const int ci = c_ch_i[ch], cj = c_ch_j[ch];
  const size_t rowoff = (size_t)r * cols;
  float *orow = fd.ii + rowoff * 10 + ch;
  float iK[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) iK[k] = fd.iKRt[k];
  const float scale = fd.depth_scale, mn = fd.min_d, mx = fd.max_d;
  const uint16_t *drow = fd.depth ? fd.depth + rowoff : nullptr;
  const float *frow = fd.depthf ? fd.depthf + rowoff : nullptr;
  float acc = 0.0f;
  int c = 0;
  for (; c + 8 <= cols; c += 8) {
    float dd[8];
    if (frow) {
#pragma unroll
      for (int k = 0; k < 8; ++k) dd[k] = frow[c + k];
    } else {
#pragma unroll
      for (int k = 0; k < 8; ++k) dd[k] = depth_of(drow[c + k], scale);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float d = dd[k];
      float v = 0.0f;
      if (depth_valid(d, mn, mx)) {
        float o[4];
        mul4_packet(iK, (float)(c + k) * d, (float)r * d, d, 1.0f, o);
        float a = ci == 0 ? o[0] : ci == 1 ? o[1] : ci == 2 ? o[2] : o[3];
        float b = cj == 0 ? o[0] : cj == 1 ? o[1] : cj == 2 ? o[2] : o[3];
        v = a * b;
      }
      acc = v + acc;
      orow[(size_t)(c + k) * 10] = acc;
    }
  }
  for (; c < cols; ++c) {
    const float d = frow ? frow[c] : depth_of(drow[c], scale);
    float v = 0.0f;
    if (depth_valid(d, mn, mx)) {
      float o[4];
      mul4_packet(iK, (float)c * d, (float)r * d, d, 1.0f, o);
      float a = ci == 0 ? o[0] : ci == 1 ? o[1] : ci == 2 ? o[2] : o[3];
      float b = cj == 0 ? o[0] : cj == 1 ? o[1] : cj == 2 ? o[2] : o[3];
      v = a * b;
    }
    acc = v + acc;
    orow[(size_t)c * 10] = acc;
  }
}

// x-prefix of a spherical frame: the points are read from the cloud through the index image,
// as PointIntegralImage::compute does (pointintegralimage.cpp:7-44), since recomputing them
// from depth would take three trigonometric functions per channel.
extern "C" __global__ __launch_bounds__(kThreads) void k_integral_x_sph(const FrameDev *__restrict__ F,
                                                                       const int *__restrict__ xblk_frame) {
  const int g = blockIdx.x;
  const FrameDev &fd = F[xblk_frame[g]];
  const int lt = (g - fd.xblk_base) * kThreads + threadIdx.x;
  const int rows = fd.rows, cols = fd.cols;
  if (lt >= rows * 10) return;
  const int r = lt / 10, ch = lt - r * 10;
  const int ci = c_ch_i[ch], cj = c_ch_j[ch];
  const size_t rowoff = (size_t)r * cols;
  float *orow = fd.ii + rowoff * 10 + ch;
  const int *irow = fd.index + rowoff;
  float acc = 0.0f;
  for (int c = 0; c < cols; ++c) {
    const int idx = irow[c];
    float v = 0.0f;
    if (idx >= 0) {
      const float4 q = fd.pts[idx];
      const float a = ci == 0 ? q.x : ci == 1 ? q.y : ci == 2 ? q.z : q.w;
      const float b = cj == 0 ? q.x : cj == 1 ? q.y : cj == 2 ? q.z : q.w;
      v = a * b;
    }
    acc = v + acc;
    orow[(size_t)c * 10] = acc;
  }
}

// y-prefix: one thread per (column, channel) slot; consecutive threads touch
// consecutive floats of a row, so every step is one coalesced row segment.
extern "C" __global__ __launch_bounds__(kThreads) void k_integral_y(const FrameDev *__restrict__ F,
                                                                   const int *__restrict__ yblk_frame) {
  const int g = blockIdx.x;
  const FrameDev &fd = F[yblk_frame[g]];
  const int j = (g - fd.yblk_base) * kThreads + threadIdx.x;
  const int stride = fd.cols * 10;
  if (j >= stride) return;
  float *col = fd.ii + j;
  const int rows = fd.rows;
  float acc = 0.0f;
  int r = 0;
  for (; r + 8 <= rows; r += 8) {
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = col[(size_t)(r + k) * stride];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      acc = v[k] + acc;
      col[(size_t)(r + k) * stride] = acc;
    }
  }
  for (; r < rows; ++r) {
    acc = col[(size_t)r * stride] + acc;
    col[(size_t)r * stride] = acc;
  }
}

// Both prefix passes in one workgroup per frame (for batches of many frames). The frame is
// walked in bands of kIBand rows and, within a band, chunks of kIChunk columns. Waves 0-7
// compute a chunk's x-prefixes exactly as k_integral_x does (one thread per (row, channel),
// sequential along the row, the carry kept across the band's chunks) into one of two LDS
// tiles, while waves 8-15 take the previous chunk's tile and carry its columns down the band
// exactly as k_integral_y does (one thread per (column, channel), sequential down the rows,
// starting from the integral image's row above the band, which the same thread wrote), and
// write every row segment of the integral image once, coalesced. The values are those of the
// two-kernel path, bit for bit; no workgroup waits for another.
constexpr int kIBand = 40, kIChunk = 48;
constexpr int kIStride = kIChunk * 10 + 1;  // LDS row stride (floats): rows fall in different banks
constexpr int kIFrameThreads = 1024;


// rows ci and cj of iKRt: a thread's channel p[ci] * p[cj] of the point p = iKRt (c d, r d, d, 1)
// needs only those two rows of mul4_packet, in its association order
struct XRows {
  float a[4], b[4];
};
__device__ __forceinline__ XRows x_rows(const float *iK, int ci, int cj) {
  XRows x;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    x.a[k] = iK[k * 4 + ci];
    x.b[k] = iK[k * 4 + cj];
  }
  return x;
}
// the channel at (row, col) from depth d (0 when invalid), bit for bit k_integral_x's value
__device__ __forceinline__ float x_channel_depth(const FrameDev &fd, const XRows &x, int row, int col, float d) {
  const float v0 = (float)col * d, v1 = (float)row * d;
  float a = x.a[0] * v0;
  a = x.a[1] * v1 + a;
  a = x.a[2] * d + a;
  a = x.a[3] * 1.0f + a;
  float b = x.b[0] * v0;
  b = x.b[1] * v1 + b;
  b = x.b[2] * d + b;
  b = x.b[3] * 1.0f + b;
  return depth_valid(d, fd.min_d, fd.max_d) ? a * b : 0.0f;
}
// component i of p by two levels of selects on the bits of i (a chain of compares is turned
// into an indexed stack copy)
__device__ __forceinline__ float sel4(int i, const float4 &p) {
  const bool odd = (i & 1) != 0, high = (i & 2) != 0;
  const float lo = odd ? p.y : p.x, hi = odd ? p.w : p.z;
  return high ? hi : lo;
}

// eight consecutive channel values of a row, the loads first
template <int M>
__device__ __forceinline__ void x_channels8(const FrameDev &fd, const XRows &x, int row, int col, int ci, int cj,
                                            float v[8]) {
  const size_t q = (size_t)row * fd.cols + col;
  if (M == kSpherical) {
    int idx[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) idx[u] = gview(fd.index)[q + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const float4 p = fd.pts[idx[u] >= 0 ? idx[u] : 0];
      v[u] = idx[u] >= 0 ? sel4(ci, p) * sel4(cj, p) : 0.0f;
    }
    return;
  }
  float d[8];
  if (fd.depthf) {
#pragma unroll
    for (int u = 0; u < 8; ++u) d[u] = gview(fd.depthf)[q + u];
  } else {
    uint16_t raw[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) raw[u] = gview(fd.depth)[q + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) d[u] = depth_of(raw[u], fd.depth_scale);
  }
#pragma unroll
  for (int u = 0; u < 8; ++u) v[u] = x_channel_depth(fd, x, row, col + u, d[u]);
}
template <int M>
__device__ __forceinline__ float x_channel1(const FrameDev &fd, const XRows &x, int row, int col, int ci, int cj) {
  const size_t q = (size_t)row * fd.cols + col;
  if (M == kSpherical) {
    const int idx = gview(fd.index)[q];
    const float4 p = fd.pts[idx >= 0 ? idx : 0];
    return idx >= 0 ? sel4(ci, p) * sel4(cj, p) : 0.0f;
  }
  const float d = fd.depthf ? gview(fd.depthf)[q] : depth_of(gview(fd.depth)[q], fd.depth_scale);
  return x_channel_depth(fd, x, row, col, d);
}

template <int M>
__device__ __forceinline__ void integral_frame_body(const FrameDev *__restrict__ F, const int *__restrict__ order) {
  __shared__ float tile[2][kIBand * kIStride];
  const FrameDev &fd = F[order ? order[blockIdx.x] : blockIdx.x];
  const int rows = fd.rows, cols = fd.cols;
  const int nchunks = (cols + kIChunk - 1) / kIChunk, nbands = (rows + kIBand - 1) / kIBand;
  const int steps = nbands * nchunks;
  const int t = threadIdx.x;
  const bool xside = t < kIFrameThreads / 2;
  // x side: (row in band, channel); y side: (column in chunk, channel) as one slot index
  const int xr = t / 10, xch = t - xr * 10;
  const int ys = t - kIFrameThreads / 2, ycc = ys / 10;
  const bool xact = xside && t < kIBand * 10, yact = !xside && ys < kIChunk * 10;
  const int ci = c_ch_i[xch < 10 ? xch : 0], cj = c_ch_j[xch < 10 ? xch : 0];
  __attribute__((address_space(1))) float *ii = (__attribute__((address_space(1))) float *)fd.ii;
  const XRows xrw = x_rows(fd.iKRt, ci, cj);
  float xacc = 0.0f;
  // the y side's carry for its next chunk, loaded a step ahead (after this step's writes,
  // which hold it when the frame has a single chunk): the row above a band was written
  // nchunks steps earlier and has usually left the cache
  float ycarry = 0.0f;
  // Pinhole frames with uint16 depth and a width that is a multiple of 8 (every row then
  // starts 16-byte aligned): the x side loads the depths of its next chunk's full groups of
  // eight during the current step
  const bool pre = M == kPinhole && !fd.depthf && (cols & 7) == 0;
  const __attribute__((address_space(1))) unsigned int *d16 =
      (const __attribute__((address_space(1))) unsigned int *)fd.depth;
  uint4 dnext[kIChunk / 8];
#pragma unroll
  for (int j = 0; j < kIChunk / 8; ++j) dnext[j] = make_uint4(0u, 0u, 0u, 0u);
  auto prefetch = [&](int st2) {
    const int b2 = st2 / nchunks, k2 = st2 - b2 * nchunks;
    const int row2 = b2 * kIBand + xr, c2 = k2 * kIChunk;
#pragma unroll
    for (int j = 0; j < kIChunk / 8; ++j)
      if (row2 < rows && c2 + 8 * j + 8 <= cols) {
        const size_t w = ((size_t)row2 * cols + c2) / 2 + 4 * j;  // 32-bit words, 16-byte aligned
        dnext[j] = make_uint4(d16[w], d16[w + 1], d16[w + 2], d16[w + 3]);
      }
  };
  if (pre && xact && steps > 0) prefetch(0);
  for (int st = 0; st <= steps; ++st) {
    if (xact && st < steps) {
      const int b = st / nchunks, k = st - b * nchunks;
      const int row = b * kIBand + xr, c0 = k * kIChunk;
      if (k == 0) xacc = 0.0f;
      float *out = &tile[st & 1][xr * kIStride + xch];
      uint4 dcur[kIChunk / 8];
      if constexpr (M == kPinhole) {
#pragma unroll
        for (int j = 0; j < kIChunk / 8; ++j) dcur[j] = dnext[j];
        if (pre && st + 1 < steps) prefetch(st + 1);
      }
      if (row < rows) {
        const int cn = min(kIChunk, cols - c0);
        int cc = 0;
        if (M == kPinhole && pre) {
#pragma unroll
          for (int j = 0; j < kIChunk / 8; ++j) {
            if (8 * j + 8 <= cn) {
              const unsigned int w[4] = {dcur[j].x, dcur[j].y, dcur[j].z, dcur[j].w};
#pragma unroll
              for (int u = 0; u < 8; ++u) {
                const float d = depth_of((uint16_t)(w[u >> 1] >> (16 * (u & 1))), fd.depth_scale);
                xacc = x_channel_depth(fd, xrw, row, c0 + 8 * j + u, d) + xacc;
                out[(8 * j + u) * 10] = xacc;
              }
              cc = 8 * j + 8;
            }
          }
        }
        // eight columns' loads in flight at a time; the prefix itself stays sequential
        for (; cc + 8 <= cn; cc += 8) {
          float v[8];
          x_channels8<M>(fd, xrw, row, c0 + cc, ci, cj, v);
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            xacc = v[u] + xacc;
            out[(cc + u) * 10] = xacc;
          }
        }
        for (; cc < cn; ++cc) {
          xacc = x_channel1<M>(fd, xrw, row, c0 + cc, ci, cj) + xacc;
          out[cc * 10] = xacc;
        }
      }
    }
    if (yact && st >= 1) {
      const int sp = st - 1, b = sp / nchunks, k = sp - b * nchunks;
      const int col = k * kIChunk + ycc, row0 = b * kIBand;
      if (col < cols) {
        // slot ys of the chunk is column c0 + ycc, channel ys % 10: consecutive threads write
        // consecutive floats of the row segment
        const size_t rowstride = (size_t)cols * 10;
        const size_t o = (size_t)row0 * rowstride + (size_t)k * kIChunk * 10 + ys;
        float acc = ycarry;
        const float *src = &tile[sp & 1][ys];
        const int rn = min(kIBand, rows - row0);
        int r = 0;
        for (; r + 8 <= rn; r += 8) {
          float v[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) v[u] = src[(r + u) * kIStride];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            acc = v[u] + acc;
            ii[o + (size_t)(r + u) * rowstride] = acc;
          }
        }
        for (; r < rn; ++r) {
          acc = src[r * kIStride] + acc;
          ii[o + (size_t)r * rowstride] = acc;
        }
      }
    }
    if (yact && st < steps) {
      const int b = st / nchunks, k = st - b * nchunks;
      const int col = k * kIChunk + ycc, row0 = b * kIBand;
      ycarry = col < cols && row0 > 0 ? ii[(size_t)(row0 - 1) * cols * 10 + (size_t)k * kIChunk * 10 + ys] : 0.0f;
    }
    __syncthreads();
  }
}
// Depth images of a batch against one packed device array (frame f at pix_base[f], rows*cols
// entries): blockIdx.y strides over the frames (gridDim.y is capped at 65535), and a frame's
// kDepthCopyBlocks blocks stride over its pixels; to_frames copies the packed array into the
// frames' depth, otherwise the frames' depth into the packed array. 16-byte accesses where both
// sides of the frame are 16-byte aligned and its pixel count is a multiple of eight, else 2-byte
// ones.
extern "C" __global__ __launch_bounds__(kThreads) void k_batch_depth_copy(const FrameDev *__restrict__ F,
                                                                         const long long *__restrict__ pix_base,
                                                                         int num_frames, const uint16_t *__restrict__ src,
                                                                         uint16_t *__restrict__ dst, int to_frames) {
  const long long stride = (long long)gridDim.x * kThreads;
  const long long t0 = (long long)blockIdx.x * kThreads + threadIdx.x;
  for (int f = blockIdx.y; f < num_frames; f += gridDim.y) {
    const FrameDev &fd = F[f];
    const long long P = (long long)fd.rows * fd.cols;
    const long long o = pix_base[f];
    uint16_t *depth = const_cast<uint16_t *>(fd.depth);
    const uint16_t *from = to_frames ? src + o : depth;
    uint16_t *to = to_frames ? depth : dst + o;
    if ((P & 7) == 0 && (((uintptr_t)from | (uintptr_t)to) & 15) == 0) {
      const __attribute__((address_space(1))) gint4 *s4 = gview((const gint4 *)from);
      __attribute__((address_space(1))) gint4 *d4 = gmut((gint4 *)to);
      for (long long i = t0; i < (P >> 3); i += stride) d4[i] = s4[i];
    } else {
      const __attribute__((address_space(1))) uint16_t *s2 = gview(from);
      __attribute__((address_space(1))) uint16_t *d2 = gmut(to);
      for (long long i = t0; i < P; i += stride) d2[i] = s2[i];
    }
  }
}
extern "C" __global__ __launch_bounds__(kIFrameThreads) void k_integral_frame(const FrameDev *__restrict__ F,
                                                                              const int *__restrict__ order) {
  integral_frame_body<kPinhole>(F, order);
}
extern "C" __global__ __launch_bounds__(kIFrameThreads) void k_integral_frame_sph(const FrameDev *__restrict__ F,
                                                                                  const int *__restrict__ order) {
  integral_frame_body<kSpherical>(F, order);
}

// ---- Eigen SelfAdjointEigenSolver<Matrix3f>::computeDirect, restated (3x3 FP32) ----
namespace {

__device__ __forceinline__ void cross3(const float *a, const float *b, float *o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ float sqn3(const float *a) { return red3(a[0] * a[0], a[1] * a[1], a[2] * a[2]); }

__device__ __forceinline__ float sel3(int i, float a, float b, float c) { return i == 0 ? a : (i == 1 ? b : c); }

// extract_kernel of Eigen's computeDirect on a shifted matrix given by its three columns
// (named values, so that the column choice stays a register select): the representative
// column i0 = argmax |diagonal| and the better-conditioned of the two cross products with
// the other columns, normalised. a x b == -(b x a) bitwise, so the cross products of the
// three column pairs serve every i0.
__device__ __forceinline__ void eig_extract_kernel(const float *mat, float *res, float *rep_out) {
  const float a0 = mat[0], a1 = mat[1], a2 = mat[2];
  const float b0 = mat[3], b1 = mat[4], b2 = mat[5];
  const float d0 = mat[6], d1 = mat[7], d2 = mat[8];
  int i0 = 0;
  float best = fabsf(a0);
  if (fabsf(b1) > best) { best = fabsf(b1); i0 = 1; }
  if (fabsf(d2) > best) { best = fabsf(d2); i0 = 2; }
  // col0 x col1, col0 x col2, col1 x col2
  const float x01 = a1 * b2 - a2 * b1, y01 = a2 * b0 - a0 * b2, z01 = a0 * b1 - a1 * b0;
  const float x02 = a1 * d2 - a2 * d1, y02 = a2 * d0 - a0 * d2, z02 = a0 * d1 - a1 * d0;
  const float x12 = b1 * d2 - b2 * d1, y12 = b2 * d0 - b0 * d2, z12 = b0 * d1 - b1 * d0;
  // i0 = 0: (c0, c1) = (c01, c02); 1: (c12, -c01); 2: (-c02, -c12)
  const float c0x = sel3(i0, x01, x12, -x02), c0y = sel3(i0, y01, y12, -y02), c0z = sel3(i0, z01, z12, -z02);
  const float c1x = sel3(i0, x02, -x01, -x12), c1y = sel3(i0, y02, -y01, -y12), c1z = sel3(i0, z02, -z01, -z12);
  const float n0 = red3(c0x * c0x, c0y * c0y, c0z * c0z), n1 = red3(c1x * c1x, c1y * c1y, c1z * c1z);
  if (n0 > n1) {
    const float sq = sqrtf(n0);
    res[0] = c0x / sq; res[1] = c0y / sq; res[2] = c0z / sq;
  } else {
    const float sq = sqrtf(n1);
    res[0] = c1x / sq; res[1] = c1y / sq; res[2] = c1z / sq;
  }
  if (rep_out) {
    rep_out[0] = sel3(i0, a0, b0, d0);
    rep_out[1] = sel3(i0, a1, b1, d1);
    rep_out[2] = sel3(i0, a2, b2, d2);
  }
}

// cov: symmetric 3x3 column-major; evals ascending, evecs column-major (col 0 = smallest).
// All = false computes column 0 only (V[0..2], the rest undefined) where it does not depend on
// the others: it is extract_kernel of (m - ev[0] I) in both branches of the reference's choice,
// unless the smallest eigenvalue is the one orthonormalised against the largest one's vector.
template <bool All = true>
__device__ __forceinline__ void eig3_direct(const float *cov, float *ev, float *V) {
  float m[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) m[i] = cov[i];
  const float shift = red3(m[0], m[4], m[8]) / 3.0f;
  m[0] -= shift; m[4] -= shift; m[8] -= shift;
  float scale = 0.0f;
#pragma unroll
  for (int i = 0; i < 9; ++i) scale = fmaxf(scale, fabsf(m[i]));
  if (scale > 0.0f) {
#pragma unroll
    for (int i = 0; i < 9; ++i) m[i] = m[i] / scale;
  }
  // roots of the characteristic polynomial
  {
    const float s_inv3 = 1.0f / 3.0f;
    const float s_sqrt3 = sqrtf(3.0f);
    float m00 = m[0], m11 = m[4], m22 = m[8], m10 = m[1], m20 = m[2], m21 = m[5];
    float c0 = m00 * m11 * m22 + 2.0f * m10 * m20 * m21 - m00 * m21 * m21 - m11 * m20 * m20 - m22 * m10 * m10;
    float c1 = m00 * m11 - m10 * m10 + m00 * m22 - m20 * m20 + m11 * m22 - m21 * m21;
    float c2 = m00 + m11 + m22;
    float c2_over_3 = c2 * s_inv3;
    float a_over_3 = (c2 * c2_over_3 - c1) * s_inv3;
    a_over_3 = fmaxf(a_over_3, 0.0f);
    float half_b = 0.5f * (c0 + c2_over_3 * (2.0f * c2_over_3 * c2_over_3 - c1));
    float q = a_over_3 * a_over_3 * a_over_3 - half_b * half_b;
    q = fmaxf(q, 0.0f);
    float rho = sqrtf(a_over_3);
    float theta = atan2f(sqrtf(q), half_b) * s_inv3;
    float cos_theta = cosf(theta);
    float sin_theta = sinf(theta);
    ev[0] = c2_over_3 - rho * (cos_theta + s_sqrt3 * sin_theta);
    ev[1] = c2_over_3 - rho * (cos_theta - s_sqrt3 * sin_theta);
    ev[2] = c2_over_3 + 2.0f * rho * cos_theta;
  }
  const float eps = FLT_EPSILON;
  if ((ev[2] - ev[0]) <= eps) {
#pragma unroll
    for (int i = 0; i < 9; ++i) V[i] = 0.0f;
    V[0] = V[4] = V[8] = 1.0f;
  } else {
    float tmp[9];
    float d0 = ev[2] - ev[1];
    float d1 = ev[1] - ev[0];
    const bool swapped = d0 > d1;  // k = 2, l = 0 when true (else k = 0, l = 2)
    if (swapped) d0 = d1;
    if (!All && !(swapped && d0 <= 2.0f * eps * d1)) {
#pragma unroll
      for (int i = 0; i < 9; ++i) tmp[i] = m[i];
      tmp[0] -= ev[0]; tmp[4] -= ev[0]; tmp[8] -= ev[0];
      eig_extract_kernel(tmp, V, nullptr);
      ev[0] = ev[0] * scale; ev[1] = ev[1] * scale; ev[2] = ev[2] * scale;
      ev[0] = ev[0] + shift; ev[1] = ev[1] + shift; ev[2] = ev[2] + shift;
      return;
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) tmp[i] = m[i];
    const float ek = swapped ? ev[2] : ev[0];
    tmp[0] -= ek; tmp[4] -= ek; tmp[8] -= ek;
    float vk[3], vl[3];
    eig_extract_kernel(tmp, vk, vl);
    if (d0 <= 2.0f * eps * d1) {
      float d = red3(vk[0] * vl[0], vk[1] * vl[1], vk[2] * vl[2]);
      float t0 = d * vl[0], t1 = d * vl[1], t2 = d * vl[2];
      vl[0] -= t0; vl[1] -= t1; vl[2] -= t2;
      float z = sqn3(vl);
      if (z > 0.0f) {
        float sz = sqrtf(z);
        vl[0] = vl[0] / sz; vl[1] = vl[1] / sz; vl[2] = vl[2] / sz;
      }
    } else {
#pragma unroll
      for (int i = 0; i < 9; ++i) tmp[i] = m[i];
      const float el = swapped ? ev[0] : ev[2];
      tmp[0] -= el; tmp[4] -= el; tmp[8] -= el;
      eig_extract_kernel(tmp, vl, nullptr);
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      V[r] = swapped ? vl[r] : vk[r];
      V[6 + r] = swapped ? vk[r] : vl[r];
    }
    float c[3];
    cross3(V + 6, V + 0, c);
    float z = sqn3(c);
    if (z > 0.0f) {
      float s = sqrtf(z);
      c[0] = c[0] / s; c[1] = c[1] / s; c[2] = c[2] / s;
    }
    V[3] = c[0]; V[4] = c[1]; V[5] = c[2];
  }
  ev[0] = ev[0] * scale; ev[1] = ev[1] * scale; ev[2] = ev[2] * scale;
  ev[0] = ev[0] + shift; ev[1] = ev[1] + shift; ev[2] = ev[2] + shift;
}

// The symmetric 3x3 eigenproblem in FP64 by cyclic Jacobi rotations (a rotation per
// off-diagonal element and sweep, t = sgn(th) / (|th| + sqrt(th^2 + 1)) with
// th = (a_qq - a_pp) / (2 a_pq), until the off-diagonal part vanishes against the diagonal),
// for the near-degenerate pixels of the FP64 option: when the two smallest eigenvalues nearly
// coincide the FP32 eigenvector of the smallest one carries an error of about
// FLT_EPSILON * l2 / (l1 - l0). cov column-major FP32 as the FP32 solve takes it; eigenvalues
// ascending and eigenvectors (column 0 = smallest) rounded to FP32.
__device__ __forceinline__ void eig3_jacobi_f64(const float *cov, float *ev, float *V) {
  double a[3][3], v[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      // the symmetric part (the FP32 matrix is symmetric up to the order of its two products)
      a[i][j] = 0.5 * ((double)cov[j * 3 + i] + (double)cov[i * 3 + j]);
      v[i][j] = i == j ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 12; ++sweep) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    const double dia = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (off <= 1e-34 * dia || off == 0.0) break;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int p = k == 2 ? 1 : 0, q = k == 0 ? 1 : 2;
      const double apq = a[p][q];
      if (apq == 0.0) continue;
      const double th = (a[q][q] - a[p][p]) / (2.0 * apq);
      const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
#pragma unroll
      for (int r = 0; r < 3; ++r) {  // columns p, q of A and V
        const double arp = a[r][p], arq = a[r][q];
        a[r][p] = c * arp - sn * arq;
        a[r][q] = sn * arp + c * arq;
        const double vrp = v[r][p], vrq = v[r][q];
        v[r][p] = c * vrp - sn * vrq;
        v[r][q] = sn * vrp + c * vrq;
      }
#pragma unroll
      for (int r = 0; r < 3; ++r) {  // rows p, q of A
        const double apr = a[p][r], aqr = a[q][r];
        a[p][r] = c * apr - sn * aqr;
        a[q][r] = sn * apr + c * aqr;
      }
      a[p][q] = a[q][p] = 0.0;
    }
  }
  // ascending order: compare-exchanges of positions (0, 1), (1, 2), (0, 1) on the eigenvalues
  // and their vectors, as selects (an index chosen at run time would put the arrays in memory)
  double e[3] = {a[0][0], a[1][1], a[2][2]}, w[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int r = 0; r < 3; ++r) w[k][r] = v[r][k];
  auto exchange = [&](int i, int j) {
    const bool sw = e[j] < e[i];
    const double ei = e[i], ej = e[j];
    e[i] = sw ? ej : ei;
    e[j] = sw ? ei : ej;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const double wi = w[i][r], wj = w[j][r];
      w[i][r] = sw ? wj : wi;
      w[j][r] = sw ? wi : wj;
    }
  };
  exchange(0, 1);
  exchange(1, 2);
  exchange(0, 1);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    ev[k] = (float)e[k];
    const double nrm = sqrt(w[k][0] * w[k][0] + w[k][1] * w[k][1] + w[k][2] * w[k][2]);
#pragma unroll
    for (int r = 0; r < 3; ++r) V[3 * k + r] = (float)(w[k][r] / nrm);
  }
}

// Omega = U diag(dg) U^T in the association order of Eigen's lazy products
// (informationmatrixcalculator.cpp:26-28); returns the upper triangle 00 01 02 11 12 22
__device__ __forceinline__ void omega6(const float *U, const float *dg, float *o) {
  const int rr[6] = {0, 0, 0, 1, 1, 2}, cc[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    int r = rr[e], c = cc[e];
    float a0 = (U[0 + r] * dg[0]) * U[0 + c];
    float a1 = (U[3 + r] * dg[1]) * U[3 + c];
    float a2 = (U[6 + r] * dg[2]) * U[6 + c];
    o[e] = (a0 + a1) + a2;
  }
}

}  // namespace

// Full = false when only the normals and curvatures are stored (no information matrices, no
// Stats): the eigensolve then computes the smallest eigenvalue's vector alone
template <int M, bool Full>
__device__ __forceinline__ void stats_body(const FrameDev *__restrict__ F, const int *__restrict__ blk_frame,
                                           const StatsParamsDev &sp) {
  const int g = blockIdx.x;
  const FrameDev &fd = F[blk_frame[g]];
  const int rows = fd.rows, cols = fd.cols, P = rows * cols;
  const int base = (g - fd.tile_base) * kTile;
  // stage 1, all four pixels at once: index and depth loads, window radius and the row-major
  // offsets of the window's corners, so that the corner reads below wait on one load only
  int idxs[4], offs[4][4], px[4];
  float ds[4];
  bool oks[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int p = base + k * kThreads + threadIdx.x;
    px[k] = p < P ? p : -1;
    idxs[k] = p < P ? *gview(fd.index + p) : -1;
  }
  frame_depths<4>(fd, px, ds);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int p = base + k * kThreads + threadIdx.x;
    const int r = p / cols, c = p - r * cols;
    const float d = ds[k];
    bool ok = idxs[k] >= 0;
    int R0 = -1, R1 = -1;
    if (d < fd.min_d || d > fd.max_d) {
      ok = false;
    } else if (M == kSpherical) {
      // SphericalPointProjector::_projectInterval (sphericalpointprojector.h:291-303); the
      // converter suppresses the image-radius limits for this projector
      // (depthimageconverterintegralimage.cpp:48-52, statscalculatorintegralimage.cpp:45-51)
      const float a = (float)atan2((double)fd.ip0, (double)d);
      R0 = (int)(fd.sph.hres * a);
      R1 = (int)(fd.sph.vres * a);
      if (R0 < 0 || R1 < 0) ok = false;
    } else {
      float inv = 1.0f / d;
      float a = fd.ip0 * inv, b = fd.ip1 * inv;
      int R = (int)fmaxf(a, b);
      if (R < 0 || R < sp.min_r) ok = false;
      if (R > sp.max_r) R = sp.max_r;
      R0 = R1 = R;
    }
    const int xmin = min(max(c - R0 - 1, 0), cols - 1), xmax = min(max(c + R0 - 1, 0), cols - 1);
    const int ymin = min(max(r - R1 - 1, 0), rows - 1), ymax = min(max(r + R1 - 1, 0), rows - 1);
    offs[k][0] = ymax * cols + xmax;  // A
    offs[k][1] = ymin * cols + xmin;  // D
    offs[k][2] = ymax * cols + xmin;  // B
    offs[k][3] = ymin * cols + xmax;  // C
    oks[k] = ok;
  }
  // stage 2: per pixel, the window sums, the eigensolve and the outputs
  for (int k = 0; k < 4; ++k) {
    const int p = base + k * kThreads + threadIdx.x;
    const int idx = idxs[k];
    if (idx < 0) continue;
    const int r = p / cols, c = p - r * cols;
    const float d = ds[k];
    bool ok = oks[k];
    float acc[10];
    int n = 0;
    if (ok) {
      const auto *A = gview((const gvec2 *)(fd.ii + (size_t)offs[k][0] * 10));
      const auto *D = gview((const gvec2 *)(fd.ii + (size_t)offs[k][1] * 10));
      const auto *B = gview((const gvec2 *)(fd.ii + (size_t)offs[k][2] * 10));
      const auto *C = gview((const gvec2 *)(fd.ii + (size_t)offs[k][3] * 10));
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        const gvec2 a = A[q], dd = D[q], bb = B[q], cc = C[q];
        acc[2 * q] = ((a.x + dd.x) - bb.x) - cc.x;
        acc[2 * q + 1] = ((a.y + dd.y) - bb.y) - cc.y;
      }
      n = (int)acc[0];
      if (n < sp.min_points) ok = false;
    }
    float4 nrm = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float op[6] = {0, 0, 0, 0, 0, 0}, on[6] = {0, 0, 0, 0, 0, 0};
    float V[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, ev[3] = {0, 0, 0}, mean[3] = {0, 0, 0};
    if (ok) {
      const float dn = 1.0f / acc[0];
      const float m0 = acc[1] * dn, m1 = acc[2] * dn, m2 = acc[3] * dn;
      float cov[9];
      cov[0] = acc[4] * dn - m0 * m0;
      cov[1] = acc[5] * dn - m1 * m0;
      cov[2] = acc[6] * dn - m2 * m0;
      cov[3] = acc[5] * dn - m0 * m1;
      cov[4] = acc[7] * dn - m1 * m1;
      cov[5] = acc[8] * dn - m2 * m1;
      cov[6] = acc[6] * dn - m0 * m2;
      cov[7] = acc[8] * dn - m1 * m2;
      cov[8] = acc[9] * dn - m2 * m2;
      eig3_direct<Full>(cov, ev, V);
      if (sp.fp64_gap > 0.0f && !(ev[1] - ev[0] >= sp.fp64_gap * fabsf(ev[2]))) eig3_jacobi_f64(cov, ev, V);
      if (ev[0] < 0.0f) ev[0] = 0.0f;
      mean[0] = m0; mean[1] = m1; mean[2] = m2;
      // Stats::curvature (stats.h:93-98): the 1e-9 literal makes the last add and the divide double
      const float curv = (float)((double)ev[0] / ((double)((ev[0] + ev[1]) + ev[2]) + 1e-9));
      float nx = V[0], ny = V[1], nz = V[2];
      if (curv < sp.curvature_threshold) {
        const float4 pt = M == kSpherical ? fd.pts[idx] : unproject_point(fd.iKRt, c, r, d);
        if (dot4(nx, ny, nz, 0.0f, pt.x, pt.y, pt.z, pt.w) > 0.0f) { nx = -nx; ny = -ny; nz = -nz; }
        nrm = make_float4(nx, ny, nz, curv);
        float dp[3], dq[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          dp[i] = curv < sp.pthr ? sp.pflat[i] : sp.pnonflat[i];
          dq[i] = curv < sp.nthr ? sp.nflat[i] : sp.nnonflat[i];
        }
        if (Full && !sp.skip_omega) {
          omega6(V, dp, op);
          omega6(V, dq, on);
        }
      } else {
        nrm = make_float4(0.0f, 0.0f, 0.0f, curv);
      }
    }
    gstore4(fd.nrm + idx, nrm);
    if (Full && !sp.skip_omega) {
      float4 *o = fd.omg + 3 * (size_t)idx;
      gstore4(o, make_float4(op[0], op[1], op[2], op[3]));
      gstore4(o + 1, make_float4(op[4], op[5], on[0], on[1]));
      gstore4(o + 2, make_float4(on[2], on[3], on[4], on[5]));
    }
    if (Full && sp.keep_stats) {
      float4 *s = fd.stats + 4 * (size_t)idx;
      gstore4(s, make_float4(V[0], V[1], V[2], V[3]));
      gstore4(s + 1, make_float4(V[4], V[5], V[6], V[7]));
      gstore4(s + 2, make_float4(V[8], ev[0], ev[1], ev[2]));
      gstore4(s + 3, make_float4(mean[0], mean[1], mean[2], __int_as_float(ok ? n : 0)));
    }
  }
}
extern "C" __global__ __launch_bounds__(kThreads) void k_stats(const FrameDev *__restrict__ F,
                                                              const int *__restrict__ blk_frame, StatsParamsDev sp) {
  if (sp.keep_stats || !sp.skip_omega) stats_body<kPinhole, true>(F, blk_frame, sp);
  else stats_body<kPinhole, false>(F, blk_frame, sp);
}
extern "C" __global__ __launch_bounds__(kThreads) void k_stats_sph(const FrameDev *__restrict__ F,
                                                                  const int *__restrict__ blk_frame, StatsParamsDev sp) {
  if (sp.keep_stats || !sp.skip_omega) stats_body<kSpherical, true>(F, blk_frame, sp);
  else stats_body<kSpherical, false>(F, blk_frame, sp);
}

// ---------------------------------------------------------------------------------------
// Registration
// ---------------------------------------------------------------------------------------

extern "C" __global__ __launch_bounds__(kThreads) void k_zfill(const PairDev *__restrict__ Pp,
                                                              const int *__restrict__ blk_pair) {
  const int g = blockIdx.x;
  const PairDev &pd = Pp[blk_pair[g]];
  const int P = pd.rows * pd.cols;
  if (pd.sub > 1) {
    // a subsampled pair (the start probe) projects into and reads only the cells of every
    // sub-th row and column: one block per tile of that grid clears them
    const int sub = pd.sub, gcols = (pd.cols + sub - 1) / sub, P2 = ((pd.rows + sub - 1) / sub) * gcols;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int q = (g - pd.tile_base) * kTile + k * kThreads + threadIdx.x;
      if (q < P2) {
        const int gr = q / gcols;
        pd.zbuf[gr * sub * pd.cols + (q - gr * gcols) * sub] = kZInit;
      }
    }
    return;
  }
  const int stride = pd.n_tiles * kTile;
  for (int base = (g - pd.tile_base) * kTile; base < P; base += stride) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      int p = base + k * kThreads + threadIdx.x;
      if (p < P) pd.zbuf[p] = kZInit;
    }
  }
}

// PinholePointProjector::_project (pinholepointprojector.h:253-262) of cloud point i (p) under
// KRt: false when the point falls outside the pair's depth range or image; else its pixel
// (x = column, y = row) and z-buffer key (depth bits << 32 | index)
__device__ __forceinline__ bool pinhole_project(const PairDev &pd, const float *KRt, float4 p, int i, int &x, int &y,
                                                unsigned long long &key) {
  float ip[4];
  mul4_packet(KRt, p.x, p.y, p.z, p.w, ip);
  const float d = ip[2];
  if (d < pd.min_d || d > pd.max_d) return false;
  const float inv = 1.0f / d;
  x = (int)roundf(ip[0] * inv);
  y = (int)roundf(ip[1] * inv);
  if (!(x >= 0 && x < pd.cols && y >= 0 && y < pd.rows)) return false;
  key = ((unsigned long long)__float_as_uint(d) << 32) | (unsigned int)i;
  return true;
}

// the z-buffer cell and key of cloud point i (p) under KRt, or cell -1 when the point
// falls outside the pair's depth range or image (or off the sampled cells, sub > 1)
template <int M>
__device__ __forceinline__ void zproject_point(const PairDev &pd, const float *KRt, float4 p, int i, int sub,
                                               int &cell, unsigned long long &key) {
  const int cols = pd.cols, rows = pd.rows;
  if (M == kSpherical) {
    int x, y;
    float d;
    if (sph_project(pd.sph, KRt, p, pd.min_d, pd.max_d, x, y, d) && x >= 0 && x < cols && y >= 0 && y < rows &&
        (sub == 1 || (x % sub == 0 && y % sub == 0))) {
      cell = y * cols + x;
      key = ((unsigned long long)__float_as_uint(d) << 32) | (unsigned int)i;
    }
  } else {
    int x, y;
    unsigned long long k;
    if (pinhole_project(pd, KRt, p, i, x, y, k) && (sub == 1 || (x % sub == 0 && y % sub == 0))) {
      cell = y * cols + x;
      key = k;
    }
  }
}

// the wavefront's atomics for its lanes' (cell, key)
__device__ __forceinline__ void zmerge(const PairDev &pd, int cell, unsigned long long key, int lane, int strategy) {
  // The lanes that hit the cell of the wavefront's first active lane send one atomicMin
  // with their smallest key. Device-scope atomics run one by one on the memory side, so
  // this keeps a pair whose points pile onto a few pixels (a diverging transform) from
  // stalling the whole launch; the z-buffer is the same.
  const unsigned long long active = __ballot(cell >= 0);
  if (active == 0) return;
  const int lead = __ffsll((long long)active) - 1;
  const int lead_cell = __shfl(cell, lead, 64);
  const bool same = cell >= 0 && cell == lead_cell;
  unsigned long long m = same ? key : kZInit;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned int lo = __shfl_xor((unsigned int)m, o, 64), hi = __shfl_xor((unsigned int)(m >> 32), o, 64);
    const unsigned long long other = ((unsigned long long)hi << 32) | lo;
    m = other < m ? other : m;
  }
  // strategy 1 reads the cell with a relaxed atomic load (one untorn 64-bit read)
  if (cell >= 0 && !same) {
    if (strategy == 0 || __hip_atomic_load(gmut(pd.zbuf + cell), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > key)
      __hip_atomic_fetch_min(gmut(pd.zbuf + cell), key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else if (lane == lead) {
    if (strategy == 0 || __hip_atomic_load(gmut(pd.zbuf + cell), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > m)
      __hip_atomic_fetch_min(gmut(pd.zbuf + cell), m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// which = 0: the reference cloud with the pair's current KRt; 1: the current cloud
// strategy 0: every merged atomic is issued; 1: the cell is read first and the atomic skipped
// when the stored key is already no larger (the read may be stale, but only ever larger than
// the cell's current value, so a skip never loses a smaller key). The z-buffer is the same.
// twin > 0 (reference cloud only): pair q + twin has the same frames as pair q, another
// transform and its own z-buffer; each point is read once and projected for both pairs.
template <int M, bool Twin>
__device__ __forceinline__ void zproject_body(const PairDev *__restrict__ Pp, const FrameDev *__restrict__ F,
                                              const int *__restrict__ blk_pair, int which, int strategy, int twin) {
  const int g = blockIdx.x;
  const int q = blk_pair[g];
  const PairDev &pd = Pp[q];
  const FrameDev &fd = F[which ? pd.cur : pd.ref];
  const int n = *gview(fd.npoints);
  const int base = (g - (which ? pd.ctile_base : pd.rtile_base)) * kTile;
  constexpr int J = Twin ? 2 : 1;
  float KRt[J][16];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const float *S = Pp[q + j * twin].state + (which ? kSlotKRtCur : kSlotKRt);
#pragma unroll
    for (int k = 0; k < 16; ++k) KRt[j][k] = *gview(S + k);
  }
  const int lane = threadIdx.x & 63;
  // a subsampled pair (the start probe) reads only the cells of every sub-th row and column:
  // points landing elsewhere are dropped before their atomics
  const int sub = which ? 1 : pd.sub;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int i = base + k * kThreads + threadIdx.x;
    float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (i < n) p = gload4(fd.pts + i);
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const PairDev &pj = Pp[q + j * twin];
      int cell = -1;
      unsigned long long key = kZInit;
      if (i < n) zproject_point<M>(pj, KRt[j], p, i, sub, cell, key);
      zmerge(pj, cell, key, lane, strategy);
    }
  }
}
extern "C" __global__ __launch_bounds__(kThreads) void k_zproject(const PairDev *__restrict__ Pp,
                                                                 const FrameDev *__restrict__ F,
                                                                 const int *__restrict__ blk_pair, int which,
                                                                 int strategy) {
  zproject_body<kPinhole, false>(Pp, F, blk_pair, which, strategy, 0);
}
extern "C" __global__ __launch_bounds__(kThreads) void k_zproject_sph(const PairDev *__restrict__ Pp,
                                                                     const FrameDev *__restrict__ F,
                                                                     const int *__restrict__ blk_pair, int which,
                                                                     int strategy) {
  zproject_body<kSpherical, false>(Pp, F, blk_pair, which, strategy, 0);
}
// the reference clouds of pairs q and q + twin (same frames) in one pass
extern "C" __global__ __launch_bounds__(kThreads) void k_zproject_twin(const PairDev *__restrict__ Pp,
                                                                      const FrameDev *__restrict__ F,
                                                                      const int *__restrict__ blk_pair, int twin) {
  zproject_body<kPinhole, true>(Pp, F, blk_pair, 0, 0, twin);
}

// z-buffer -> index image (and depth image); resets the z-buffer. Pairs that use the
// current frame's own index image are skipped.
extern "C" __global__ __launch_bounds__(kThreads) void k_zresolve(const PairDev *__restrict__ Pp,
                                                                 const int *__restrict__ blk_pair,
                                                                 float *__restrict__ depth_out) {
  const int g = blockIdx.x;
  const PairDev &pd = Pp[blk_pair[g]];
  if (pd.cur_own) return;
  const int P = pd.rows * pd.cols;
  const int base = (g - pd.tile_base) * kTile;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int p = base + k * kThreads + threadIdx.x;
    if (p >= P) break;
    unsigned long long key = pd.zbuf[p];
    pd.cur_index[p] = key == kZInit ? -1 : (int)(unsigned int)(key & 0xffffffffull);
    if (depth_out) depth_out[p] = key == kZInit ? FLT_MAX : __uint_as_float((unsigned int)(key >> 32));
    pd.zbuf[p] = kZInit;
  }
}

namespace {

// Per-thread accumulation of the FP32 per-correspondence terms: a thread sums at most the
// four correspondences of its pixels in FP32 (relative error of a slot at most 3 ulp of
// its terms' magnitudes), and every sum beyond the thread is FP64 (block_partial, then the
// per-pair reduction of k_solve). FP32 slots take 31 registers instead of 62, which lets
// the linearising kernels run three wavefronts per SIMD instead of two.
typedef float Acc2;

// The per-correspondence terms of Linearizer::update (linearizer.cpp:55-112) for
// J_p = [I | S(p)], J_n = [0 | S(n)], S(v) = -2[v]x (bm_se3.h:55-66), p and n the transformed
// reference point and normal. With C = [p]x and D = [n]x:
//   H = [[Wp, -2 Wp C], [., 4 (C^T Wp C + D^T Wn D)]],
//   b = [Wp e_p; -2 (C^T Wp e_p + D^T Wn e_n)].
// The slots hold the terms without their constant factors -2 and 4 (slot_scale), which the
// FP64 block reduction applies: powers of two, so the scaling is exact. The arithmetic is
// contracted into FMAs: the sums only have to meet a tolerance against the FP64 twin, while the
// gates that decide the correspondences keep the reference's uncontracted order.
// Output order: H upper triangle row-major (21), b (6), chi2, inliers, count; slot 30 counts
// the pixels where both index images hold a point.
__device__ __forceinline__ void lin_terms(const float *rp, const float *rn, float4 cp, float4 cn, float4 o0, float4 o1,
                                          float4 o2, const AlignParamsDev &ap, Acc2 *acc) {
  // Omega_p / Omega_n (symmetric)
  const float P00 = o0.x, P01 = o0.y, P02 = o0.z, P11 = o0.w, P12 = o1.x, P22 = o1.y;
  const float ns = ap.gicp ? 0.0f : ap.nscale;
  const float N00 = ns * o1.z, N01 = ns * o1.w, N02 = ns * o2.x, N11 = ns * o2.y, N12 = ns * o2.z, N22 = ns * o2.w;
  const float pe0 = rp[0] - cp.x, pe1 = rp[1] - cp.y, pe2 = rp[2] - cp.z;
  const float ne0 = rn[0] - cn.x, ne1 = rn[1] - cn.y, ne2 = rn[2] - cn.z;
  const float ep0 = fmaf(P02, pe2, fmaf(P01, pe1, P00 * pe0));
  const float ep1 = fmaf(P12, pe2, fmaf(P11, pe1, P01 * pe0));
  const float ep2 = fmaf(P22, pe2, fmaf(P12, pe1, P02 * pe0));
  const float en0 = fmaf(N02, ne2, fmaf(N01, ne1, N00 * ne0));
  const float en1 = fmaf(N12, ne2, fmaf(N11, ne1, N01 * ne0));
  const float en2 = fmaf(N22, ne2, fmaf(N12, ne1, N02 * ne0));
  const float chi = fmaf(ne2, en2, fmaf(ne1, en1, fmaf(ne0, en0, fmaf(pe2, ep2, fmaf(pe1, ep1, pe0 * ep0)))));
  float k = 1.0f;
  acc[29] += 1.0f;
  if (chi > ap.max_chi2) {
    if (ap.robust) k = ap.max_chi2 / chi;
    else return;
  }
  const float px = rp[0], py = rp[1], pz = rp[2];
  const float nx = rn[0], ny = rn[1], nz = rn[2];
  // A = Wp C: column 0 = Wp (0, pz, -py), column 1 = Wp (-pz, 0, px), column 2 = Wp (py, -px, 0)
  const float A00 = fmaf(P01, pz, -P02 * py), A10 = fmaf(P11, pz, -P12 * py), A20 = fmaf(P12, pz, -P22 * py);
  const float A01 = fmaf(P02, px, -P00 * pz), A11 = fmaf(P12, px, -P01 * pz), A21 = fmaf(P22, px, -P02 * pz);
  const float A02 = fmaf(P00, py, -P01 * px), A12 = fmaf(P01, py, -P11 * px), A22 = fmaf(P02, py, -P12 * px);
  const float B00 = fmaf(N01, nz, -N02 * ny), B10 = fmaf(N11, nz, -N12 * ny), B20 = fmaf(N12, nz, -N22 * ny);
  const float B01 = fmaf(N02, nx, -N00 * nz), B11 = fmaf(N12, nx, -N01 * nz), B21 = fmaf(N22, nx, -N02 * nz);
  const float B02 = fmaf(N00, ny, -N01 * nx), B12 = fmaf(N01, ny, -N11 * nx), B22 = fmaf(N02, ny, -N12 * nx);
  // C^T A + [n]x^T B: row i of C^T is column i of C
#define CTA(sx, sy, sz, a0, a1, a2, i) \
  ((i) == 0 ? fmaf(sz, a1, -(sy) * (a2)) : (i) == 1 ? fmaf(sx, a2, -(sz) * (a0)) : fmaf(sy, a0, -(sx) * (a1)))
#define CTA_ADD(sx, sy, sz, a0, a1, a2, i, t)                                    \
  ((i) == 0 ? fmaf(sz, a1, fmaf(-(sy), a2, t)) : (i) == 1 ? fmaf(sx, a2, fmaf(-(sz), a0, t)) \
            : fmaf(sy, a0, fmaf(-(sx), a1, t)))
  const float R00 = CTA_ADD(nx, ny, nz, B00, B10, B20, 0, CTA(px, py, pz, A00, A10, A20, 0));
  const float R01 = CTA_ADD(nx, ny, nz, B01, B11, B21, 0, CTA(px, py, pz, A01, A11, A21, 0));
  const float R02 = CTA_ADD(nx, ny, nz, B02, B12, B22, 0, CTA(px, py, pz, A02, A12, A22, 0));
  const float R11 = CTA_ADD(nx, ny, nz, B01, B11, B21, 1, CTA(px, py, pz, A01, A11, A21, 1));
  const float R12 = CTA_ADD(nx, ny, nz, B02, B12, B22, 1, CTA(px, py, pz, A02, A12, A22, 1));
  const float R22 = CTA_ADD(nx, ny, nz, B02, B12, B22, 2, CTA(px, py, pz, A02, A12, A22, 2));
  // C^T e_p + D^T e_n
  const float br0 = CTA_ADD(nx, ny, nz, en0, en1, en2, 0, CTA(px, py, pz, ep0, ep1, ep2, 0));
  const float br1 = CTA_ADD(nx, ny, nz, en0, en1, en2, 1, CTA(px, py, pz, ep0, ep1, ep2, 1));
  const float br2 = CTA_ADD(nx, ny, nz, en0, en1, en2, 2, CTA(px, py, pz, ep0, ep1, ep2, 2));
#undef CTA
#undef CTA_ADD
  // H upper triangle, row-major over the 6x6: rows 0-2 = [Htt | Htr], rows 3-5 = [. | Hrr]
  acc[0] = fmaf(k, P00, acc[0]); acc[1] = fmaf(k, P01, acc[1]); acc[2] = fmaf(k, P02, acc[2]);
  acc[3] = fmaf(k, A00, acc[3]); acc[4] = fmaf(k, A01, acc[4]); acc[5] = fmaf(k, A02, acc[5]);
  acc[6] = fmaf(k, P11, acc[6]); acc[7] = fmaf(k, P12, acc[7]);
  acc[8] = fmaf(k, A10, acc[8]); acc[9] = fmaf(k, A11, acc[9]); acc[10] = fmaf(k, A12, acc[10]);
  acc[11] = fmaf(k, P22, acc[11]);
  acc[12] = fmaf(k, A20, acc[12]); acc[13] = fmaf(k, A21, acc[13]); acc[14] = fmaf(k, A22, acc[14]);
  acc[15] = fmaf(k, R00, acc[15]); acc[16] = fmaf(k, R01, acc[16]); acc[17] = fmaf(k, R02, acc[17]);
  acc[18] = fmaf(k, R11, acc[18]); acc[19] = fmaf(k, R12, acc[19]);
  acc[20] = fmaf(k, R22, acc[20]);
  acc[21] = fmaf(k, ep0, acc[21]); acc[22] = fmaf(k, ep1, acc[22]); acc[23] = fmaf(k, ep2, acc[23]);
  acc[24] = fmaf(k, br0, acc[24]); acc[25] = fmaf(k, br1, acc[25]); acc[26] = fmaf(k, br2, acc[26]);
  acc[27] = fmaf(k, chi, acc[27]);
  acc[28] += 1.0f;
}

// A frame's closed-form information matrices (FrameDev::omega_closed), held in registers
struct OmegaModel {
  float thr[2], a[4], b[4];
};
__device__ __forceinline__ OmegaModel omega_model(const FrameDev &f) {
  OmegaModel m;
#pragma unroll
  for (int i = 0; i < 2; ++i) m.thr[i] = f.om_thr[i];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    m.a[i] = f.om_a[i];
    m.b[i] = f.om_b[i];
  }
  return m;
}

// The information matrices of a point with normal n (curvature in w) in closed form: b I + (a - b)
// n n^T with the flat pair where the curvature is below the threshold, else the non-flat pair
// (PointInformationMatrixCalculator / NormalInformationMatrixCalculator::compute,
// informationmatrixcalculator.cpp:9-57, with equal second and third diagonal entries); zero where
// the normal is zero, as the conversion stores them. Same layout as the stored omg records.
__device__ __forceinline__ void closed_omega(const OmegaModel &m, float4 n, float4 &o0, float4 &o1, float4 &o2) {
  if (n.x == 0.0f && n.y == 0.0f && n.z == 0.0f) {
    o0 = o1 = o2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  const bool pf = n.w < m.thr[0], nf = n.w < m.thr[1];
  const float pa = pf ? m.a[0] : m.a[1], pb = pf ? m.b[0] : m.b[1];
  const float na = nf ? m.a[2] : m.a[3], nb = nf ? m.b[2] : m.b[3];
  const float kp = pa - pb, kn = na - nb;
  const float px = kp * n.x, py = kp * n.y, pz = kp * n.z;
  const float qx = kn * n.x, qy = kn * n.y, qz = kn * n.z;
  o0 = make_float4(fmaf(px, n.x, pb), px * n.y, px * n.z, fmaf(py, n.y, pb));
  o1 = make_float4(py * n.z, fmaf(pz, n.z, pb), fmaf(qx, n.x, nb), qx * n.y);
  o2 = make_float4(qx * n.z, fmaf(qy, n.y, nb), qy * n.z, fmaf(qz, n.z, nb));
}

// the constant factor of each slot's terms (lin_terms): -2 for H_tr and b_r, 4 for H_rr
__device__ __forceinline__ double slot_scale(int i) {
  return (i == 3 || i == 4 || i == 5 || i == 8 || i == 9 || i == 10 || i == 12 || i == 13 || i == 14 ||
          (i >= 24 && i <= 26))
             ? -2.0
             : (i >= 15 && i <= 20) ? 4.0 : 1.0;
}

// the value of lane ^ O (O = 1, 2, 4, 8: within a DPP row of 16 lanes)
template <int O>
__device__ __forceinline__ int dpp_xor(int x) {
  if (O == 1) return __builtin_amdgcn_update_dpp(x, x, 0xb1, 0xf, 0xf, false);  // quad_perm [1, 0, 3, 2]
  if (O == 2) return __builtin_amdgcn_update_dpp(x, x, 0x4e, 0xf, 0xf, false);  // quad_perm [2, 3, 0, 1]
  if (O == 8) return __builtin_amdgcn_update_dpp(x, x, 0x128, 0xf, 0xf, false);  // row_ror:8
  const int from_below = __builtin_amdgcn_update_dpp(x, x, 0x114, 0xf, 0xf, false);  // row_shr:4, lane - 4
  const int from_above = __builtin_amdgcn_update_dpp(x, x, 0x104, 0xf, 0xf, false);  // row_shl:4, lane + 4
  return (threadIdx.x & 4) ? from_below : from_above;
}
template <int O>
__device__ __forceinline__ double dpp_xor_f64(double x) {
  const unsigned long long b = __double_as_longlong(x);
  const unsigned int lo = (unsigned int)dpp_xor<O>((int)(unsigned int)b);
  const unsigned int hi = (unsigned int)dpp_xor<O>((int)(unsigned int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// halving step for partner offsets 8, 4, 2 (DPP within a row)
template <int O, int H>
__device__ __forceinline__ void dpp_halve(double *v) {
  const bool up = (threadIdx.x & O) != 0;
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const double keep = up ? v[H + i] : v[i], send = up ? v[i] : v[H + i];
    v[i] = keep + dpp_xor_f64<O>(send);
  }
}

// one halving step of the reduce-scatter on FP32 slots: the permlane swap of (x, y) leaves the
// lane its kept slot and its partner's copy, one in each register
template <int O>
__device__ __forceinline__ float swap_add_f32(float x, float y) {
  const auto r = O == 32 ? __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false)
                         : __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// block-wide reduction of the kVals accumulators into one FP64 partial. Within a wavefront a
// reduce-scatter: the kPartial (32) slots are halved five times; at partner offset o a lane
// keeps the half of its slots selected by its lane bit o and adds its partner's copy of it
// (32 exchanges instead of 6 per slot). Lane l then holds slot l >> 1 summed over the half of
// the wavefront with its lane bit 0, and one more exchange with lane l ^ 1 completes it. The
// first two halvings (offsets 32 and 16, permlane swaps) add in FP32, so an FP32 sum covers at
// most the sixteen correspondences of four lanes; the rest (DPP moves, offsets 8 to 1), the sum
// over the wavefronts and each slot's constant factor (slot_scale, a power of two) are FP64. The
// order of the additions is fixed. The wavefronts' sums go through one of two LDS buffers
// (buf): a block that reduces several tiles in turn alternates them, so the next tile's writes
// never meet the previous tile's reads and need no barrier of their own (the barrier inside the
// next call orders a buffer's reads before its reuse two calls later).
__device__ void block_partial(Acc2 *acc, double *out, int buf = 0) {
  static_assert(kPartial == 32, "the reduce-scatter halves 32 slots five times");
  __shared__ double red[2][kThreads / 64][kPartial];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  float f[kPartial];
#pragma unroll
  for (int i = 0; i < kPartial; ++i) f[i] = i < kVals ? acc[i] : 0.0f;
#pragma unroll
  for (int i = 0; i < 16; ++i) f[i] = swap_add_f32<32>(f[i], f[16 + i]);
#pragma unroll
  for (int i = 0; i < 8; ++i) f[i] = swap_add_f32<16>(f[i], f[8 + i]);
  double v[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = (double)f[i];
  dpp_halve<8, 4>(v);
  dpp_halve<4, 2>(v);
  dpp_halve<2, 1>(v);
  v[0] += dpp_xor_f64<1>(v[0]);
  if ((lane & 1) == 0) red[buf][wave][lane >> 1] = v[0];
  __syncthreads();
  if (threadIdx.x < kPartial) {
    double s = 0.0;
    if (threadIdx.x < kVals)
      for (int w = 0; w < kThreads / 64; ++w) s += red[buf][w][threadIdx.x];
    *gmut(out + threadIdx.x) = s * slot_scale(threadIdx.x);
  }
}

// Workgroups are dispatched to the eight XCDs in turn. Block b takes tile
// start(b % 8) + b / 8, where XCD x owns the contiguous range of q + (x < r) tiles starting at
// x q + min(x, r) (q = grid / 8, r = grid % 8): neighbouring tiles of a pair, whose reference
// gathers overlap, then run on the same XCD and share its L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + min(x, r) + (b >> 3);
}

// CorrespondenceFinderProjective gates on already loaded values (correspondencefinderprojective.cpp:62-95).
// Converted: both clouds come from depth conversions, whose normals are unit vectors or exactly
// zero, so a normal's squared norm is zero exactly when its three components are.
template <bool Converted>
__device__ __forceinline__ bool gates_loaded(const float4 &cn, const float4 &rn0, const float4 &cp, const float4 &rp0,
                                             const float *T, const AlignParamsDev &ap, float *rp, float *rn) {
  if (Converted) {
    if (!ap.gicp && ((cn.x == 0.0f && cn.y == 0.0f && cn.z == 0.0f) || (rn0.x == 0.0f && rn0.y == 0.0f && rn0.z == 0.0f)))
      return false;
  } else if (!ap.gicp && (dot4(cn.x, cn.y, cn.z, 0.0f, cn.x, cn.y, cn.z, 0.0f) == 0.0f ||
                          dot4(rn0.x, rn0.y, rn0.z, 0.0f, rn0.x, rn0.y, rn0.z, 0.0f) == 0.0f)) {
    return false;
  }
  iso_mul3(T, rp0.x, rp0.y, rp0.z, rp0.w, rp);
  iso_rot3(T, rn0.x, rn0.y, rn0.z, rn);
  const float dv0 = cp.x - rp[0], dv1 = cp.y - rp[1], dv2 = cp.z - rp[2], dv3 = cp.w - rp0.w;
  if (dot4(dv0, dv1, dv2, dv3, dv0, dv1, dv2, dv3) > ap.sq_thr) return false;
  if (!ap.gicp) {
    if (dot4(cn.x, cn.y, cn.z, 0.0f, rn[0], rn[1], rn[2], 0.0f) < ap.cos_thr) return false;
    float rc = rn0.w, cc = cn.w;
    if (rc < ap.flat) rc = ap.flat;
    if (cc < ap.flat) cc = ap.flat;
    const float ratio = (float)(((double)rc + 1e-5) / ((double)cc + 1e-5));
    if (ratio < ap.min_ratio || ratio > ap.max_ratio) return false;
  }
  return true;
}

// CorrespondenceFinderProjective gates for (ref index ri, cur index ci); T = finder transform
__device__ __forceinline__ bool gates(const FrameDev &fr, const FrameDev &fc, int ri, int ci, const float *T,
                                      const AlignParamsDev &ap, float *rp, float *rn, float4 &cp, float4 &cn) {
  cn = fc.nrm[ci];
  const float4 rn0 = fr.nrm[ri];
  if (!ap.gicp && (dot4(cn.x, cn.y, cn.z, 0.0f, cn.x, cn.y, cn.z, 0.0f) == 0.0f ||
                   dot4(rn0.x, rn0.y, rn0.z, 0.0f, rn0.x, rn0.y, rn0.z, 0.0f) == 0.0f))
    return false;
  cp = fc.pts[ci];
  const float4 rp0 = fr.pts[ri];
  iso_mul3(T, rp0.x, rp0.y, rp0.z, rp0.w, rp);
  iso_rot3(T, rn0.x, rn0.y, rn0.z, rn);
  const float dv0 = cp.x - rp[0], dv1 = cp.y - rp[1], dv2 = cp.z - rp[2], dv3 = cp.w - rp0.w;
  if (dot4(dv0, dv1, dv2, dv3, dv0, dv1, dv2, dv3) > ap.sq_thr) return false;
  if (!ap.gicp) {
    if (dot4(cn.x, cn.y, cn.z, 0.0f, rn[0], rn[1], rn[2], 0.0f) < ap.cos_thr) return false;
    float rc = rn0.w, cc = cn.w;
    if (rc < ap.flat) rc = ap.flat;
    if (cc < ap.flat) cc = ap.flat;
    const float ratio = (float)(((double)rc + 1e-5) / ((double)cc + 1e-5));
    if (ratio < ap.min_ratio || ratio > ap.max_ratio) return false;
  }
  return true;
}

// windowed search of _computeMulti (correspondencefinderprojective.cpp:144-236)
__device__ int window_search(const FrameDev &fr, const FrameDev &fc, const int *cur_index, int rows, int cols, int r,
                             int c, int ri, const float *T, const AlignParamsDev &ap, float *rp, float *rn) {
  const float4 rn0 = fr.nrm[ri];
  if (dot4(rn0.x, rn0.y, rn0.z, 0.0f, rn0.x, rn0.y, rn0.z, 0.0f) == 0.0f) return -1;
  const float4 rp0 = fr.pts[ri];
  iso_mul3(T, rp0.x, rp0.y, rp0.z, rp0.w, rp);
  iso_rot3(T, rn0.x, rn0.y, rn0.z, rn);
  float rc = rn0.w;
  if (rc < ap.flat) rc = ap.flat;
  const int r0 = max(r - ap.row_region, 0), r1 = min(r + ap.row_region + 1, rows);
  const int c0 = max(c - ap.col_region, 0), c1 = min(c + ap.col_region + 1, cols);
  int best = -1;
  float best_nd = 0.0f, best_pd = 1e20f;
  for (int r2 = r0; r2 < r1; ++r2)
    for (int c2 = c0; c2 < c1; ++c2) {
      const int ci = cur_index[r2 * cols + c2];
      if (ci < 0) continue;
      const float4 cn = fc.nrm[ci];
      if (dot4(cn.x, cn.y, cn.z, 0.0f, cn.x, cn.y, cn.z, 0.0f) == 0.0f) continue;
      const float4 cp = fc.pts[ci];
      const float nd = dot4(cn.x, cn.y, cn.z, 0.0f, rn[0], rn[1], rn[2], 0.0f);
      if (nd < ap.cos_thr) continue;
      const float dv0 = cp.x - rp[0], dv1 = cp.y - rp[1], dv2 = cp.z - rp[2], dv3 = cp.w - rp0.w;
      const float pd = dot4(dv0, dv1, dv2, dv3, dv0, dv1, dv2, dv3);
      if (pd > ap.sq_thr) continue;
      float cc = cn.w;
      if (cc < ap.flat) cc = ap.flat;
      const float ratio = (float)(((double)rc + 1e-5) / ((double)cc + 1e-5));
      if (ratio < ap.min_ratio || ratio > ap.max_ratio) continue;
      if (best == -1 || (pd < best_pd && nd > best_nd)) { best = ci; best_nd = nd; best_pd = pd; }
    }
  return best;
}

}  // namespace

// The stored information matrices of frame F[0]'s points from their normals in closed form
// (FrameDev::omega_closed), for a frame converted without storing them; one thread per point
extern "C" __global__ __launch_bounds__(kThreads) void k_omega_fill(const FrameDev *__restrict__ F) {
  const FrameDev &fd = F[0];
  const int n = *gview(fd.npoints);
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  float4 o0, o1, o2;
  closed_omega(omega_model(fd), gload4(fd.nrm + i), o0, o1, o2);
  float4 *o = fd.omg + 3 * (size_t)i;
  gstore4(o, o0);
  gstore4(o + 1, o1);
  gstore4(o + 2, o2);
}

// The correspondences and linearisation terms of one kDTH x kDTW tile t of a pair's (sampled)
// projector image, accumulated into the thread's slots. Thread x of the block takes tile column
// x & 63 and the tile rows (x >> 6) + 4 k: a wavefront reads 64 consecutive pixels of a row. The
// reference index of a pixel comes from the z-buffer key: in global memory (Lds = false; the cell
// is reset for the next projection) or in the block's LDS tile (Lds = true, cell = tile offset).
// Both variants run the same arithmetic in the same order, so they give the same partial sums.
// Closed: both clouds are depth conversions (FrameDev::omega_closed), so the current points'
// information matrices are formed from their normals and the zero-normal gate tests components.
template <bool Lds, bool Closed>
__device__ __forceinline__ void corr_lin_tile(const PairDev &pd, const FrameDev &fr, const FrameDev &fc, const float *T,
                                              const AlignParamsDev &ap, int store_pairs, int t,
                                              const unsigned long long *zl, const OmegaModel &om_model, Acc2 *acc) {
  const int sub = pd.sub;
  const int GR = (pd.rows + sub - 1) / sub, GC = (pd.cols + sub - 1) / sub;
  const int ty = t / pd.dtiles_x, tx = t - ty * pd.dtiles_x;
  const int lc = threadIdx.x & 63, lr = threadIdx.x >> 6;
  const int gc = tx * kDTW + lc;
#pragma unroll 1
  for (int grp = 0; grp < kDTH / 16; ++grp) {
    // all loads of the thread's four pixels of this group are issued before any of them is used
    int ri[4], ci[4], pix[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int lrow = grp * 16 + k * 4 + lr;
      const int gr = ty * kDTH + lrow;
      pix[k] = gr < GR && gc < GC ? (gr * sub) * pd.cols + gc * sub : -1;
      unsigned long long key = kZInit;
      if (pix[k] >= 0) key = Lds ? zl[lrow * kDTW + lc] : *gview(pd.zbuf + pix[k]);
      ri[k] = key == kZInit ? -1 : (int)(unsigned int)(key & 0xffffffffull);
      ci[k] = pix[k] >= 0 ? *gview(pd.cur_index + pix[k]) : -1;
    }
    if (!Lds) {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (pix[k] >= 0) *gmut(pd.zbuf + pix[k]) = kZInit;
    }
    float4 cn[4], rn0[4], cp[4], rp0[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (ri[k] >= 0 && ci[k] >= 0) {
        cn[k] = gload4(fc.nrm + ci[k]);
        rn0[k] = gload4(fr.nrm + ri[k]);
        cp[k] = gload4(fc.pts + ci[k]);
        rp0[k] = gload4(fr.pts + ri[k]);
      }
    }
    bool ok[4];
    float rp[4][3], rn[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      ok[k] = false;
      if (ri[k] >= 0 && ci[k] >= 0) {
        acc[30] += 1.0f;
        ok[k] = gates_loaded<Closed>(cn[k], rn0[k], cp[k], rp0[k], T, ap, rp[k], rn[k]);
      }
    }
    // the current points' information matrices: loaded, or in closed form from their normals
    float4 om[4][3];
    if (!Closed) {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (ok[k]) {
          const float4 *o = fc.omg + 3 * (size_t)ci[k];
          om[k][0] = gload4(o);
          om[k][1] = gload4(o + 1);
          om[k][2] = gload4(o + 2);
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (store_pairs && pix[k] >= 0)
        *gmut((unsigned long long *)pd.pairs + pix[k]) =
            ok[k] ? ((unsigned long long)(unsigned int)ci[k] << 32) | (unsigned int)ri[k] : ~0ull;
      if (ok[k]) {
        if (Closed) closed_omega(om_model, cn[k], om[k][0], om[k][1], om[k][2]);
        lin_terms(rp[k], rn[k], cp[k], cn[k], om[k][0], om[k][1], om[k][2], ap, acc);
      }
    }
  }
}

// Fused correspondence + linearisation over the projector image of each pair, one block per
// kDTH x kDTW tile (blk_dpair maps blocks to pairs). Reads and resets the z-buffer of the
// reference projection.
extern "C" __global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_corr_lin(
    const PairDev *__restrict__ Pp, const FrameDev *__restrict__ F, const int *__restrict__ blk_dpair,
    AlignParamsDev ap, int store_pairs) {
  const int g = xcd_tile(blockIdx.x, gridDim.x);
  const PairDev &pd = Pp[blk_dpair[g]];
  const FrameDev &fr = F[pd.ref];
  const FrameDev &fc = F[pd.cur];
  const int t = g - pd.dtile_base;
  float T[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) T[k] = *gview(pd.state + kSlotTinv + k);
  const bool multi = ap.row_region != 0 || ap.col_region != 0;
  Acc2 acc[kVals];
#pragma unroll
  for (int v = 0; v < kVals; ++v) acc[v] = 0.0f;
  if (!multi) {
    if (fc.omega_closed && fr.omega_closed)
      corr_lin_tile<false, true>(pd, fr, fc, T, ap, store_pairs, t, nullptr, omega_model(fc), acc);
    else corr_lin_tile<false, false>(pd, fr, fc, T, ap, store_pairs, t, nullptr, OmegaModel{}, acc);
  } else {
    const int sub = pd.sub;
    const int GR = (pd.rows + sub - 1) / sub, GC = (pd.cols + sub - 1) / sub;
    const int ty = t / pd.dtiles_x, tx = t - ty * pd.dtiles_x;
    const int gc = tx * kDTW + (threadIdx.x & 63);
    for (int lrow = threadIdx.x >> 6; lrow < kDTH; lrow += 4) {
      const int gr = ty * kDTH + lrow;
      if (gr >= GR || gc >= GC) continue;
      const int r = gr * sub, c = gc * sub;
      const int p = r * pd.cols + c;
      const unsigned long long key = pd.zbuf[p];
      pd.zbuf[p] = kZInit;
      const int ri = key == kZInit ? -1 : (int)(unsigned int)(key & 0xffffffffull);
      int ci = -1;
      float rp[3], rn[3];
      float4 cp, cn;
      bool ok = false;
      if (ri >= 0) {
        ci = window_search(fr, fc, pd.cur_index, pd.rows, pd.cols, r, c, ri, T, ap, rp, rn);
        if (ci >= 0) {
          ok = true;
          cp = fc.pts[ci];
          cn = fc.nrm[ci];
        }
      }
      if (store_pairs) pd.pairs[p] = ok ? make_int2(ri, ci) : make_int2(-1, -1);
      if (ok) {
        float4 o0, o1, o2;
        if (fc.omega_closed) {
          closed_omega(omega_model(fc), cn, o0, o1, o2);
        } else {
          const float4 *o = fc.omg + 3 * (size_t)ci;
          o0 = o[0];
          o1 = o[1];
          o2 = o[2];
        }
        lin_terms(rp, rn, cp, cn, o0, o1, o2, ap, acc);
      }
    }
  }
  block_partial(acc, pd.partials + (size_t)t * kPartial);
}

// ---- tile-binned projection fused with the correspondences and the linearisation ----------
// The reference cloud's z-buffer of a pair (PinholePointProjector::project,
// pinholepointprojector.cpp:33-66) is built per fused tile in LDS: k_sbox bounds, for every source
// tile of the reference frame, the fused tiles its projection may reach, and one k_proj_corr_lin
// block per fused tile lists the source tiles whose bound holds it, projects exactly their points
// into its LDS z-buffer with the 64-bit minimum on (depth bits << 32 | index), then runs
// k_corr_lin's correspondence and linearisation code on it. Every point that lands in the tile
// comes from a listed source tile and the minimum does not depend on the order, so the z-buffer,
// the correspondences and the partial sums are those of k_zproject + k_corr_lin, bit for bit.

// One thread per source tile of each pair's reference frame: the bounding box of the tile's
// projection under the pair's KRt, from the eight corners of its frustum (the pixel corners of
// the tile at its minimum and maximum depth: the tile's points lie in their convex hull, and a
// perspective projection maps the hull into the hull of the projected corners when every corner
// lies in front of the camera), widened by kBinMargin pixels, as the range of fused tiles
// (fx0, fy0, fx1, fy1) it touches; empty (fx0 > fx1) when no point of the tile can project into
// the image. A tile whose frustum may reach the camera plane covers every fused tile.
extern "C" __global__ __launch_bounds__(kThreads) void k_sbox(const PairDev *__restrict__ Pp,
                                                             const FrameDev *__restrict__ F,
                                                             const int *__restrict__ blk_spair) {
  const int g = blockIdx.x;
  const PairDev &pd = Pp[blk_spair[g]];
  const FrameDev &fr = F[pd.ref];
  const int sx = stiles_x(fr.cols);
  const int s = (g - pd.sblk_base) * kThreads + threadIdx.x;
  if (s >= stiles(fr.rows, fr.cols)) return;
  int4 box = make_int4(1, 1, 0, 0);
  const gvec2 dr = *gview((const gvec2 *)(fr.stile_d + s));
  if (dr.x <= dr.y) {  // the tile has valid pixels
    float KRt[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) KRt[k] = *gview(pd.state + kSlotKRt + k);
    const int ty = s / sx, tx = s - ty * sx;
    const int c0 = tx * kSTW, c1 = min(c0 + kSTW, fr.cols) - 1, r0 = ty * kSTH, r1 = min(r0 + kSTH, fr.rows) - 1;
    float umin = INFINITY, umax = -INFINITY, vmin = INFINITY, vmax = -INFINITY, zmin = INFINITY, zmax = -INFINITY;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float4 p = unproject_point(fr.iKRt, (k & 1) ? c1 : c0, (k & 2) ? r1 : r0, (k & 4) ? dr.y : dr.x);
      float ip[4];
      mul4_packet(KRt, p.x, p.y, p.z, 1.0f, ip);
      zmin = fminf(zmin, ip[2]);
      zmax = fmaxf(zmax, ip[2]);
      const float u = ip[0] / ip[2], v = ip[1] / ip[2];
      umin = fminf(umin, u);
      umax = fmaxf(umax, u);
      vmin = fminf(vmin, v);
      vmax = fmaxf(vmax, v);
    }
    // depth range with a relative margin for the rounding of the points' own depths
    if (!(zmax < pd.min_d * (1.0f - 1e-4f) || zmin > pd.max_d * (1.0f + 1e-4f))) {
      int x0 = 0, x1 = pd.cols - 1, y0 = 0, y1 = pd.rows - 1;
      if (zmin > 0.1f * pd.min_d && umin >= -1e6f && umax <= 1e6f && vmin >= -1e6f && vmax <= 1e6f) {
        x0 = max(x0, (int)floorf(umin) - kBinMargin);
        x1 = min(x1, (int)ceilf(umax) + kBinMargin);
        y0 = max(y0, (int)floorf(vmin) - kBinMargin);
        y1 = min(y1, (int)ceilf(vmax) + kBinMargin);
      }
      if (x0 <= x1 && y0 <= y1) box = make_int4(x0 / kDTW, y0 / kFH, x1 / kDTW, y1 / kFH);
    }
  }
  *gmut((gint4 *)pd.sbox + s) = gint4{box.x, box.y, box.z, box.w};
}

// One block per fused tile (blk_fpair, in the XCD-aware order of k_corr_lin): the list of the
// reference frame's source tiles whose box holds the tile (every source tile when more than
// list_cap do), the LDS z-buffer of the tile's kFH x kDTW pixels from their points, then the
// kFH / kDTH correspondence tiles' correspondences and linearisation terms, one block partial each.
extern "C" __global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_proj_corr_lin(
    const PairDev *__restrict__ Pp, const FrameDev *__restrict__ F, const int *__restrict__ blk_fpair,
    AlignParamsDev ap, int store_pairs, int list_cap) {
  __shared__ unsigned long long zl[kFH * kDTW];
  __shared__ int sl[kListCap];
  __shared__ int scount;
  const int g = xcd_tile(blockIdx.x, gridDim.x);
  const PairDev &pd = Pp[blk_fpair[g]];
  const FrameDev &fr = F[pd.ref];
  const FrameDev &fc = F[pd.cur];
  const int ft = g - pd.ftile_base;
  const int fy = ft / pd.dtiles_x, fx = ft - fy * pd.dtiles_x;
  for (int i = threadIdx.x; i < kFH * kDTW; i += kThreads) zl[i] = kZInit;
  if (threadIdx.x == 0) scount = 0;
  __syncthreads();
  {
    // the list, compacted per wavefront: one LDS atomic per wavefront and round
    const int ns = stiles(fr.rows, fr.cols), lane = threadIdx.x & 63;
    const __attribute__((address_space(1))) gint4 *box = gview((const gint4 *)pd.sbox);
    for (int s0 = threadIdx.x - lane; s0 < ns; s0 += kThreads) {
      const int s = s0 + lane;
      bool hit = false;
      if (s < ns) {
        const gint4 b = box[s];
        hit = fx >= b.x && fx <= b.z && fy >= b.y && fy <= b.w;
      }
      const unsigned long long m = __ballot(hit);
      int base = 0;
      if (lane == 0 && m) base = atomicAdd(&scount, __popcll(m));
      base = __shfl(base, 0, 64);
      const int slot = base + __popcll(m & (lane ? ~0ull >> (64 - lane) : 0ull));
      if (hit && slot < list_cap) sl[slot] = s;
    }
  }
  __syncthreads();
  const int cnt = scount;
  const bool all = cnt > list_cap;
  {
    // the frame and pair fields the loop needs, held in registers: a field read from the tables
    // inside the loop is a scalar load, and waiting for it would also wait for the LDS atomics
    float KRt[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) KRt[k] = *gview(pd.state + kSlotKRt + k);
    const int frows = fr.rows, fcols = fr.cols, prows = pd.rows, pcols = pd.cols;
    float iKRt[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) iKRt[k] = fr.iKRt[k];
    const float mind = pd.min_d, maxd = pd.max_d;
    const __attribute__((address_space(1))) gvec2u *spix = gview((const gvec2u *)fr.spix);
    const __attribute__((address_space(1))) int *stn = gview(fr.stile_n);
    const int sx = stiles_x(fcols);
    const int nsrc = all ? stiles(frows, fcols) : cnt;
    const int xlo = fx * kDTW, ylo = fy * kFH;
    const int xspan = min(kDTW, pcols - xlo), yspan = min(kFH, prows - ylo);  // the tile's part of the image
    // wavefront w takes source tiles w, w + 4, ...: its list entries and their point counts are
    // read once (lane i holds entry w + 4 i) and taken with readlane, so no LDS wait inside the
    // loop; lane l projects the tile's packed pixels l, l + 64, ... (at most kPhaseB of them),
    // all loads in flight together. A pixel's point is unproject_point of its column, row and
    // depth, as the conversion computed it: the cloud point's bits without loading the point.
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nw = kThreads / 64;
    const int mine = wave + nw * lane;
    const int entry = !all && mine < nsrc ? sl[mine] : 0;
    const int ecount = !all && mine < nsrc ? stn[entry] : 0;
    const int ntiles = nsrc > wave ? (nsrc - wave + nw - 1) / nw : 0;
    for (int i = 0; i < ntiles; ++i) {
      const int st = all ? wave + nw * i : __builtin_amdgcn_readlane(entry, i);
      const int n = all ? stn[st] : __builtin_amdgcn_readlane(ecount, i);
      const int sty = st / sx, stx = st - sty * sx;
      const __attribute__((address_space(1))) gvec2u *tp = spix + (size_t)st * (kSTH * kSTW);
      gvec2u q[kPhaseB];
#pragma unroll
      for (int u = 0; u < kPhaseB; ++u)
        if (u * 64 + lane < n) q[u] = tp[u * 64 + lane];
#pragma unroll
      for (int u = 0; u < kPhaseB; ++u) {
        if (u * 64 + lane >= n) continue;
        const unsigned int at = q[u].y >> kSPosShift, idx = q[u].y & ((1u << kSPosShift) - 1u);
        // the cloud point's w is 1 exactly: the last row of iKRt is (0, 0, 0, 1) and the column,
        // row and depth are finite and non-negative
        const float4 p = unproject_point(iKRt, stx * kSTW + (int)(at % kSTW), sty * kSTH + (int)(at / kSTW),
                                         __uint_as_float(q[u].x));
        // pinhole_project's arithmetic on the register copies; the image bounds and the tile's
        // bounds together are one range per axis
        float ip[4];
        mul4_packet(KRt, p.x, p.y, p.z, 1.0f, ip);
        const float d = ip[2];
        if (d < mind || d > maxd) continue;
        const float inv = 1.0f / d;
        const int lx = (int)roundf(ip[0] * inv) - xlo, ly = (int)roundf(ip[1] * inv) - ylo;
        if ((unsigned int)lx < (unsigned int)xspan && (unsigned int)ly < (unsigned int)yspan)
          atomicMin(&zl[ly * kDTW + lx], ((unsigned long long)__float_as_uint(d) << 32) | idx);
      }
    }
  }
  __syncthreads();
  float T[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) T[k] = *gview(pd.state + kSlotTinv + k);
  const int nty = dtiles_y(pd.rows);
  const bool closed = fc.omega_closed != 0 && fr.omega_closed != 0;
  const OmegaModel om_model = closed ? omega_model(fc) : OmegaModel{};
#pragma unroll 1
  for (int sub_t = 0; sub_t < kFH / kDTH; ++sub_t) {
    const int ty = fy * (kFH / kDTH) + sub_t;
    if (ty >= nty) break;
    Acc2 acc[kVals];
#pragma unroll
    for (int v = 0; v < kVals; ++v) acc[v] = 0.0f;
    const int t = ty * pd.dtiles_x + fx;
    if (closed) corr_lin_tile<true, true>(pd, fr, fc, T, ap, store_pairs, t, zl + sub_t * kDTH * kDTW, om_model, acc);
    else corr_lin_tile<true, false>(pd, fr, fc, T, ap, store_pairs, t, zl + sub_t * kDTH * kDTW, om_model, acc);
    block_partial(acc, pd.partials + (size_t)t * kPartial, sub_t & 1);
  }
}

// Linearizer::update over stored pairs at the transform in state slot `slot`.
extern "C" __global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_lin_pairs(
    const PairDev *__restrict__ Pp, const FrameDev *__restrict__ F, const int *__restrict__ blk_pair, AlignParamsDev ap,
    int slot, int num_override) {
  const int g = blockIdx.x;
  const PairDev &pd = Pp[blk_pair[g]];
  const FrameDev &fr = F[pd.ref];
  const FrameDev &fc = F[pd.cur];
  const int P = num_override > 0 ? num_override : pd.rows * pd.cols;
  const int t = g - pd.tile_base;
  const int base = t * kTile;
  float T[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) T[k] = *gview(pd.state + slot + k);
  Acc2 acc[kVals];
#pragma unroll
  for (int v = 0; v < kVals; ++v) acc[v] = 0.0f;
  int2 pr[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int p = base + k * kThreads + threadIdx.x;
    if (p < P) {
      const unsigned long long v = *gview((const unsigned long long *)pd.pairs + p);
      pr[k] = make_int2((int)(unsigned int)v, (int)(unsigned int)(v >> 32));
    } else {
      pr[k] = make_int2(-1, -1);
    }
  }
  // the current points' information matrices: loaded, or in closed form from their normals
  const bool closed = fc.omega_closed != 0;
  const OmegaModel om_model = closed ? omega_model(fc) : OmegaModel{};
  float4 rp0[4], rn0[4], cp[4], cn[4], om[4][3];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (pr[k].x >= 0) {
      rp0[k] = gload4(fr.pts + pr[k].x);
      rn0[k] = gload4(fr.nrm + pr[k].x);
      cp[k] = gload4(fc.pts + pr[k].y);
      cn[k] = gload4(fc.nrm + pr[k].y);
      if (!closed) {
        const float4 *o = fc.omg + 3 * (size_t)pr[k].y;
        om[k][0] = gload4(o);
        om[k][1] = gload4(o + 1);
        om[k][2] = gload4(o + 2);
      }
    }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (pr[k].x >= 0) {
      float rp[3], rn[3];
      iso_mul3(T, rp0[k].x, rp0[k].y, rp0[k].z, rp0[k].w, rp);
      iso_rot3(T, rn0[k].x, rn0[k].y, rn0[k].z, rn);
      if (closed) closed_omega(om_model, cn[k], om[k][0], om[k][1], om[k][2]);
      lin_terms(rp, rn, cp[k], cn[k], om[k][0], om[k][1], om[k][2], ap, acc);
    }
  block_partial(acc, pd.partials + (size_t)t * kPartial);
}

namespace {
// The pair's FP64 totals: kSolveLanes consecutive threads per slot v, thread `part` adding
// tiles part, part + kSolveLanes, ... in order with eight loads in flight, then a fixed tree
// over the kSolveLanes sums. The order is fixed, so every call on the same partials gives
// the same bits.
__device__ void reduce_partials(const PairDev &pd, int nparts, double *tot /* shared, kPartial */) {
  const int v = threadIdx.x / kSolveLanes, part = threadIdx.x - v * kSolveLanes;
  double s = 0.0;
  if (v < kVals) {
    const double *__restrict__ src = pd.partials + v;
    int i = part;
    for (; i + 7 * kSolveLanes < nparts; i += 8 * kSolveLanes) {
      double x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) x[u] = src[(size_t)(i + u * kSolveLanes) * kPartial];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += x[u];
    }
    for (; i < nparts; i += kSolveLanes) s += src[(size_t)i * kPartial];
  }
  // the slot's lanes are consecutive lanes of one wavefront
#pragma unroll
  for (int o = kSolveLanes / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (v < kVals && part == 0) tot[v] = s;
}
}  // namespace

// One workgroup per pair: the FP64 totals by all its threads, then on thread 0 H + lambda I,
// LDLT and the update of the SE(3) state.
// mode 0: last inner iteration (updates T, Tinv, KRt); 1: inner iteration (updates invT_inner);
// 2: statistics linearisation (stores the totals only). dtiles: the partials come from
// k_corr_lin's n_dtiles tiles (1) or from k_lin_pairs' n_tiles tiles (0).
extern "C" __global__ __launch_bounds__(kSolveThreads) void k_solve(const PairDev *__restrict__ Pp, float lambda,
                                                                   int mode, int from_inner, int dtiles) {
  const PairDev &pd = Pp[blockIdx.x];
  __shared__ double tot[kPartial];
  reduce_partials(pd, dtiles ? pd.n_dtiles : pd.n_tiles, tot);
  __syncthreads();
  if (threadIdx.x != 0) return;
  double *res = pd.result + (mode == 2 ? kPartial : 0);
  for (int v = 0; v < kVals; ++v) res[v] = tot[v];
  if (mode == 2) return;
  // H from the row-major upper triangle
  double H[36], b[6];
  int e = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 6; ++j) {
      H[j * 6 + i] = tot[e];
      H[i * 6 + j] = tot[e];
      ++e;
    }
  for (int i = 0; i < 6; ++i) b[i] = -tot[21 + i];
  for (int i = 0; i < 6; ++i) H[i * 6 + i] += (double)lambda;
  double dx[6];
  ldlt_solve6<double>(H, b, dx);
  float dxf[6];
  for (int i = 0; i < 6; ++i) dxf[i] = (float)dx[i];
  float *S = pd.state;
  float invT[16];
  for (int k = 0; k < 16; ++k) invT[k] = S[(from_inner ? kSlotInvTInner : kSlotTinv) + k];
  set_row3(invT);
  float dT[16];
  v2t<float>(dxf, dT);
  iso_mul(dT, invT, invT);
  if (mode == 1) {
    set_row3(invT);
    for (int k = 0; k < 16; ++k) S[kSlotInvTInner + k] = invT[k];
    return;
  }
  float T[16];
  iso_inverse(invT, T);
  float x[6];
  t2v<float>(T, x);
  v2t<float>(x, T);
  set_row3(T);
  // the reference projector's transform T * reference sensor offset (alignerprojective.cpp:51)
  float Ti[16], RO[16], TR[16], KRt[16];
  iso_inverse(T, Ti);
  for (int k = 0; k < 16; ++k) RO[k] = S[kSlotRefOffset + k];
  iso_mul(T, RO, TR);
  projector_state(pd.model, pd.K, TR, KRt);
  for (int k = 0; k < 16; ++k) {
    S[kSlotT + k] = T[k];
    S[kSlotTinv + k] = Ti[k];
    S[kSlotKRt + k] = KRt[k];
  }
}

// Aligner::_computeStatistics (aligner.cpp:82-135) of every pair, one wavefront per pair: the
// FP64 unscented statistics from the statistics linearisation's totals (result slots
// kPartial..kPartial + 20, H's row-major upper triangle) at the final transform, into result
// slots 2 kPartial ...
static_assert(kResultDoubles == 2 * kPartial + nicp_stats::kSolStats, "result record layout");
extern "C" __global__ __launch_bounds__(64) void k_solution_stats(const PairDev *__restrict__ Pp,
                                                                 float trans_max_ratio, float rot_max_ratio) {
  const PairDev &pd = Pp[blockIdx.x];
  __shared__ double H[36];
  const int lane = threadIdx.x;
  if (lane < 36) {
    const int r = lane % 6, c = lane / 6, i = min(r, c), j = max(r, c);
    H[lane] = pd.result[kPartial + 6 * i - i * (i - 1) / 2 + (j - i)];
  }
  __syncthreads();
  float T[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) T[k] = pd.state[kSlotT + k];
  nicp_stats::solution_statistics_wave(H, T, trans_max_ratio, rot_max_ratio, pd.result + 2 * kPartial);
}

// ---------------------------------------------------------------------------------------
// kNN correspondences (CorrespondenceFinderNN, correspondencefindernn.cpp:13-97): an exact
// 6D nearest neighbour [p; s n] within the squared inlier distance, found on a uniform 3D
// grid over the current cloud, followed by the reference's gates.
// ---------------------------------------------------------------------------------------

// order-preserving float <-> uint for atomic min/max
__device__ __forceinline__ unsigned int f2ord(float f) {
  unsigned int u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// bounding boxes of pair blockIdx.y's current (blockIdx.z = 0) and reference (1) clouds:
// bbox[12 y + 6 z + 0..2] = min keys (init 0xffffffff), bbox[12 y + 6 z + 3..5] = max keys (init 0)
extern "C" __global__ __launch_bounds__(kThreads) void k_nn_bbox(const PairDev *__restrict__ Pp,
                                                                const FrameDev *__restrict__ F,
                                                                unsigned int *__restrict__ bbox) {
  const FrameDev &fc = F[blockIdx.z ? Pp[blockIdx.y].ref : Pp[blockIdx.y].cur];
  const int n = *fc.npoints;
  const float4 *__restrict__ pts = fc.pts;
  bbox += 12 * blockIdx.y + 6 * blockIdx.z;
  float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const float4 p = pts[i];
    mn[0] = fminf(mn[0], p.x); mn[1] = fminf(mn[1], p.y); mn[2] = fminf(mn[2], p.z);
    mx[0] = fmaxf(mx[0], p.x); mx[1] = fmaxf(mx[1], p.y); mx[2] = fmaxf(mx[2], p.z);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    for (int o = 32; o > 0; o >>= 1) {
      mn[k] = fminf(mn[k], __shfl_xor(mn[k], o, 64));
      mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], o, 64));
    }
  }
  // the block's waves meet in LDS; one set of atomics per block (all blocks of a cloud
  // update the same six words)
  __shared__ float red[kThreads / 64][6];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      red[w][k] = mn[k];
      red[w][3 + k] = mx[k];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float a = red[0][k], b = red[0][3 + k];
      for (int v = 1; v < kThreads / 64; ++v) {
        a = fminf(a, red[v][k]);
        b = fmaxf(b, red[v][3 + k]);
      }
      atomicMin(bbox + k, f2ord(a));
      atomicMax(bbox + 3 + k, f2ord(b));
    }
  }
}

// one pair's grid; its nx*ny*nz cell starts and a sentinel live at cell_base of the
// batch's cell arena
struct GridDev {
  float ox, oy, oz, h, inv_h;
  int nx, ny, nz;
  int cell_base;
  int order_base;  // the pair's first entry in the query order
};

__device__ __forceinline__ int cell_coord(float v, float o, float inv_h) { return (int)floorf((v - o) * inv_h); }

__device__ __forceinline__ int grid_cell(const GridDev &g, const float4 &p) {
  const int cx = min(max(cell_coord(p.x, g.ox, g.inv_h), 0), g.nx - 1);
  const int cy = min(max(cell_coord(p.y, g.oy, g.inv_h), 0), g.ny - 1);
  const int cz = min(max(cell_coord(p.z, g.oz, g.inv_h), 0), g.nz - 1);
  return g.cell_base + (cz * g.ny + cy) * g.nx + cx;
}

// points per cell of pair blockIdx.y's current cloud
extern "C" __global__ __launch_bounds__(kThreads) void k_nn_count(const PairDev *__restrict__ Pp,
                                                                 const FrameDev *__restrict__ F,
                                                                 const GridDev *__restrict__ grids,
                                                                 int *__restrict__ counts) {
  const FrameDev &fc = F[Pp[blockIdx.y].cur];
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= *fc.npoints) return;
  atomicAdd(counts + grid_cell(grids[blockIdx.y], fc.pts[i]), 1);
}

// exclusive scan of n ints, pass 1: each block of 1024 threads scans 4096 entries
extern "C" __global__ __launch_bounds__(1024) void k_scan_local(int *__restrict__ data, int n, int *__restrict__ block_sums) {
  const int base = blockIdx.x * 4096 + threadIdx.x * 4;
  int v[4], s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = base + k < n ? data[base + k] : 0;
    s += v[k];
  }
  __shared__ int wsum[16];
  int incl = wave_incl_scan(s);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wave; ++w) before += wsum[w];
  int run = before + incl - s;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) data[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 1023) block_sums[blockIdx.x] = before + incl;
}

// pass 2: one block scans the block sums (up to 1024 * 16 of them); total to *total
extern "C" __global__ __launch_bounds__(1024) void k_scan_sums(int *__restrict__ sums, int nb, int *__restrict__ total) {
  const int chunk = (nb + 1023) / 1024, start = threadIdx.x * chunk;
  int local = 0;
  for (int i = 0; i < chunk; ++i)
    if (start + i < nb) local += sums[start + i];
  __shared__ int wsum[16];
  int incl = wave_incl_scan(local);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wave; ++w) before += wsum[w];
  int run = before + incl - local;
  for (int i = 0; i < chunk; ++i)
    if (start + i < nb) {
      int c = sums[start + i];
      sums[start + i] = run;
      run += c;
    }
  if (threadIdx.x == 1023) *total = before + incl;
}

// pass 3: add the block offsets; also writes the cursor copy used by the scatter
extern "C" __global__ __launch_bounds__(1024) void k_scan_add(int *__restrict__ data, int n, const int *__restrict__ sums,
                                                             int *__restrict__ cursor) {
  const int base = blockIdx.x * 4096 + threadIdx.x * 4;
  const int off = sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + k < n) {
      int v = data[base + k] + off;
      data[base + k] = v;
      cursor[base + k] = v;
    }
}

// sorted current clouds: (x, y, z, index bits) and (s nx, s ny, s nz, 0) per point, in cell
// order; the cell arena's exclusive scan puts every pair's points after the previous pair's
extern "C" __global__ __launch_bounds__(kThreads) void k_nn_scatter(const PairDev *__restrict__ Pp,
                                                                   const FrameDev *__restrict__ F,
                                                                   const GridDev *__restrict__ grids, float ns,
                                                                   int *__restrict__ cursor, float4 *__restrict__ sorted) {
  const FrameDev &fc = F[Pp[blockIdx.y].cur];
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= *fc.npoints) return;
  const float4 p = fc.pts[i];
  const float4 n = fc.nrm[i];
  const int slot = atomicAdd(cursor + grid_cell(grids[blockIdx.y], p), 1);
  sorted[2 * (size_t)slot] = make_float4(p.x, p.y, p.z, __int_as_float(i));
  sorted[2 * (size_t)slot + 1] = make_float4(ns * n.x, ns * n.y, ns * n.z, 0.0f);
}

// Query order: the reference points of every pair sorted by a Morton code of a coarse cubic
// grid (2^bits cells per axis, bits in nx) over the reference cloud's bounding box, so that the
// 64 queries of a wavefront lie close together. The order changes the speed only.
__device__ __forceinline__ unsigned int spread3(unsigned int v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__device__ __forceinline__ int order_key(const GridDev &o, const float4 &p) {
  const int top = (1 << o.nx) - 1;
  const unsigned int cx = (unsigned int)min(max(cell_coord(p.x, o.ox, o.inv_h), 0), top);
  const unsigned int cy = (unsigned int)min(max(cell_coord(p.y, o.oy, o.inv_h), 0), top);
  const unsigned int cz = (unsigned int)min(max(cell_coord(p.z, o.oz, o.inv_h), 0), top);
  return o.cell_base + (int)(spread3(cx) | (spread3(cy) << 1) | (spread3(cz) << 2));
}
extern "C" __global__ __launch_bounds__(kThreads) void k_nn_order_count(const PairDev *__restrict__ Pp,
                                                                       const FrameDev *__restrict__ F,
                                                                       const GridDev *__restrict__ ogrids,
                                                                       int *__restrict__ counts) {
  const FrameDev &fr = F[Pp[blockIdx.y].ref];
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= *fr.npoints) return;
  atomicAdd(counts + order_key(ogrids[blockIdx.y], fr.pts[i]), 1);
}
extern "C" __global__ __launch_bounds__(kThreads) void k_nn_order_scatter(const PairDev *__restrict__ Pp,
                                                                         const FrameDev *__restrict__ F,
                                                                         const GridDev *__restrict__ ogrids,
                                                                         int *__restrict__ cursor,
                                                                         int *__restrict__ order) {
  const FrameDev &fr = F[Pp[blockIdx.y].ref];
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= *fr.npoints) return;
  order[atomicAdd(cursor + order_key(ogrids[blockIdx.y], fr.pts[i]), 1)] = i;
}

// 6D squared distance in FLANN's L2 functor order for 6 dimensions
__device__ __forceinline__ float nn_dist6(const float q[6], float ax, float ay, float az, float bx, float by,
                                          float bz) {
  const float d0 = q[0] - ax, d1 = q[1] - ay, d2 = q[2] - az, d3 = q[3] - bx, d4 = q[4] - by, d5 = q[5] - bz;
  float d = 0.0f + (d0 * d0 + d1 * d1 + d2 * d2 + d3 * d3);
  d += d4 * d4;
  d += d5 * d5;
  return d;
}
// keeps the smallest distance below the initial bound, ties to the lower index: the result
// does not depend on the order in which candidates are seen
__device__ __forceinline__ void nn_take(float d, int idx, float &best, int &bi) {
  // non-short-circuit form: two selects instead of branches
  const bool take = (d < best) | ((d == best) & (bi >= 0) & (idx < bi));
  best = take ? d : best;
  bi = take ? idx : bi;
}
// shell k >= 1 may be skipped once every point outside shells 0..k-1 is farther than best (a
// point in a cell at Chebyshev cell distance m >= 1 is at least (m - 1) h away)
__device__ __forceinline__ bool nn_done(int k, float h, float r2, float best) {
  if (k < 1) return false;
  const float lb = (float)(k - 1) * h;
  const float lb2 = lb * lb * (1.0f - 1e-5f);
  return lb2 >= r2 || best < lb2;
}

// one lane's search: shells of cells around its own cell (cx, cy, cz)
__device__ void nn_walk(const float q[6], int cx, int cy, int cz, const GridDev &g, const int *__restrict__ cell_start,
                        const float4 *__restrict__ sorted, float r2, float &best, int &bi) {
  // Chebyshev distance from the query's cell to the farthest grid cell bounds the shells
  const int kcap = max(max(max(abs(cx), abs(cx - (g.nx - 1))), max(abs(cy), abs(cy - (g.ny - 1)))),
                       max(abs(cz), abs(cz - (g.nz - 1))));
  for (int k = 0; k <= kcap; ++k) {
    if (nn_done(k, g.h, r2, best)) break;
    for (int dz = -k; dz <= k; ++dz) {
      const int z = cz + dz;
      if (z < 0 || z >= g.nz) continue;
      for (int dy = -k; dy <= k; ++dy) {
        const int y = cy + dy;
        if (y < 0 || y >= g.ny) continue;
        const bool face = (dz == -k || dz == k || dy == -k || dy == k);
        const int step = face ? 1 : 2 * k;
        for (int dx = -k; dx <= k; dx += (step > 0 ? step : 1)) {
          const int x = cx + dx;
          if (x >= 0 && x < g.nx) {
            const size_t cell = ((size_t)z * g.ny + y) * g.nx + x;
            const int s = cell_start[cell], e = cell_start[cell + 1];
            for (int j = s; j < e; ++j) {
              const float4 a = sorted[2 * (size_t)j], b = sorted[2 * (size_t)j + 1];
              nn_take(nn_dist6(q, a.x, a.y, a.z, b.x, b.y, b.z), __float_as_int(a.w), best, bi);
            }
          }
          if (k == 0) break;
        }
      }
    }
  }
}

// The points of up to two ranges per lane ([s1, s1 + n1) and [s2, s2 + n2) of the sorted
// arena) against every lane's query. The ranges are packed into rounds of 64 points: point
// `pos` of the concatenation belongs to the last lane whose exclusive offset is <= pos (a
// binary search over the lanes' offsets). Each lane loads one point per round into the
// wavefront's LDS slots, from where every lane reads each point at one (broadcast) address.
// The running best is the key (distance bits << 32 | index): distances are >= +0, so one
// unsigned compare orders by distance and then by index, as nn_take does.
__device__ __forceinline__ void nn_scan_packed(const float q[6], const float4 *__restrict__ sorted, int s1, int n1,
                                               int s2, int n2, int lane, float4 (*sh)[64],
                                               unsigned long long &key) {
  const int len = n1 + n2;
  const int incl = wave_incl_scan(len);
  const int total = __builtin_amdgcn_readlane(incl, 63);
  const int excl = incl - len;
  for (int b0 = 0; b0 < total; b0 += 64) {
    const int pos = b0 + lane;
    int lo = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
      const int v = __shfl(excl, lo + step, 64);
      lo = v <= pos ? lo + step : lo;
    }
    // every shuffle outside any branch: ds_bpermute reads 0 from a lane that EXEC disables
    const int o = pos - __shfl(excl, lo, 64);
    const int m1 = __shfl(n1, lo, 64);
    const int a1 = __shfl(s1, lo, 64), a2 = __shfl(s2, lo, 64);
    const int j = o < m1 ? a1 + o : a2 + (o - m1);
    float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f), b = a;
    if (pos < total) {
      a = sorted[2 * (size_t)j];
      b = sorted[2 * (size_t)j + 1];
    }
    sh[0][lane] = a;
    sh[1][lane] = b;
    // one wavefront writes and reads its own slots: wait for the writes, keep the order
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    const int cnt = min(64, total - b0);
#pragma unroll 4
    for (int t = 0; t < cnt; ++t) {
      const float4 pa = sh[0][t], pb = sh[1][t];
      const float d = nn_dist6(q, pa.x, pa.y, pa.z, pb.x, pb.y, pb.z);
      const unsigned long long c = ((unsigned long long)__float_as_uint(d) << 32) | __float_as_uint(pa.w);
      key = c < key ? c : key;
    }
    __builtin_amdgcn_wave_barrier();
  }
}

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return __builtin_amdgcn_readfirstlane(v);
}

constexpr int kNNWaveBox = 8;  // widest box of query cells (per axis) a wavefront searches together

// the wavefront's search: shells of cells around the box [lo, hi] of its active queries' cells.
// Every cell within Chebyshev distance k of a lane's cell lies in the box grown by k, so after
// shells 0..k-1 a lane has seen every cell the per-lane walk sees up to the same k, and stops
// by the same bound. Lanes keep scanning (with no effect on their exact minimum) until all
// lanes are done. 64 rows of a shell look up their cell ranges at once and scan them packed.
__device__ void nn_wave(const float q[6], int lx, int ly, int lz, int hx, int hy, int hz, const GridDev &g,
                        const int *__restrict__ cell_start, const float4 *__restrict__ sorted, float r2, bool active,
                        int lane, float4 (*sh)[64], float &best, int &bi) {
  const int kcap = max(max(max(lx, g.nx - 1 - hx), max(ly, g.ny - 1 - hy)), max(max(lz, g.nz - 1 - hz), 0));
  // the initial key admits exactly the distances below r2 (none when r2 is NaN)
  const unsigned long long key0 = r2 == r2 ? (unsigned long long)__float_as_uint(r2) << 32 : 0ull;
  unsigned long long key = key0;
  for (int k = 0; k <= kcap; ++k) {
    best = __uint_as_float((unsigned int)(key >> 32));
    if (__ballot(active && !nn_done(k, g.h, r2, best)) == 0) break;
    const int zlo = max(lz - k, 0), zhi = min(hz + k, g.nz - 1), ylo = max(ly - k, 0), yhi = min(hy + k, g.ny - 1);
    if (zlo > zhi || ylo > yhi) continue;
    const int yc = yhi - ylo + 1, rows = yc * (zhi - zlo + 1);
    for (int r0 = 0; r0 < rows; r0 += 64) {
      const int r = r0 + lane;
      int s1 = 0, e1 = 0, s2 = 0, e2 = 0;
      if (r < rows) {
        const int y = ylo + r % yc, z = zlo + r / yc;
        const int row = (z * g.ny + y) * g.nx;
        if (k == 0 || z == lz - k || z == hz + k || y == ly - k || y == hy + k) {
          const int x0 = max(lx - k, 0), x1 = min(hx + k, g.nx - 1);
          if (x0 <= x1) {
            s1 = cell_start[row + x0];
            e1 = cell_start[row + x1 + 1];
          }
        } else {
          const int xa = lx - k, xb = hx + k;
          if (xa >= 0 && xa < g.nx) {
            s1 = cell_start[row + xa];
            e1 = cell_start[row + xa + 1];
          }
          if (xb >= 0 && xb < g.nx) {
            s2 = cell_start[row + xb];
            e2 = cell_start[row + xb + 1];
          }
        }
      }
      nn_scan_packed(q, sorted, s1, max(e1 - s1, 0), s2, max(e2 - s2, 0), lane, sh, key);
    }
  }
  best = __uint_as_float((unsigned int)(key >> 32));
  bi = key < key0 ? (int)(unsigned int)key : -1;
}

// one thread per reference point of pair blockIdx-mapped pd: nearest neighbour + gates.
// The pair's "image" is 1 x N_ref; pairs[i] = (i, j) or (-1, -1). Threads take the queries in
// the pair's order (identity when order is null); every lane of a wavefront stays in the loop,
// as the wave search needs all of them.
extern "C" __global__ __launch_bounds__(kThreads) void k_nn_query(const PairDev *__restrict__ Pp,
                                                                 const FrameDev *__restrict__ F,
                                                                 const int *__restrict__ blk_pair,
                                                                 const GridDev *__restrict__ grids,
                                                                 const int *__restrict__ cell_arena,
                                                                 const float4 *__restrict__ sorted,
                                                                 const int *__restrict__ order, AlignParamsDev ap,
                                                                 float ns, int *__restrict__ nn_out) {
  const int gb = blockIdx.x;
  const int pi = blk_pair[gb];
  const PairDev &pd = Pp[pi];
  const GridDev g = grids[pi];
  const int *__restrict__ cell_start = cell_arena + g.cell_base;
  const FrameDev &fr = F[pd.ref];
  const FrameDev &fc = F[pd.cur];
  const int nref = *fr.npoints;
  const int base = (gb - pd.tile_base) * kTile;
  const int lane = threadIdx.x & 63;
  __shared__ float4 nn_sh[kThreads / 64][2][64];  // per wavefront: the points of one scan round
  float T[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) T[k] = *gview(pd.state + kSlotTinv + k);
  // sR = T.linear() * ns
  float sR[9];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) sR[c * 3 + r] = mc(T, r, c) * ns;
  const float r2 = ap.sq_thr;
  for (int kk = 0; kk < 4; ++kk) {
    const int j = base + kk * kThreads + threadIdx.x;
    const bool active = j < nref;
    const int i = active && order ? order[g.order_base + j] : j;
    float4 p = make_float4(0.0f, 0.0f, 0.0f, 1.0f), n = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (active) {
      p = fr.pts[i];
      n = fr.nrm[i];
    }
    float q[6], pt[3];
    iso_mul3(T, p.x, p.y, p.z, p.w, pt);
    q[0] = pt[0]; q[1] = pt[1]; q[2] = pt[2];
#pragma unroll
    for (int r = 0; r < 3; ++r) q[3 + r] = red3(sR[r] * n.x, sR[3 + r] * n.y, sR[6 + r] * n.z);
    // cell coordinates, held within +-2^20 (a farther query only searches more shells than it
    // needs: its true cell distances are larger than the held ones)
    constexpr int kFar = 1 << 20;
    const int cx = min(max(cell_coord(q[0], g.ox, g.inv_h), -kFar), kFar);
    const int cy = min(max(cell_coord(q[1], g.oy, g.inv_h), -kFar), kFar);
    const int cz = min(max(cell_coord(q[2], g.oz, g.inv_h), -kFar), kFar);
    float best = r2;
    int bi = -1;
    const int lx = wave_min_i(active ? cx : INT_MAX), hx = wave_max_i(active ? cx : INT_MIN);
    if (lx <= hx) {  // some lane has a query
      const int ly = wave_min_i(active ? cy : INT_MAX), hy = wave_max_i(active ? cy : INT_MIN);
      const int lz = wave_min_i(active ? cz : INT_MAX), hz = wave_max_i(active ? cz : INT_MIN);
      if (hx - lx < kNNWaveBox && hy - ly < kNNWaveBox && hz - lz < kNNWaveBox)
        nn_wave(q, lx, ly, lz, hx, hy, hz, g, cell_start, sorted, r2, active, lane, nn_sh[threadIdx.x >> 6], best, bi);
      else if (active)
        nn_walk(q, cx, cy, cz, g, cell_start, sorted, r2, best, bi);
    }
    if (!active) {
      if (j < pd.cols) pd.pairs[j] = make_int2(-1, -1);
      continue;
    }
    if (nn_out) nn_out[i] = bi;
    bool ok = bi >= 0;
    if (ok) {
      float rn[3], rp[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        rn[r] = red3(mc(T, r, 0) * n.x, mc(T, r, 1) * n.y, mc(T, r, 2) * n.z);
        rp[r] = red3(mc(T, r, 0) * p.x, mc(T, r, 1) * p.y, mc(T, r, 2) * p.z) + mc(T, r, 3);
      }
      const float4 cn = fc.nrm[bi], cp = fc.pts[bi];
      if (!ap.gicp && (red3(rn[0] * rn[0], rn[1] * rn[1], rn[2] * rn[2]) == 0.0f ||
                       red3(cn.x * cn.x, cn.y * cn.y, cn.z * cn.z) == 0.0f))
        ok = false;
      const float e0 = rp[0] - cp.x, e1 = rp[1] - cp.y, e2 = rp[2] - cp.z;
      if (ok && red3(e0 * e0, e1 * e1, e2 * e2) > r2) ok = false;
      if (ok && (isnan(cn.x) || isnan(cn.y) || isnan(cn.z) || isnan(rn[0]) || isnan(rn[1]) || isnan(rn[2]))) ok = false;
      if (ok && !ap.gicp && red3(rn[0] * cn.x, rn[1] * cn.y, rn[2] * cn.z) < ap.cos_thr) ok = false;
    }
    pd.pairs[i] = ok ? make_int2(i, bi) : make_int2(-1, -1);
  }
}

// DepthImage_scale (imageutils.cpp:6-37): block mean of step x step pixels over the
// non-zero ones, rejected (0) when the block's depth variance exceeds max_cov. The
// input is uint16 times scale (the reference scales the converted float image; the
// conversion is exact per pixel). One thread per output pixel.
extern "C" __global__ __launch_bounds__(kThreads) void k_depth_scale(const uint16_t *__restrict__ src, const float *__restrict__ srcf,
                                                                    int rows, int cols, float scale, int step,
                                                                    float max_cov, float *__restrict__ dst) {
  const int orows = rows / step, ocols = cols / step;
  const int o = blockIdx.x * kThreads + threadIdx.x;
  if (o >= orows * ocols) return;
  const int r = o / ocols, c = o - r * ocols;
  float acc = 0.0f, acc2 = 0.0f;
  int np = 0;
  const int sr = r * step, sc = c * step;
  for (int i = 0; i < step; ++i)
    for (int j = 0; j < step; ++j)
      if (sr + i < rows && sc + j < cols) {
        const size_t q = (size_t)(sr + i) * cols + sc + j;
        const float f = srcf ? srcf[q] : depth_of(src[q], scale);
        acc += f;
        acc2 += f * f;
        np += f > 0.0f;
      }
  float out = 0.0f;
  if (np) {
    const float mu = acc / (float)np;
    const float sigma = acc2 / (float)np - mu * mu;
    if (!(sigma > max_cov)) out = mu;
  }
  dst[o] = out;
}
