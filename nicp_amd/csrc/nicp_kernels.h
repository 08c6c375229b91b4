// Device-side descriptors shared by the kernels and the host launcher code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nicp_dev {

constexpr int kTile = 1024;          // pixels (or points) per per-pixel block
constexpr int kThreads = 256;        // threads per block (4 wavefronts of 64)
constexpr int kPartial = 32;         // doubles per block partial of the linearizer
constexpr int kVals = 31;            // used partial slots: H (21), b (6), chi2, inliers, count, both-valid pixels
constexpr int kSolveThreads = 256;   // threads per pair of k_solve (kSolveLanes per partial slot)
constexpr int kSolveLanes = 8;
constexpr uint64_t kZInit = ~0ull;   // empty z-buffer cell
// doubles of a pair's result record: the last in-loop linearisation's totals (kPartial), the
// statistics linearisation's totals (kPartial) and the solution statistics (48)
constexpr int kResultDoubles = 2 * kPartial + 48;

// per-pixel blocks of a pair: 1024-pixel tiles of its projector image
__host__ __device__ inline int pair_tiles(int rows, int cols) { return (rows * cols + kTile - 1) / kTile; }

// correspondence + linearisation blocks: kDTH x kDTW pixel tiles of a pair's projector image
// (of its sampled grid when the pair is subsampled); a wavefront covers one 64-pixel tile row
constexpr int kDTH = 16, kDTW = 64;
__host__ __device__ inline int dtiles_x(int cols) { return (cols + kDTW - 1) / kDTW; }
__host__ __device__ inline int dtiles_y(int rows) { return (rows + kDTH - 1) / kDTH; }

// Tile-binned projection (k_sbox, k_proj_corr_lin). Source tiles: kSTH x kSTW pixel blocks of a
// reference frame, each with the depth range of its valid pixels. Fused tiles: kFH x kDTW pixels
// of the projector image, kFH / kDTH correspondence tiles each, z-buffered in LDS; a fused tile
// lists up to a capacity of at most kListCap source tiles whose projection may reach it (more:
// every source tile of the reference frame is scanned).
constexpr int kSTH = 16, kSTW = 32;
constexpr int kBinMargin = 2;  // pixels added to each side of a projected source-tile box
constexpr int kPhaseB = kSTH * kSTW / 64;  // source pixels per lane of a wavefront's source tile
constexpr int kFH = 64;
constexpr int kListCap = 256;
constexpr int kDepthCopyBlocks = 32;  // blocks per frame of k_batch_depth_copy
constexpr int kCountTiles = 4;        // 1024-pixel tiles per block of k_count
constexpr int kSPosShift = 23;  // bit of a packed source pixel's offset in the tile (FrameDev::spix)
static_assert(kSTH * kSTW <= (1 << (32 - kSPosShift)), "a packed source pixel's offset fits above its index");
static_assert(kSTH * kSTW % 64 == 0 && kListCap <= kThreads, "k_proj_corr_lin's source-tile loop");
static_assert(kFH % kDTH == 0, "a fused tile holds whole correspondence tiles");
__host__ __device__ inline int stiles_x(int cols) { return (cols + kSTW - 1) / kSTW; }
__host__ __device__ inline int stiles(int rows, int cols) { return ((rows + kSTH - 1) / kSTH) * stiles_x(cols); }
__host__ __device__ inline int ftiles_y(int rows) { return (rows + kFH - 1) / kFH; }

// SphericalPointProjector parameters (sphericalpointprojector.h:304-313): half fields of view,
// image centre, angular resolutions in pixels per radian and their inverses.
struct SphDev {
  float hfov, vfov, hc, vc, hres, vres, ihres, ivres;
};
constexpr int kPinhole = 0, kSpherical = 1;

// One frame (a depth image and the cloud made from it). Every pointer is a
// device pointer; arrays indexed by pixel hold rows*cols entries, arrays
// indexed by point hold up to rows*cols entries (the capacity).
struct FrameDev {
  int rows, cols;
  int n_tiles;          // ceil(rows*cols / kTile)
  int tile_base;        // first global block of this frame in per-pixel launches
  int xblk_base;        // first block of the x-prefix pass (rows*10 threads, padded to kThreads)
  int yblk_base;        // first block of the y-prefix pass (cols*10 threads, padded to kThreads)
  float depth_scale;
  float min_d, max_d;
  float ip0, ip1;       // K * (worldRadius, worldRadius, 0): rows 0 and 1 (interval numerators)
  float K[9];           // camera matrix (column-major)
  float iKRt[16];       // unprojection matrix (pinhole) or the projector transform (spherical), identity pose
  int model;            // kPinhole or kSpherical
  SphDev sph;
  const uint16_t *depth;
  const float *depthf;  // float depth in metres (used instead of depth when set)
  int *index;           // index image
  float *ii;            // integral image, 10 floats per pixel
  float4 *pts;          // (x, y, z, 1)
  float4 *nrm;          // (nx, ny, nz, curvature)
  float4 *omg;          // 3 float4 per point: Omega_p (00 01 02 11 12 22), Omega_n (same)
  float4 *stats;        // 4 float4 per point or null: eigvecs (9), eigvals (3), mean (3), n
  float2 *stile_d;      // per source tile: (min, max) depth of its valid pixels, or null
  // per source tile, kSTH * kSTW slots: its valid pixels compacted in the tile's row-major order,
  // each as (depth as float bits, cloud index | pixel offset in the tile << kSPosShift), and
  // their count; or null (frames of more than 2^kSPosShift pixels have none)
  uint2 *spix;
  int *stile_n;
  int *npoints;         // device scalar
  int *tile_count;      // per tile
  int *tile_off;        // per tile
  // Set by the depth conversion when both information-matrix calculators' diagonals have equal
  // second and third entries: a point's information matrices are then b I + (a - b) n n^T in
  // the eigenbasis's closed form, from its normal n and curvature alone (the flat pair where the
  // curvature is below om_thr, else the non-flat pair; zero where the normal is zero), and the
  // linearisers compute them instead of loading omg. Cleared by anything that changes omg.
  int omega_closed;
  float om_thr[2];      // curvature thresholds: points, normals
  float om_a[4], om_b[4];  // (a, b) of: point flat, point non-flat, normal flat, normal non-flat
};

// One registration pair: reference and current frame ids, the projector
// image and the per-pair loop state.
struct PairDev {
  int ref, cur;
  int rows, cols;       // projector image size
  int tile_base, n_tiles;          // per-pixel blocks over the projector image
  int dtile_base, n_dtiles, dtiles_x;  // kDTH x kDTW tiles of the (sampled) projector image
  int ftile_base, n_ftiles;        // kFH x kDTW fused tiles of the projector image
  int sblk_base, n_sblk;           // k_sbox blocks over the reference frame's source tiles
  int4 *sbox;                      // per source tile of the reference frame: its fused-tile box
  int rtile_base, n_rtiles;        // blocks over the reference cloud capacity
  int ctile_base, n_ctiles;        // blocks over the current cloud capacity
  int cur_own;                     // 1: cur_index is the current frame's own index image
  int sub;                         // k_corr_lin evaluates every sub-th row and column (1: all)
  float K[9];
  float min_d, max_d;
  int model;                       // kPinhole or kSpherical
  SphDev sph;                      // spherical parameters at the pair's image size
  unsigned long long *zbuf;        // rows*cols z-buffer keys (depth bits << 32 | index)
  int *cur_index;                  // rows*cols
  int2 *pairs;                     // rows*cols stored correspondences (or (-1,-1))
  float *state;                    // kStateFloats: T, Tinv, KRt (reference), KRt (current), invT (inner),
                                   // reference sensor offset
  double *partials;                // max(n_tiles, n_dtiles) * kPartial
  double *result;                  // kResultDoubles: last linearisation, statistics linearisation, solution statistics
};

constexpr int kStateFloats = 96;
constexpr int kSlotT = 0, kSlotTinv = 16, kSlotKRt = 32, kSlotKRtCur = 48, kSlotInvTInner = 64, kSlotRefOffset = 80;

struct StatsParamsDev {
  int min_r, max_r, min_points;
  float curvature_threshold;
  float pflat[3], pnonflat[3], nflat[3], nnonflat[3];
  float pthr, nthr;
  int keep_stats;
  // set: the information matrices are not stored (the frames' closed form stands for them until
  // something asks for the stored arrays, which are then filled from the normals)
  int skip_omega;
  // > 0: pixels whose FP32 eigenvalues have a relative gap (l1 - l0) / l2 below this value are
  // solved again in FP64 (nicp_ctx_set_fp64_eigen_gap); 0 keeps the reference's FP32 solve
  float fp64_gap;
};

struct AlignParamsDev {
  float sq_thr, cos_thr, flat, min_ratio, max_ratio;
  int row_region, col_region, gicp;
  float max_chi2;
  int robust;
  float nscale;
  float lambda;
};

}  // namespace nicp_dev
