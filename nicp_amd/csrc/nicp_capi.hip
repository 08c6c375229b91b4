// Host side of the nicp_amd C ABI (include/nicp_amd.h): device memory of
// clouds and batches, launch sequences of the conversion and registration
// stages, hipGraph capture of the iteration chain, and the FP64 solution
// statistics of Aligner::_computeStatistics (nicp/nicp/aligner.cpp:82-135).
// Single translation unit: the kernels are compiled here too.
#include "nicp_kernels.hip"

#include <float.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/nicp_amd.h"

using namespace nicp_dev;
using namespace nicp_math;

// ---------------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------------
static thread_local char g_err[512] = "";

static int set_err(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) return set_err(NICP_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

#define CHK(x)                  \
  do {                          \
    int rc_ = (x);              \
    if (rc_ != NICP_OK) return rc_; \
  } while (0)

extern "C" const char *nicp_last_error(void) { return g_err; }

// ---------------------------------------------------------------------------------------
// device memory helpers
// ---------------------------------------------------------------------------------------
namespace {

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// a growable device buffer, or (fixed) a view of preallocated memory that never grows
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  bool fixed = false;
  int reserve(size_t n) {
    if (n <= bytes) return NICP_OK;
    if (fixed) return set_err(NICP_ERR_INVALID, "%zu bytes needed, %zu available in the fixed workspace", n, bytes);
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    size_t want = align_up(std::max<size_t>(n, 256), 1 << 16);
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      p = nullptr;
      return set_err(NICP_ERR_OOM, "hipMalloc(%zu): %s", want, hipGetErrorString(e));
    }
    bytes = want;
    return NICP_OK;
  }
  void release() {
    if (p && !fixed) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T *as() const { return (T *)p; }
};

// carves aligned sub-ranges out of one allocation
struct Carver {
  size_t off = 0;
  template <typename T>
  size_t take(size_t count) {
    size_t o = align_up(off, 256);
    off = o + count * sizeof(T);
    return o;
  }
};

}  // namespace

// ---------------------------------------------------------------------------------------
// the fixed workspace of nicp_align_noalloc: one statically sized array in device memory,
// which every GPU the library runs on holds until the process ends. Host code never passes
// the array's name to a kernel or a copy (on the host it is only a stand-in): it uses the
// device address hipGetSymbolAddress returns for the current device, kept per device.
// ---------------------------------------------------------------------------------------
constexpr int kNoAllocMaxPixels = 640 * 480;            // projector pixels and points per cloud
constexpr int kNoAllocMaxRows = 960, kNoAllocMaxCols = 1280;  // projector image sides
constexpr size_t kNoAllocScratchBytes = 7u << 20;        // 7 MiB
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {nullptr};

// the device address of `scratch` on `device` (which must be current), looked up once per device
int scratch_address(int device, void **out) {
  if (device < 0 || device >= kMaxDevices) return set_err(NICP_ERR_INVALID, "device %d beyond the workspace table", device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return set_err(NICP_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return NICP_OK;
}
}  // namespace


// ---------------------------------------------------------------------------------------
// per-kernel device timing with HIP events (recorded around each launch on the
// launching stream; inside a captured graph the records become event nodes)
// ---------------------------------------------------------------------------------------
enum KernelId {
  KID_COUNT, KID_SCAN, KID_UNPROJECT, KID_INTEGRAL_X, KID_INTEGRAL_Y, KID_STATS, KID_ZFILL, KID_ZPROJECT_CUR,
  KID_ZRESOLVE, KID_ZPROJECT_REF, KID_CORR_LIN, KID_SOLVE, KID_LIN_PAIRS, KID_SOLVE_STATS, KID_NN_QUERY,
  KID_NET_INPUT, KID_NET_FORWARD, KID_START_PROBE, KID_INTEGRAL_FRAME, KID_SOLUTION_STATS, KID_SRC_TILES, KID_BIN,
  KID_PROJ_CORR_LIN, KID_N
};
// the last three are stages: the network input kernel, the network's forward pass (its six
// GEMM launches) and the one-iteration probe of both starts of the learned-start aligner
static const char *kKernelNames[KID_N] = {"k_count", "k_scan", "k_unproject", "k_integral_x", "k_integral_y",
                                          "k_stats", "k_zfill", "k_zproject(current)", "k_zresolve",
                                          "k_zproject(reference)", "k_corr_lin", "k_solve", "k_lin_pairs",
                                          "k_solve(statistics)", "k_nn_query", "k_net_input", "initnet_forward",
                                          "start_probe", "k_integral_frame", "k_solution_stats", "k_src_tiles",
                                          "k_sbox", "k_proj_corr_lin"};

namespace {
struct ProfEntry {
  int id;
  hipEvent_t a, b;
};
struct Prof {
  bool on = false;
  hipError_t err = hipSuccess;     // first failure of a timing call (profiling then stops)
  const char *err_where = "";
  std::vector<hipEvent_t> pool;
  std::vector<ProfEntry> *sink = nullptr;  // where new entries go (eager list or a graph's list)
  std::vector<ProfEntry> eager;
  double total_ms[KID_N] = {0};
  long long launches[KID_N] = {0};
  void fail(hipError_t e, const char *where) {
    if (err == hipSuccess) {
      err = e;
      err_where = where;
    }
    on = false;
    (void)hipGetLastError();  // keep the failure out of the kernel-launch checks
  }
  hipEvent_t take() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    hipError_t r = hipEventCreateWithFlags(&e, hipEventDefault);
    if (r != hipSuccess) {
      fail(r, "hipEventCreateWithFlags");
      e = nullptr;
    }
    return e;
  }
  void begin(int id, hipStream_t s, ProfEntry &en) {
    en.id = id;
    en.a = take();
    en.b = take();
    if (en.a) {
      hipError_t r = hipEventRecord(en.a, s);
      if (r != hipSuccess) fail(r, "hipEventRecord(start)");
    }
  }
  void end(hipStream_t s, ProfEntry &en) {
    if (!en.a || !en.b) return;
    hipError_t r = hipEventRecord(en.b, s);
    if (r != hipSuccess) {
      fail(r, "hipEventRecord(end)");
      return;
    }
    if (sink) sink->push_back(en);
  }
  // reads finished entries; recycle = return their events to the pool
  void read(std::vector<ProfEntry> &list, bool recycle) {
    for (const ProfEntry &en : list) {
      float ms = 0.0f;
      hipError_t r = hipEventElapsedTime(&ms, en.a, en.b);
      if (r == hipSuccess) {
        total_ms[en.id] += ms;
        launches[en.id] += 1;
      } else {
        fail(r, "hipEventElapsedTime");
      }
      if (recycle) {
        pool.push_back(en.a);
        pool.push_back(en.b);
      }
    }
    if (recycle) list.clear();
  }
  void destroy_list(std::vector<ProfEntry> &list) {
    for (const ProfEntry &en : list) {
      (void)hipEventDestroy(en.a);
      (void)hipEventDestroy(en.b);
    }
    list.clear();
  }
  void release() {
    destroy_list(eager);
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    pool.clear();
  }
};
}  // namespace

#define LAUNCH(prof, id, kern, grid, block, stream, ...)                                   \
  do {                                                                                     \
    ProfEntry en_{};                                                                       \
    const bool p_ = (prof) != nullptr && (prof)->on;                                      \
    if (p_) (prof)->begin((id), (stream), en_);                                            \
    hipLaunchKernelGGL(kern, grid, block, 0, stream, __VA_ARGS__);                         \
    hipError_t le_ = hipGetLastError();                                                    \
    if (le_ != hipSuccess) return set_err(NICP_ERR_HIP, "launch %s: %s", #kern, hipGetErrorString(le_)); \
    if (p_) (prof)->end((stream), en_);                                                    \
  } while (0)
// ---------------------------------------------------------------------------------------
// objects
// ---------------------------------------------------------------------------------------
struct nicp_ctx {
  int device = 0;
  float fp64_eigen_gap = 0.0f;  // nicp_ctx_set_fp64_eigen_gap: applied to every conversion
  void *noalloc_ws = nullptr;  // device address of the fixed workspace on this device
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // workspace of single-pair calls
  DevBuf frames_tbl, pairs_tbl, maps, zbuf, cur_index, pairs_buf, state, partials, result, depth_tmp, scratch_in;
  // last single-pair correspondences (device pointer into pairs_buf) and its image size
  int last_rows = 0, last_cols = 0;
  bool last_valid = false;
  Prof prof;
  std::vector<std::vector<ProfEntry> *> pending_graphs;  // replayed graphs whose events are not yet read
  void graph_pending(std::vector<ProfEntry> *ev) {
    if (prof.on && std::find(pending_graphs.begin(), pending_graphs.end(), ev) == pending_graphs.end())
      pending_graphs.push_back(ev);
  }
  void read_pending() {
    for (std::vector<ProfEntry> *ev : pending_graphs) prof.read(*ev, false);
    pending_graphs.clear();
  }
  // kNN grid of the current cloud
  DevBuf grid_cells, grid_cursor, grid_sorted, grid_aux, nn_out, nn_order, grid_cells_c, grid_sorted_c;
  // allocations of destroyed clouds (size, pointer), reused for new clouds of the same size;
  // every user of them runs on `stream`, so a reuse is ordered after the last use
  std::vector<std::pair<size_t, void *>> cloud_pool;
};

struct nicp_cloud {
  nicp_ctx *ctx = nullptr;
  FrameDev fd{};          // host copy holding device pointers
  int n = 0;
  int has_index = 0;
  int has_stats = 0;
  void *alloc = nullptr;  // owned allocation (null for batch views)
  size_t alloc_bytes = 0; // its size when it may go back to the context's pool (else 0)
  int cap = 0;            // points the arrays hold
  float4 *gauss = nullptr;        // nicp_gaussian records (7 float4 each) or null
  void *gauss_alloc = nullptr;    // owned allocation of gauss when it is not inside alloc
};

// Pinned host staging of a batch's table uploads. Copies from pinned memory are asynchronous,
// so a call can enqueue its uploads and kernels while earlier work (the conversion) still runs.
// A call reserves all its bytes first (waiting for the previous call's copies, which read the
// same memory), puts its tables, and records `done` after its copies.
struct PinnedStage {
  char *p = nullptr;
  size_t cap = 0, used = 0;
  hipEvent_t done = nullptr;
  bool pending = false;
  static size_t slot(size_t n) { return align_up(n, 256); }
  int begin(size_t bytes) {
    if (pending) {
      HIPCHK(hipEventSynchronize(done));
      pending = false;
    }
    used = 0;
    if (bytes > cap) {
      if (p) (void)hipHostFree(p);
      p = nullptr;
      cap = 0;
      const size_t want = align_up(std::max<size_t>(bytes, 1 << 16), 1 << 16);
      HIPCHK(hipHostMalloc((void **)&p, want, hipHostMallocDefault));
      cap = want;
    }
    if (!done) HIPCHK(hipEventCreateWithFlags(&done, hipEventDisableTiming));
    return NICP_OK;
  }
  // a copy of n bytes of src in the staging memory (the caller reserved room with begin)
  const void *put(const void *src, size_t n) {
    char *dst = p + used;
    if (n) memcpy(dst, src, n);
    used += slot(n);
    return dst;
  }
  int end(hipStream_t s) {
    HIPCHK(hipEventRecord(done, s));
    pending = true;
    return NICP_OK;
  }
  void release() {
    if (done) (void)hipEventDestroy(done);
    if (p) (void)hipHostFree(p);
    p = nullptr;
    done = nullptr;
    cap = used = 0;
    pending = false;
  }
};

struct GraphCache {
  hipGraphExec_t exec = nullptr;
  std::vector<int> key;
  std::vector<ProfEntry> events;  // timing events recorded inside the graph
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    key.clear();
    for (const ProfEntry &en : events) {
      (void)hipEventDestroy(en.a);
      (void)hipEventDestroy(en.b);
    }
    events.clear();
  }
};

struct nicp_batch {
  nicp_ctx *ctx = nullptr;
  int num_frames = 0;
  std::vector<FrameDev> frames;
  std::vector<int> npoints_host;
  StatsParamsDev sp{};
  void *arena = nullptr;
  FrameDev *d_frames = nullptr;
  int *d_blk_frame = nullptr, *d_xblk_frame = nullptr, *d_yblk_frame = nullptr;
  // frames by pixel count, largest first: the order in which k_integral_frame takes them (one
  // workgroup per frame), so the largest frames of an uneven batch start first
  int *d_frame_order = nullptr;
  // the conversion skipped storing the information matrices (the frames' closed form stands for
  // them); nicp_batch_frame_cloud fills a frame's arrays from its normals before handing it out
  bool omega_deferred = false;
  std::vector<char> omega_filled;
  // first pixel of each frame in the packed depth arrays of nicp_batch_upload/download_depth_device
  long long *d_pix_base = nullptr;
  long long total_pix = 0;
  int total_tiles = 0, total_xblk = 0, total_yblk = 0;
  // align workspace
  DevBuf pairs_tbl, maps, zbuf, cur_index, pairs_buf, state, partials, result, sbox;
  GraphCache graph;
  GraphCache conv_graph;
  DevBuf synth_buf;  // render job tables of nicp_batch_render_synthetic
  // the learned aligner's start probe, prepared next to the refinement (its own tables, state
  // and results), and its starts, choices and flags (learned_aux)
  DevBuf pr_pairs_tbl, pr_maps, pr_zbuf, pr_cur_index, pr_pairs_buf, pr_state, pr_partials, pr_result, learned_aux;
  PinnedStage stage;  // table uploads of align calls
  GraphCache learned_graph;  // nicp_batch_align_learned: network, start probe, choice and refinement
  // projection strategy (nicp_batch_set_projection): -1 the autotuner's choice when it is on,
  // else the tile-binned projection where it applies; 0, 1, 2 forced
  int projection = -1;
  // set by the first nicp_batch_convert: before it the frames' index images, points and point
  // counts are uninitialised memory, which align calls and frame clouds would index with
  bool converted = false;
  int list_cap = kListCap;  // list capacity of the tile-binned projection (nicp_batch_set_bin_capacity)
  // autotuner of the reference projection (nicp_batch_set_autotune): per plan shape, the
  // strategy chosen from the timed iterations of the shape's first call
  int autotune = 0;
  std::vector<std::vector<int>> tuned_keys;
  std::vector<int> tuned_choice;
  std::vector<float> tuned_ms;  // kNumStrategies per shape: average ms per iteration (-1 unmeasured)
};

namespace {
// Replays `graph`, captured from what enqueue() launches on the context's stream unless its key
// differs from `key` (then it is captured again). Timing events recorded during the capture
// become nodes of the graph and are read after the replay.
template <typename F>
int graph_replay(nicp_ctx *ctx, GraphCache &graph, const std::vector<int> &key, F enqueue) {
  if (graph.exec == nullptr || graph.key != key) {
    graph.reset();
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->prof.sink = &graph.events;
    const int rc = enqueue();
    ctx->prof.sink = &ctx->prof.eager;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc != NICP_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return set_err(NICP_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&graph.exec, g, nullptr, nullptr, 0);
    hipError_t de = hipGraphDestroy(g);
    if (e == hipSuccess && de != hipSuccess) e = de;
    if (e != hipSuccess) {
      graph.exec = nullptr;
      return set_err(NICP_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    }
    graph.key = key;
  }
  HIPCHK(hipGraphLaunch(graph.exec, ctx->stream));
  ctx->graph_pending(&graph.events);
  return NICP_OK;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// parameter conversion
// ---------------------------------------------------------------------------------------
namespace {

int check_camera(const nicp_camera *cam) {
  if (!cam) return set_err(NICP_ERR_INVALID, "camera is null");
  if (!(cam->min_distance > 0.0f) || !(cam->max_distance >= cam->min_distance))
    return set_err(NICP_ERR_INVALID, "camera distances must satisfy 0 < min_distance <= max_distance");
  if (cam->model != NICP_PINHOLE && cam->model != NICP_SPHERICAL)
    return set_err(NICP_ERR_INVALID, "unknown projector model %d", cam->model);
  if (cam->model == NICP_SPHERICAL && !(cam->horizontal_fov > 0.0f && cam->vertical_fov > 0.0f))
    return set_err(NICP_ERR_INVALID, "spherical fields of view must be > 0");
  return NICP_OK;
}

// SphericalPointProjector::_updateParameters (sphericalpointprojector.cpp:19-27) at an image size
SphDev sph_params(const nicp_camera *cam, int rows, int cols) {
  SphDev s{};
  if (cam->model != NICP_SPHERICAL) return s;
  s.hfov = cam->horizontal_fov;
  s.vfov = cam->vertical_fov;
  s.hc = cam->horizontal_center;
  s.vc = cam->vertical_center;
  s.hres = (float)cols / (2.0f * s.hfov);
  s.ihres = 1.0f / s.hres;
  s.vres = (float)rows / (2.0f * s.vfov);
  s.ivres = 1.0f / s.vres;
  return s;
}

StatsParamsDev to_dev(const nicp_stats_params *s, int keep_stats) {
  StatsParamsDev d{};
  d.min_r = s->min_image_radius;
  d.max_r = s->max_image_radius;
  d.min_points = s->min_points;
  d.curvature_threshold = s->curvature_threshold;
  for (int i = 0; i < 3; ++i) {
    d.pflat[i] = s->point_flat[i];
    d.pnonflat[i] = s->point_nonflat[i];
    d.nflat[i] = s->normal_flat[i];
    d.nnonflat[i] = s->normal_nonflat[i];
  }
  d.pthr = s->point_curvature_threshold;
  d.nthr = s->normal_curvature_threshold;
  d.keep_stats = keep_stats;
  return d;
}

AlignParamsDev to_dev(const nicp_align_params *a) {
  AlignParamsDev d{};
  // CorrespondenceFinder::setInlierDistanceThreshold (correspondencefinder.h:86-89)
  d.sq_thr = a->inlier_distance_threshold * a->inlier_distance_threshold;
  d.cos_thr = a->inlier_normal_angular_threshold;
  d.flat = a->flat_curvature_threshold;
  // correspondencefinderprojective.cpp:33-34
  d.min_ratio = 1.0f / a->inlier_curvature_ratio_threshold;
  d.max_ratio = a->inlier_curvature_ratio_threshold;
  d.row_region = a->row_search_region;
  d.col_region = a->col_search_region;
  d.gicp = a->demoted_to_gicp;
  d.max_chi2 = a->inlier_max_chi2;
  d.robust = a->robust_kernel;
  d.nscale = a->normal_omega_scale;
  d.lambda = a->lambda;
  return d;
}

int check_align_params(const nicp_align_params *a) {
  if (!a) return set_err(NICP_ERR_INVALID, "align params are null");
  if (a->outer_iterations < 0 || a->inner_iterations < 1 || a->row_search_region < 0 || a->col_search_region < 0)
    return set_err(NICP_ERR_INVALID, "invalid iteration counts or search regions");
  return NICP_OK;
}

// fills the geometric members of a frame descriptor
// The closed form of the information matrices a conversion with these parameters writes
// (FrameDev::omega_closed): PointInformationMatrixCalculator and NormalInformationMatrixCalculator
// (informationmatrixcalculator.cpp:9-57) give U diag(d) U^T with U the window's eigenvectors,
// the normal first; with d1 == d2 that is d1 I + (d0 - d1) n n^T.
void set_omega_model(FrameDev &f, const StatsParamsDev &sp) {
  const float *d[4] = {sp.pflat, sp.pnonflat, sp.nflat, sp.nnonflat};
  f.omega_closed = 1;
  for (int i = 0; i < 4; ++i) {
    f.omega_closed = f.omega_closed && d[i][1] == d[i][2];
    f.om_a[i] = d[i][0];
    f.om_b[i] = d[i][1];
  }
  f.om_thr[0] = sp.pthr;
  f.om_thr[1] = sp.nthr;
}

void frame_geometry(FrameDev &f, int rows, int cols, const nicp_camera *cam, float world_radius, float depth_scale) {
  f.rows = rows;
  f.cols = cols;
  f.n_tiles = (rows * cols + kTile - 1) / kTile;
  f.depth_scale = depth_scale;
  f.min_d = cam->min_distance;
  f.max_d = cam->max_distance;
  for (int k = 0; k < 9; ++k) f.K[k] = cam->K[k];
  f.model = cam->model;
  f.sph = sph_params(cam, rows, cols);
  float I4[16], KRt[16];
  set_identity4(I4);
  if (cam->model == NICP_SPHERICAL) {
    // the converter's projector transform is the identity (depthimageconverterintegralimage.cpp:41);
    // the interval takes the world radius itself
    set_identity4(f.iKRt);
    f.ip0 = f.ip1 = world_radius;
    return;
  }
  projector_matrices(cam->K, I4, KRt, f.iKRt);
  // _projectInterval (pinholepointprojector.h:293-302): K * (R, R, 0), 3-term redux
  f.ip0 = red3(cam->K[0] * world_radius, cam->K[3] * world_radius, cam->K[6] * 0.0f);
  f.ip1 = red3(cam->K[1] * world_radius, cam->K[4] * world_radius, cam->K[7] * 0.0f);
}

int xblocks(int rows) { return (rows * 10 + kThreads - 1) / kThreads; }
int yblocks(int cols) { return (cols * 10 + kThreads - 1) / kThreads; }

// Whether a batch frame carries packed source-tile pixels: pinhole frames whose cloud indices
// fit below the pixel offset of a packed entry (kSPosShift bits)
bool spack_frame(int model, int rows, int cols) {
  return model == NICP_PINHOLE && (long long)rows * cols <= (1ll << kSPosShift);
}

// Layout of the per-frame device arrays inside one allocation.
struct FrameLayout {
  size_t depth, depthf, index, ii, pts, nrm, omg, stats, stile, spix, stn, npoints, tcount, toff, total;
};
// with_spack: the packed source-tile pixels of the tile-binned projection (batch frames for which
// spack_frame holds)
FrameLayout frame_layout(int rows, int cols, bool with_stats, bool with_ii, size_t base, bool with_f32 = false,
                         bool with_spack = false) {
  Carver cv;
  cv.off = base;
  const size_t P = (size_t)rows * cols;
  const size_t T = (P + kTile - 1) / kTile;
  FrameLayout L;
  L.depth = cv.take<uint16_t>(P);
  L.depthf = with_f32 ? cv.take<float>(P) : 0;
  L.index = cv.take<int>(P);
  L.ii = with_ii ? cv.take<float>(P * 10) : 0;
  L.pts = cv.take<float4>(P);
  L.nrm = cv.take<float4>(P);
  L.omg = cv.take<float4>(P * 3);
  L.stats = with_stats ? cv.take<float4>(P * 4) : 0;
  L.stile = with_ii ? cv.take<float2>((size_t)stiles(rows, cols)) : 0;
  L.spix = with_ii && with_spack ? cv.take<uint2>((size_t)stiles(rows, cols) * kSTH * kSTW) : 0;
  L.stn = with_ii && with_spack ? cv.take<int>((size_t)stiles(rows, cols)) : 0;
  L.npoints = cv.take<int>(1);
  L.tcount = cv.take<int>(T);
  L.toff = cv.take<int>(T);
  L.total = cv.off;
  return L;
}

void frame_pointers(FrameDev &f, char *base, const FrameLayout &L, bool with_stats, bool with_ii,
                    bool with_f32 = false, bool with_spack = false) {
  f.depth = (const uint16_t *)(base + L.depth);
  f.depthf = with_f32 ? (const float *)(base + L.depthf) : nullptr;
  f.index = (int *)(base + L.index);
  f.ii = with_ii ? (float *)(base + L.ii) : nullptr;
  f.pts = (float4 *)(base + L.pts);
  f.nrm = (float4 *)(base + L.nrm);
  f.omg = (float4 *)(base + L.omg);
  f.stats = with_stats ? (float4 *)(base + L.stats) : nullptr;
  // source-tile depth ranges of the tile-binned projection (pinhole frames)
  f.stile_d = with_ii && f.model == NICP_PINHOLE ? (float2 *)(base + L.stile) : nullptr;
  const bool spack = with_ii && with_spack && spack_frame(f.model, f.rows, f.cols);
  f.spix = spack ? (uint2 *)(base + L.spix) : nullptr;
  f.stile_n = spack ? (int *)(base + L.stn) : nullptr;
  f.npoints = (int *)(base + L.npoints);
  f.tile_count = (int *)(base + L.tcount);
  f.tile_off = (int *)(base + L.toff);
}

// Frames per conversion from which both prefix passes run as one workgroup per frame
// (k_integral_frame, the same values): with fewer frames the per-row and per-column
// kernels spread one frame over more compute units.
constexpr int kFusedIntegralFrames = 16;

// Launch sequence of the conversion for a frame table already on the device.
// All frames of one launch share the projector model.
int launch_convert(Prof *pf, hipStream_t s, const FrameDev *d_frames, int num_frames, const int *d_blk, int total_tiles,
                   const int *d_x, int total_x, const int *d_y, int total_y, const StatsParamsDev &sp, int model,
                   const int *d_order = nullptr) {
  if (total_tiles == 0) return NICP_OK;
  (void)hipGetLastError();
  const bool sph = model == NICP_SPHERICAL;
  const dim3 count_grid((total_tiles + kCountTiles - 1) / kCountTiles);
  if (sph) LAUNCH(pf, KID_COUNT, k_count_sph, count_grid, dim3(kThreads), s, d_frames, d_blk, total_tiles);
  else LAUNCH(pf, KID_COUNT, k_count, count_grid, dim3(kThreads), s, d_frames, d_blk, total_tiles);
  LAUNCH(pf, KID_SCAN, k_scan, dim3(num_frames), dim3(1024), s, d_frames);
  if (sph) LAUNCH(pf, KID_UNPROJECT, k_unproject_sph, dim3(total_tiles), dim3(kThreads), s, d_frames, d_blk);
  else LAUNCH(pf, KID_UNPROJECT, k_unproject, dim3(total_tiles), dim3(kThreads), s, d_frames, d_blk);
  if (!sph) LAUNCH(pf, KID_SRC_TILES, k_src_tiles, dim3(total_tiles), dim3(kThreads), s, d_frames, d_blk);
  if (num_frames >= kFusedIntegralFrames) {
    if (sph)
      LAUNCH(pf, KID_INTEGRAL_FRAME, k_integral_frame_sph, dim3(num_frames), dim3(kIFrameThreads), s, d_frames, d_order);
    else LAUNCH(pf, KID_INTEGRAL_FRAME, k_integral_frame, dim3(num_frames), dim3(kIFrameThreads), s, d_frames, d_order);
  } else {
    if (sph) LAUNCH(pf, KID_INTEGRAL_X, k_integral_x_sph, dim3(total_x), dim3(kThreads), s, d_frames, d_x);
    else LAUNCH(pf, KID_INTEGRAL_X, k_integral_x, dim3(total_x), dim3(kThreads), s, d_frames, d_x);
    LAUNCH(pf, KID_INTEGRAL_Y, k_integral_y, dim3(total_y), dim3(kThreads), s, d_frames, d_y);
  }
  if (sph) LAUNCH(pf, KID_STATS, k_stats_sph, dim3(total_tiles), dim3(kThreads), s, d_frames, d_blk, sp);
  else LAUNCH(pf, KID_STATS, k_stats, dim3(total_tiles), dim3(kThreads), s, d_frames, d_blk, sp);
  HIPCHK(hipGetLastError());
  return NICP_OK;
}

// ---------------------------------------------------------------------------------------
// registration workspace shared by single-pair and batched calls
// ---------------------------------------------------------------------------------------
struct PairPlan {
  std::vector<PairDev> pairs;
  std::vector<int> blk_pair, blk_ref, blk_cur, blk_dpair, blk_spair, blk_fpair;  // block maps
  int total_pix_tiles = 0, total_rtiles = 0, total_ctiles = 0, total_dtiles = 0, total_sblk = 0, total_ftiles = 0;
  size_t total_pix = 0;
  size_t total_parts = 0;       // partial slots: max(n_tiles, n_dtiles) per pair
  size_t total_sbox = 0;        // source-tile boxes (int4) of all pairs' reference frames
  std::vector<int> n_stiles;    // source tiles of each pair's reference frame
  // device copies of blk_dpair, blk_spair, blk_fpair (set by prepare_align)
  int *d_dpair = nullptr, *d_spair = nullptr, *d_fpair = nullptr;
  // the tile-binned projection applies: pinhole pairs, every current index image the frame's own,
  // every reference frame with source-tile depth ranges, no subsampling (set by batch_align)
  bool fused_ok = false;
  int list_cap = kListCap;      // source tiles a fused tile lists before it scans them all
  // the aligner's sensor offsets (null: identity), the same for every pair of the plan
  const float *ref_offset = nullptr, *cur_offset = nullptr;
  // > 0: pairs q and q + twin share both frames (the two starts of the start probe), and
  // one k_zproject_twin pass over the first half's reference tiles projects both
  int twin = 0;
};

// the per-pixel block maps of every pair from its projector image and sampling: linear
// 1024-pixel tiles (k_zfill, k_zresolve, k_lin_pairs) and kDTH x kDTW tiles (k_corr_lin) of
// the sampled grid
void plan_pixel_tiles(PairPlan &plan) {
  plan.blk_pair.clear();
  plan.blk_dpair.clear();
  plan.blk_fpair.clear();
  plan.total_parts = 0;
  for (size_t i = 0; i < plan.pairs.size(); ++i) {
    PairDev &pd = plan.pairs[i];
    const int GR = (pd.rows + pd.sub - 1) / pd.sub, GC = (pd.cols + pd.sub - 1) / pd.sub;
    pd.n_tiles = pair_tiles(GR, GC);
    pd.tile_base = (int)plan.blk_pair.size();
    plan.blk_pair.insert(plan.blk_pair.end(), pd.n_tiles, (int)i);
    pd.dtiles_x = dtiles_x(GC);
    pd.n_dtiles = dtiles_y(GR) * pd.dtiles_x;
    pd.dtile_base = (int)plan.blk_dpair.size();
    plan.blk_dpair.insert(plan.blk_dpair.end(), pd.n_dtiles, (int)i);
    plan.total_parts += (size_t)std::max(pd.n_tiles, pd.n_dtiles);
    pd.n_ftiles = ftiles_y(GR) * pd.dtiles_x;
    pd.ftile_base = (int)plan.blk_fpair.size();
    plan.blk_fpair.insert(plan.blk_fpair.end(), pd.n_ftiles, (int)i);
  }
  plan.total_pix_tiles = (int)plan.blk_pair.size();
  plan.total_dtiles = (int)plan.blk_dpair.size();
  plan.total_ftiles = (int)plan.blk_fpair.size();
}

// hosts frames[] (host copies) and pair list -> plan with block maps and buffer offsets
void make_plan(const std::vector<FrameDev> &frames, int num_pairs, const int *ref, const int *cur,
               const nicp_camera *const *cams, PairPlan &plan) {
  plan.pairs.resize(num_pairs);
  plan.blk_pair.clear();
  plan.blk_ref.clear();
  plan.blk_cur.clear();
  plan.blk_spair.clear();
  plan.n_stiles.clear();
  plan.total_sbox = 0;
  size_t pix = 0;
  for (int i = 0; i < num_pairs; ++i) {
    PairDev &pd = plan.pairs[i];
    const FrameDev &fr = frames[ref[i]], &fc = frames[cur[i]];
    pd.ref = ref[i];
    pd.cur = cur[i];
    pd.rows = fc.rows;
    pd.cols = fc.cols;
    const int P = pd.rows * pd.cols;
    pd.n_rtiles = (fr.rows * fr.cols + kTile - 1) / kTile;
    pd.rtile_base = (int)plan.blk_ref.size();
    for (int t = 0; t < pd.n_rtiles; ++t) plan.blk_ref.push_back(i);
    pd.n_ctiles = (fc.rows * fc.cols + kTile - 1) / kTile;
    pd.ctile_base = (int)plan.blk_cur.size();
    for (int t = 0; t < pd.n_ctiles; ++t) plan.blk_cur.push_back(i);
    const int ns = stiles(fr.rows, fr.cols);
    pd.n_sblk = (ns + kThreads - 1) / kThreads;
    pd.sblk_base = (int)plan.blk_spair.size();
    plan.blk_spair.insert(plan.blk_spair.end(), pd.n_sblk, i);
    // box offset of this pair, turned into a pointer by prepare_align
    pd.sbox = (int4 *)(uintptr_t)plan.total_sbox;
    plan.total_sbox += (size_t)ns;
    plan.n_stiles.push_back(ns);
    for (int k = 0; k < 9; ++k) pd.K[k] = cams[i]->K[k];
    pd.min_d = cams[i]->min_distance;
    pd.max_d = cams[i]->max_distance;
    pd.model = cams[i]->model;
    pd.sph = sph_params(cams[i], pd.rows, pd.cols);
    pd.cur_own = 0;
    pd.sub = 1;
    // pixel offset of this pair (aligned to 64 entries), turned into pointers by prepare_align
    pd.zbuf = (unsigned long long *)(uintptr_t)pix;
    pix += align_up((size_t)P, 64);
  }
  plan_pixel_tiles(plan);
  plan.total_sblk = (int)plan.blk_spair.size();
  plan.total_rtiles = (int)plan.blk_ref.size();
  plan.total_ctiles = (int)plan.blk_cur.size();
  plan.total_pix = pix;
}

// Batch frames all carry their index images and every pair's projector is its current
// frame's camera and image (nicp_batch_align). Then the current cloud projected at the
// identity lands every point on its own pixel, so the current index image of
// AlignerProjective::align (alignerprojective.cpp:44-47) is the frame's own index image.
// A spherical projection does not return every point to its own pixel (azimuths are truncated),
// so spherical pairs keep projecting their current cloud.
void plan_batch_frames(const std::vector<FrameDev> &frames, PairPlan &plan) {
  plan.blk_cur.clear();
  for (size_t i = 0; i < plan.pairs.size(); ++i) {
    PairDev &pd = plan.pairs[i];
    if (pd.model == NICP_SPHERICAL) {
      pd.ctile_base = (int)plan.blk_cur.size();
      for (int t = 0; t < pd.n_ctiles; ++t) plan.blk_cur.push_back((int)i);
      continue;
    }
    pd.cur_own = 1;
    pd.cur_index = frames[pd.cur].index;
    pd.ctile_base = 0;
    pd.n_ctiles = 0;
  }
  plan.total_ctiles = (int)plan.blk_cur.size();
}

struct AlignBuffers {
  DevBuf *pairs_tbl, *maps, *zbuf, *cur_index, *pairs_buf, *state, *partials, *result;
  DevBuf *sbox = nullptr;  // source-tile boxes of the tile-binned projection (batches)
};

// allocates/assigns the per-pair buffers and uploads tables and the initial state
// The staged bytes prepare_align uploads for a plan (through a PinnedStage)
size_t prepare_stage_bytes(const PairPlan &plan) {
  const size_t np = plan.pairs.size();
  const size_t nb = plan.blk_pair.size() + plan.blk_ref.size() + plan.blk_cur.size() + plan.blk_dpair.size() +
                    plan.blk_spair.size() + plan.blk_fpair.size();
  return PinnedStage::slot(np * sizeof(PairDev)) + PinnedStage::slot(nb * sizeof(int)) +
         PinnedStage::slot(np * kStateFloats * sizeof(float));
}

int prepare_align(hipStream_t s, PairPlan &plan, AlignBuffers &B, const float *init16, int **d_blk_pair,
                  int **d_blk_ref, int **d_blk_cur, PairDev **d_pairs, PinnedStage *stage = nullptr) {
  const int np = (int)plan.pairs.size();
  CHK(B.zbuf->reserve(plan.total_pix * sizeof(unsigned long long)));
  CHK(B.cur_index->reserve(plan.total_pix * sizeof(int)));
  CHK(B.pairs_buf->reserve(plan.total_pix * sizeof(int2)));
  CHK(B.state->reserve((size_t)np * kStateFloats * sizeof(float)));
  CHK(B.partials->reserve(std::max<size_t>(plan.total_parts, 1) * kPartial * sizeof(double)));
  CHK(B.result->reserve((size_t)np * kResultDoubles * sizeof(double)));
  if (B.sbox && plan.fused_ok) CHK(B.sbox->reserve(std::max<size_t>(plan.total_sbox, 1) * sizeof(int4)));
  size_t tile_part = 0;
  std::vector<float> st((size_t)np * kStateFloats, 0.0f);
  for (int i = 0; i < np; ++i) {
    PairDev &pd = plan.pairs[i];
    size_t off = (size_t)(uintptr_t)pd.zbuf;
    pd.zbuf = B.zbuf->as<unsigned long long>() + off;
    if (!pd.cur_own) pd.cur_index = B.cur_index->as<int>() + off;
    pd.pairs = B.pairs_buf->as<int2>() + off;
    pd.state = B.state->as<float>() + (size_t)i * kStateFloats;
    pd.partials = B.partials->as<double>() + tile_part * kPartial;
    tile_part += (size_t)std::max(pd.n_tiles, pd.n_dtiles);
    pd.result = B.result->as<double>() + (size_t)i * kResultDoubles;
    if (B.sbox && plan.fused_ok) {
      // the pair's boxes: one per source tile of its reference frame, inside the buffer
      const size_t off = (size_t)(uintptr_t)pd.sbox;
      if (off + (size_t)plan.n_stiles[i] > plan.total_sbox || plan.total_sbox * sizeof(int4) > B.sbox->bytes)
        return set_err(NICP_ERR_STATE, "source-tile boxes of pair %d outside their buffer", i);
      pd.sbox = B.sbox->as<int4>() + off;
    } else {
      pd.sbox = nullptr;
    }
    // initial state: AlignerProjective::align, alignerprojective.cpp:36-54; the sensor offsets
    // with their last rows set to 0 0 0 1 (aligner.h:177,194)
    float *S = &st[(size_t)i * kStateFloats];
    float T[16], RO[16], CO[16];
    for (int k = 0; k < 16; ++k) T[k] = init16 ? init16[(size_t)i * 16 + k] : (k % 5 == 0 ? 1.0f : 0.0f);
    set_row3(T);
    set_identity4(RO);
    set_identity4(CO);
    if (plan.ref_offset) memcpy(RO, plan.ref_offset, sizeof(RO));
    if (plan.cur_offset) memcpy(CO, plan.cur_offset, sizeof(CO));
    set_row3(RO);
    set_row3(CO);
    float Ti[16], TR[16], KRt[16];
    iso_inverse(T, Ti);
    iso_mul(T, RO, TR);
    projector_state(pd.model, pd.K, TR, KRt);
    for (int k = 0; k < 16; ++k) {
      S[kSlotT + k] = T[k];
      S[kSlotTinv + k] = Ti[k];
      S[kSlotKRt + k] = KRt[k];
      S[kSlotRefOffset + k] = RO[k];
    }
    projector_state(pd.model, pd.K, CO, KRt);
    for (int k = 0; k < 16; ++k) S[kSlotKRtCur + k] = KRt[k];
  }
  const size_t nb = plan.blk_pair.size() + plan.blk_ref.size() + plan.blk_cur.size() + plan.blk_dpair.size() +
                    plan.blk_spair.size() + plan.blk_fpair.size();
  CHK(B.pairs_tbl->reserve((size_t)np * sizeof(PairDev)));
  CHK(B.maps->reserve(std::max<size_t>(nb, 1) * sizeof(int)));
  std::vector<int> maps;
  maps.reserve(nb);
  maps.insert(maps.end(), plan.blk_pair.begin(), plan.blk_pair.end());
  maps.insert(maps.end(), plan.blk_ref.begin(), plan.blk_ref.end());
  maps.insert(maps.end(), plan.blk_cur.begin(), plan.blk_cur.end());
  maps.insert(maps.end(), plan.blk_dpair.begin(), plan.blk_dpair.end());
  maps.insert(maps.end(), plan.blk_spair.begin(), plan.blk_spair.end());
  maps.insert(maps.end(), plan.blk_fpair.begin(), plan.blk_fpair.end());
  *d_blk_pair = B.maps->as<int>();
  *d_blk_ref = *d_blk_pair + plan.blk_pair.size();
  *d_blk_cur = *d_blk_ref + plan.blk_ref.size();
  plan.d_dpair = *d_blk_cur + plan.blk_cur.size();
  plan.d_spair = plan.d_dpair + plan.blk_dpair.size();
  plan.d_fpair = plan.d_spair + plan.blk_spair.size();
  *d_pairs = B.pairs_tbl->as<PairDev>();
  const void *src_pairs = plan.pairs.data(), *src_maps = maps.data(), *src_st = st.data();
  if (stage) {
    src_pairs = stage->put(plan.pairs.data(), (size_t)np * sizeof(PairDev));
    src_maps = stage->put(maps.data(), nb * sizeof(int));
    src_st = stage->put(st.data(), st.size() * sizeof(float));
  }
  HIPCHK(hipMemcpyAsync(*d_pairs, src_pairs, (size_t)np * sizeof(PairDev), hipMemcpyHostToDevice, s));
  if (nb) HIPCHK(hipMemcpyAsync(*d_blk_pair, src_maps, nb * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(B.state->p, src_st, st.size() * sizeof(float), hipMemcpyHostToDevice, s));
  return NICP_OK;
}

// Strategies of the reference projection and correspondence search of an iteration:
//   0  k_zproject (merged 64-bit atomics into the global z-buffer) + k_corr_lin;
//   1  the same, reading each cell before its atomic;
//   2  tile-binned: k_bin + k_proj_corr_lin (the z-buffer of each fused tile in LDS).
// All three give the same z-buffer, correspondences and partial sums.
constexpr int kNumStrategies = 3;
constexpr int kFusedMinTiles = 1024;  // fused tiles from which strategy 2 is the default

// Timing of a tuning call: iteration i uses strategy i % (number of applicable strategies) and
// records its projection and correspondence launches between two events.
struct ZTune {
  std::vector<hipEvent_t> ev;  // 2 per outer iteration
  std::vector<int> strategy;
  void release() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    ev.clear();
    strategy.clear();
  }
};

// the launch sequence of AlignerProjective::align for all pairs of a plan
int launch_align(Prof *pf, hipStream_t s, const PairPlan &plan, const PairDev *d_pairs, const FrameDev *d_frames,
                 const int *d_blk_pair, const int *d_blk_ref, const int *d_blk_cur, const nicp_align_params *ap,
                 bool with_stats = true, int zstrategy = 0, ZTune *tune = nullptr) {
  const int np = (int)plan.pairs.size();
  const AlignParamsDev a = to_dev(ap);
  const dim3 tb(kThreads);
  if (plan.total_pix_tiles == 0) return NICP_OK;
  // one projector model per plan (a batch's frames share it)
  const bool sph = plan.pairs[0].model == NICP_SPHERICAL;
  const bool fused_ok = plan.fused_ok && !sph && a.row_region == 0 && a.col_region == 0;
  const int nstrat = fused_ok ? 3 : 2;
  if (!fused_ok && zstrategy == 2) zstrategy = 0;
  const bool only_fused = fused_ok && !tune && zstrategy == 2;
  (void)hipGetLastError();
  if (!only_fused || plan.total_ctiles) LAUNCH(pf, KID_ZFILL, k_zfill, dim3(plan.total_pix_tiles), tb, s, d_pairs, d_blk_pair);
  if (plan.total_ctiles) {
    if (sph) LAUNCH(pf, KID_ZPROJECT_CUR, k_zproject_sph, dim3(plan.total_ctiles), tb, s, d_pairs, d_frames, d_blk_cur, 1, zstrategy);
    else LAUNCH(pf, KID_ZPROJECT_CUR, k_zproject, dim3(plan.total_ctiles), tb, s, d_pairs, d_frames, d_blk_cur, 1, zstrategy);
    LAUNCH(pf, KID_ZRESOLVE, k_zresolve, dim3(plan.total_pix_tiles), tb, s, d_pairs, d_blk_pair, (float *)nullptr);
  }
  const int inner = ap->inner_iterations;
  for (int it = 0; it < ap->outer_iterations; ++it) {
    const bool last = it == ap->outer_iterations - 1;
    const int store = (int)(last || inner > 1);
    const int zs = tune ? it % nstrat : zstrategy;
    if (tune) {
      hipEvent_t e0 = nullptr, e1 = nullptr;
      HIPCHK(hipEventCreate(&e0));
      tune->ev.push_back(e0);
      HIPCHK(hipEventCreate(&e1));
      tune->ev.push_back(e1);
      tune->strategy.push_back(zs);
      HIPCHK(hipEventRecord(e0, s));
    }
    if (zs == 2) {
      if (plan.total_sblk) LAUNCH(pf, KID_BIN, k_sbox, dim3(plan.total_sblk), tb, s, d_pairs, d_frames, plan.d_spair);
      LAUNCH(pf, KID_PROJ_CORR_LIN, k_proj_corr_lin, dim3(plan.total_ftiles), tb, s, d_pairs, d_frames, plan.d_fpair, a,
             store, plan.list_cap);
    } else {
      if (plan.total_rtiles) {
        if (sph)
          LAUNCH(pf, KID_ZPROJECT_REF, k_zproject_sph, dim3(plan.total_rtiles), tb, s, d_pairs, d_frames, d_blk_ref, 0, zs);
        else if (plan.twin > 0)
          LAUNCH(pf, KID_ZPROJECT_REF, k_zproject_twin, dim3(plan.pairs[plan.twin].rtile_base), tb, s, d_pairs, d_frames,
                 d_blk_ref, plan.twin);
        else LAUNCH(pf, KID_ZPROJECT_REF, k_zproject, dim3(plan.total_rtiles), tb, s, d_pairs, d_frames, d_blk_ref, 0, zs);
      }
      LAUNCH(pf, KID_CORR_LIN, k_corr_lin, dim3(plan.total_dtiles), tb, s, d_pairs, d_frames, plan.d_dpair, a, store);
    }
    if (tune) HIPCHK(hipEventRecord(tune->ev.back(), s));
    LAUNCH(pf, KID_SOLVE, k_solve, dim3(np), dim3(kSolveThreads), s, d_pairs, ap->lambda, inner > 1 ? 1 : 0, 0, 1);
    for (int k = 1; k < inner; ++k) {
      LAUNCH(pf, KID_LIN_PAIRS, k_lin_pairs, dim3(plan.total_pix_tiles), tb, s, d_pairs, d_frames, d_blk_pair, a,
             kSlotInvTInner, 0);
      LAUNCH(pf, KID_SOLVE, k_solve, dim3(np), dim3(kSolveThreads), s, d_pairs, ap->lambda, k == inner - 1 ? 0 : 1, 1,
             0);
    }
  }
  // Aligner::_computeStatistics: one more linearisation at the final T over the last correspondences
  if (ap->outer_iterations > 0 && with_stats) {
    LAUNCH(pf, KID_LIN_PAIRS, k_lin_pairs, dim3(plan.total_pix_tiles), tb, s, d_pairs, d_frames, d_blk_pair, a,
           kSlotTinv, 0);
    LAUNCH(pf, KID_SOLVE_STATS, k_solve, dim3(np), dim3(kSolveThreads), s, d_pairs, 0.0f, 2, 0, 0);
    LAUNCH(pf, KID_SOLUTION_STATS, k_solution_stats, dim3(np), dim3(64), s, d_pairs,
           ap->translational_min_eigen_ratio, ap->rotational_min_eigen_ratio);
  }
  HIPCHK(hipGetLastError());
  return NICP_OK;
}

// r: a pair's result record (kResultDoubles): the last in-loop linearisation, the statistics
// linearisation and the solution statistics k_solution_stats computed on the device
void fill_result(const double *r, const nicp_align_params *ap, nicp_align_result *out, float ms) {
  double H[36];
  int e = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 6; ++j) {
      H[j * 6 + i] = H[i * 6 + j] = r[e];
      ++e;
    }
  memcpy(out->H, H, sizeof(H));
  for (int i = 0; i < 6; ++i) out->b[i] = r[21 + i];
  out->error = (float)r[27];
  out->inliers = (int)llround(r[28]);
  out->num_correspondences = (int)llround(r[29]);
  out->overlap_pixels = (int)llround(r[30]);
  out->total_time_ms = ms;
  if (ap->outer_iterations > 0) {
    const double *st = r + 2 * kPartial;
    memcpy(out->omega, st + nicp_stats::kSolOmega, sizeof(out->omega));
    memcpy(out->mean, st + nicp_stats::kSolMean, sizeof(out->mean));
    out->translational_ratio = st[nicp_stats::kSolTrans];
    out->rotational_ratio = st[nicp_stats::kSolRot];
    out->solution_valid = st[nicp_stats::kSolValid] != 0.0;
  } else {
    memset(out->mean, 0, sizeof(out->mean));
    memset(out->omega, 0, sizeof(out->omega));
    out->translational_ratio = out->rotational_ratio = INFINITY;
    out->solution_valid = 0;
  }
}

}  // namespace

// ---------------------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------------------
extern "C" int nicp_ctx_create(int device, nicp_ctx **out) {
  if (!out) return set_err(NICP_ERR_INVALID, "out is null");
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return set_err(NICP_ERR_INVALID, "device %d out of range (%d devices)", device, n);
  HIPCHK(hipSetDevice(device));
  nicp_ctx *c = new (std::nothrow) nicp_ctx();
  if (!c) return set_err(NICP_ERR_OOM, "out of host memory");
  c->device = device;
  {
    const int rc = scratch_address(device, &c->noalloc_ws);
    if (rc != NICP_OK) {
      delete c;
      return rc;
    }
  }
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e != hipSuccess) {
    delete c;
    return set_err(NICP_ERR_HIP, "stream/event creation: %s", hipGetErrorString(e));
  }
  *out = c;
  return NICP_OK;
}

extern "C" int nicp_ctx_destroy(nicp_ctx *c) {
  if (!c) return NICP_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  DevBuf *bufs[] = {&c->frames_tbl, &c->pairs_tbl, &c->maps, &c->zbuf, &c->cur_index, &c->pairs_buf,
                    &c->state, &c->partials, &c->result, &c->depth_tmp, &c->scratch_in, &c->grid_cells,
                    &c->grid_cursor, &c->grid_sorted, &c->grid_aux, &c->nn_out, &c->nn_order, &c->grid_cells_c,
                    &c->grid_sorted_c};
  for (DevBuf *b : bufs) b->release();
  for (auto &pe : c->cloud_pool) (void)hipFree(pe.second);
  c->cloud_pool.clear();
  c->prof.release();
  (void)hipEventDestroy(c->ev0);
  (void)hipEventDestroy(c->ev1);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return NICP_OK;
}

extern "C" int nicp_ctx_synchronize(nicp_ctx *c) {
  if (!c) return set_err(NICP_ERR_INVALID, "ctx is null");
  HIPCHK(hipStreamSynchronize(c->stream));
  return NICP_OK;
}

extern "C" void *nicp_ctx_stream(nicp_ctx *c) { return c ? (void *)c->stream : nullptr; }

extern "C" int nicp_ctx_set_fp64_eigen_gap(nicp_ctx *c, float gap) {
  if (!c) return set_err(NICP_ERR_INVALID, "ctx is null");
  if (!(gap >= 0.0f && gap <= 1.0f)) return set_err(NICP_ERR_INVALID, "the FP64 eigen gap must lie in [0, 1]");
  c->fp64_eigen_gap = gap;
  return NICP_OK;
}

// plain device buffers for callers that keep data resident (datasets, network inputs)
extern "C" int nicp_device_alloc(nicp_ctx *c, size_t bytes, void **ptr) {
  if (!c || !ptr) return set_err(NICP_ERR_INVALID, "bad arguments");
  *ptr = nullptr;
  HIPCHK(hipSetDevice(c->device));
  hipError_t e = hipMalloc(ptr, std::max<size_t>(bytes, 256));
  if (e != hipSuccess) {
    *ptr = nullptr;
    return set_err(NICP_ERR_OOM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  return NICP_OK;
}

extern "C" int nicp_device_free(nicp_ctx *c, void *ptr) {
  if (!c) return set_err(NICP_ERR_INVALID, "context is null");
  if (!ptr) return NICP_OK;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipFree(ptr));
  return NICP_OK;
}

extern "C" int nicp_device_copy(nicp_ctx *c, void *dst, const void *src, size_t bytes, int kind) {
  if (!c || (bytes && (!dst || !src)) || kind < 0 || kind > 2) return set_err(NICP_ERR_INVALID, "bad arguments");
  if (!bytes) return NICP_OK;
  HIPCHK(hipSetDevice(c->device));
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, k, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return NICP_OK;
}

// ---------------------------------------------------------------------------------------
// clouds
// ---------------------------------------------------------------------------------------
namespace {
constexpr size_t kCloudPoolEntries = 8;

// a cloud allocation of exactly `bytes`: from the context's pool when it holds one
hipError_t cloud_alloc(nicp_ctx *ctx, size_t bytes, void **p) {
  for (size_t i = 0; i < ctx->cloud_pool.size(); ++i)
    if (ctx->cloud_pool[i].first == bytes) {
      *p = ctx->cloud_pool[i].second;
      ctx->cloud_pool.erase(ctx->cloud_pool.begin() + (long)i);
      return hipSuccess;
    }
  return hipMalloc(p, bytes);
}
// back to the pool; the oldest entry is freed when the pool is full
void cloud_release(nicp_ctx *ctx, void *p, size_t bytes) {
  if (ctx->cloud_pool.size() >= kCloudPoolEntries) {
    (void)hipFree(ctx->cloud_pool.front().second);
    ctx->cloud_pool.erase(ctx->cloud_pool.begin());
  }
  ctx->cloud_pool.emplace_back(bytes, p);
}

int cloud_from_depth(nicp_ctx *ctx, const uint16_t *depth, const float *depthf, int rows, int cols, float depth_scale,
                     const nicp_camera *cam, const nicp_stats_params *sp, int keep_stats, nicp_cloud **out) {
  if (!ctx || (!depth && !depthf) || !sp || !out) return set_err(NICP_ERR_INVALID, "null argument");
  if (rows <= 0 || cols <= 0 || (long long)rows * cols > (1ll << 28))
    return set_err(NICP_ERR_INVALID, "bad image size %dx%d", rows, cols);
  CHK(check_camera(cam));
  HIPCHK(hipSetDevice(ctx->device));
  nicp_cloud *cl = new (std::nothrow) nicp_cloud();
  if (!cl) return set_err(NICP_ERR_OOM, "out of host memory");
  cl->ctx = ctx;
  FrameLayout L = frame_layout(rows, cols, keep_stats != 0, true, 0, depthf != nullptr);
  void *mem = nullptr;
  hipError_t e = cloud_alloc(ctx, L.total, &mem);
  if (e != hipSuccess) {
    delete cl;
    return set_err(NICP_ERR_OOM, "hipMalloc(%zu): %s", L.total, hipGetErrorString(e));
  }
  cl->alloc = mem;
  cl->alloc_bytes = L.total;
  cl->cap = rows * cols;
  FrameDev &f = cl->fd;
  frame_geometry(f, rows, cols, cam, sp->world_radius, depth_scale);
  frame_pointers(f, (char *)mem, L, keep_stats != 0, true, depthf != nullptr);
  set_omega_model(f, to_dev(sp, keep_stats != 0));
  f.tile_base = 0;
  f.xblk_base = 0;
  f.yblk_base = 0;
  cl->has_index = 1;
  cl->has_stats = keep_stats != 0;
  hipStream_t s = ctx->stream;
  int rc = NICP_OK;
  do {
    if ((rc = ctx->frames_tbl.reserve(sizeof(FrameDev))) != NICP_OK) break;
    const int nt = f.n_tiles, nx = xblocks(rows), ny = yblocks(cols);
    if ((rc = ctx->maps.reserve((size_t)(nt + nx + ny) * sizeof(int))) != NICP_OK) break;
    std::vector<int> zeros((size_t)(nt + nx + ny), 0);
    FrameDev *d_f = ctx->frames_tbl.as<FrameDev>();
    int *d_maps = ctx->maps.as<int>();
    e = depthf ? hipMemcpyAsync((void *)f.depthf, depthf, (size_t)rows * cols * sizeof(float), hipMemcpyHostToDevice, s)
               : hipMemcpyAsync((void *)f.depth, depth, (size_t)rows * cols * sizeof(uint16_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_f, &f, sizeof(FrameDev), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_maps, zeros.data(), zeros.size() * sizeof(int), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) { rc = set_err(NICP_ERR_HIP, "upload: %s", hipGetErrorString(e)); break; }
    ctx->prof.sink = &ctx->prof.eager;
    StatsParamsDev spd = to_dev(sp, keep_stats != 0);
    spd.fp64_gap = ctx->fp64_eigen_gap;
    rc = launch_convert(&ctx->prof, s, d_f, 1, d_maps, nt, d_maps + nt, nx, d_maps + nt + nx, ny, spd, cam->model);
    if (rc != NICP_OK) break;
    e = hipMemcpyAsync(&cl->n, f.npoints, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) { rc = set_err(NICP_ERR_HIP, "conversion: %s", hipGetErrorString(e)); break; }
  } while (0);
  if (rc != NICP_OK) {
    (void)hipFree(mem);
    delete cl;
    return rc;
  }
  *out = cl;
  return NICP_OK;
}
}  // namespace

extern "C" int nicp_cloud_from_depth_u16(nicp_ctx *ctx, const uint16_t *depth, int rows, int cols, float depth_scale,
                                         const nicp_camera *cam, const nicp_stats_params *sp, int keep_stats,
                                         nicp_cloud **out) {
  if (!depth) return set_err(NICP_ERR_INVALID, "depth is null");
  return cloud_from_depth(ctx, depth, nullptr, rows, cols, depth_scale, cam, sp, keep_stats, out);
}

extern "C" int nicp_cloud_from_depth_f32(nicp_ctx *ctx, const float *depth, int rows, int cols, const nicp_camera *cam,
                                         const nicp_stats_params *sp, int keep_stats, nicp_cloud **out) {
  if (!depth) return set_err(NICP_ERR_INVALID, "depth is null");
  return cloud_from_depth(ctx, nullptr, depth, rows, cols, 1.0f, cam, sp, keep_stats, out);
}

extern "C" int nicp_depth_scale(nicp_ctx *ctx, const uint16_t *depth, const float *depthf, int rows, int cols,
                                float depth_scale, int step, float max_depth_cov, float *out) {
  if (!ctx || (!depth && !depthf) || !out || rows <= 0 || cols <= 0 || step <= 0)
    return set_err(NICP_ERR_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(ctx->device));
  const size_t P = (size_t)rows * cols;
  const int orows = rows / step, ocols = cols / step;
  const size_t O = (size_t)orows * ocols;
  if (O == 0) return NICP_OK;
  CHK(ctx->scratch_in.reserve(P * sizeof(float)));
  CHK(ctx->depth_tmp.reserve(O * sizeof(float)));
  hipStream_t s = ctx->stream;
  if (depthf) HIPCHK(hipMemcpyAsync(ctx->scratch_in.p, depthf, P * sizeof(float), hipMemcpyHostToDevice, s));
  else HIPCHK(hipMemcpyAsync(ctx->scratch_in.p, depth, P * sizeof(uint16_t), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_depth_scale, dim3((unsigned)((O + kThreads - 1) / kThreads)), dim3(kThreads), 0, s,
                     depthf ? (const uint16_t *)nullptr : ctx->scratch_in.as<uint16_t>(),
                     depthf ? ctx->scratch_in.as<float>() : (const float *)nullptr, rows, cols, depth_scale, step,
                     max_depth_cov, ctx->depth_tmp.as<float>());
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, ctx->depth_tmp.p, O * sizeof(float), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return NICP_OK;
}

extern "C" int nicp_cloud_from_host(nicp_ctx *ctx, int n, const float *points4, const float *normals4,
                                    const float *curvature, const float *omega_p16, const float *omega_n16,
                                    nicp_cloud **out) {
  if (!ctx || !out || (n > 0 && !points4) || n < 0) return set_err(NICP_ERR_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(ctx->device));
  nicp_cloud *cl = new (std::nothrow) nicp_cloud();
  if (!cl) return set_err(NICP_ERR_OOM, "out of host memory");
  cl->ctx = ctx;
  Carver cv;
  const size_t cap = std::max(n, 1);
  size_t o_pts = cv.take<float4>(cap), o_nrm = cv.take<float4>(cap), o_omg = cv.take<float4>(cap * 3),
         o_np = cv.take<int>(1);
  void *mem = nullptr;
  hipError_t e = cloud_alloc(ctx, cv.off, &mem);
  if (e != hipSuccess) {
    delete cl;
    return set_err(NICP_ERR_OOM, "hipMalloc: %s", hipGetErrorString(e));
  }
  cl->alloc = mem;
  cl->alloc_bytes = cv.off;
  cl->n = n;
  cl->cap = (int)cap;
  FrameDev &f = cl->fd;
  memset(&f, 0, sizeof(f));
  f.pts = (float4 *)((char *)mem + o_pts);
  f.nrm = (float4 *)((char *)mem + o_nrm);
  f.omg = (float4 *)((char *)mem + o_omg);
  f.npoints = (int *)((char *)mem + o_np);
  std::vector<float4> nrm(cap), omg(cap * 3);
  for (int i = 0; i < n; ++i) {
    const float *nv = normals4 ? normals4 + 4 * (size_t)i : nullptr;
    nrm[i] = make_float4(nv ? nv[0] : 0.0f, nv ? nv[1] : 0.0f, nv ? nv[2] : 0.0f, curvature ? curvature[i] : 0.0f);
    float op[6] = {0, 0, 0, 0, 0, 0}, on[6] = {0, 0, 0, 0, 0, 0};
    const int up[6] = {0, 4, 8, 5, 9, 10};
    for (int k = 0; k < 6; ++k) {
      if (omega_p16) op[k] = omega_p16[16 * (size_t)i + up[k]];
      if (omega_n16) on[k] = omega_n16[16 * (size_t)i + up[k]];
    }
    omg[3 * i] = make_float4(op[0], op[1], op[2], op[3]);
    omg[3 * i + 1] = make_float4(op[4], op[5], on[0], on[1]);
    omg[3 * i + 2] = make_float4(on[2], on[3], on[4], on[5]);
  }
  hipStream_t s = ctx->stream;
  if (n > 0) {
    e = hipMemcpyAsync(f.pts, points4, (size_t)n * sizeof(float4), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(f.nrm, nrm.data(), (size_t)n * sizeof(float4), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(f.omg, omg.data(), (size_t)n * 3 * sizeof(float4), hipMemcpyHostToDevice, s);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(f.npoints, &cl->n, sizeof(int), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(mem);
    delete cl;
    return set_err(NICP_ERR_HIP, "upload: %s", hipGetErrorString(e));
  }
  *out = cl;
  return NICP_OK;
}

extern "C" int nicp_cloud_destroy(nicp_cloud *cl) {
  if (!cl) return NICP_OK;
  if (cl->alloc || cl->gauss_alloc) {
    (void)hipSetDevice(cl->ctx->device);
    if (cl->alloc && cl->alloc_bytes) {
      cloud_release(cl->ctx, cl->alloc, cl->alloc_bytes);
    } else if (cl->alloc) {
      (void)hipStreamSynchronize(cl->ctx->stream);
      (void)hipFree(cl->alloc);
    }
    if (cl->gauss_alloc) {
      (void)hipStreamSynchronize(cl->ctx->stream);
      (void)hipFree(cl->gauss_alloc);
    }
  }
  delete cl;
  return NICP_OK;
}

extern "C" int nicp_cloud_size(const nicp_cloud *cl, int *n) {
  if (!cl || !n) return set_err(NICP_ERR_INVALID, "null argument");
  *n = cl->n;
  return NICP_OK;
}

extern "C" int nicp_cloud_download(nicp_cloud *cl, float *points4, float *normals4, float *curvature,
                                   float *omega_p16, float *omega_n16) {
  if (!cl) return set_err(NICP_ERR_INVALID, "cloud is null");
  const int n = cl->n;
  if (n == 0) return NICP_OK;
  hipStream_t s = cl->ctx->stream;
  HIPCHK(hipSetDevice(cl->ctx->device));
  std::vector<float4> pts, nrm, omg;
  if (points4) HIPCHK(hipMemcpyAsync(points4, cl->fd.pts, (size_t)n * sizeof(float4), hipMemcpyDeviceToHost, s));
  if (normals4 || curvature) {
    nrm.resize(n);
    HIPCHK(hipMemcpyAsync(nrm.data(), cl->fd.nrm, (size_t)n * sizeof(float4), hipMemcpyDeviceToHost, s));
  }
  if (omega_p16 || omega_n16) {
    omg.resize((size_t)n * 3);
    HIPCHK(hipMemcpyAsync(omg.data(), cl->fd.omg, (size_t)n * 3 * sizeof(float4), hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  for (int i = 0; i < n; ++i) {
    if (normals4) {
      normals4[4 * (size_t)i] = nrm[i].x;
      normals4[4 * (size_t)i + 1] = nrm[i].y;
      normals4[4 * (size_t)i + 2] = nrm[i].z;
      normals4[4 * (size_t)i + 3] = 0.0f;
    }
    if (curvature) curvature[i] = nrm[i].w;
    const float *o = (const float *)&omg[3 * (size_t)i];
    for (int which = 0; which < 2; ++which) {
      float *O = which ? omega_n16 : omega_p16;
      if (!O) continue;
      O += 16 * (size_t)i;
      const float *u = o + 6 * which;
      for (int k = 0; k < 16; ++k) O[k] = 0.0f;
      O[0] = u[0]; O[4] = O[1] = u[1]; O[8] = O[2] = u[2];
      O[5] = u[3]; O[9] = O[6] = u[4]; O[10] = u[5];
    }
  }
  return NICP_OK;
}

extern "C" int nicp_cloud_download_stats(nicp_cloud *cl, float *eigvecs9, float *eigvals3, float *means4,
                                         int *npts) {
  if (!cl) return set_err(NICP_ERR_INVALID, "cloud is null");
  if (!cl->has_stats || !cl->fd.stats) return set_err(NICP_ERR_STATE, "the cloud was converted without keep_stats");
  const int n = cl->n;
  std::vector<float4> st((size_t)n * 4);
  HIPCHK(hipSetDevice(cl->ctx->device));
  if (n) HIPCHK(hipMemcpyAsync(st.data(), cl->fd.stats, st.size() * sizeof(float4), hipMemcpyDeviceToHost, cl->ctx->stream));
  HIPCHK(hipStreamSynchronize(cl->ctx->stream));
  for (int i = 0; i < n; ++i) {
    const float *s = (const float *)&st[4 * (size_t)i];
    if (eigvecs9)
      for (int k = 0; k < 9; ++k) eigvecs9[9 * (size_t)i + k] = s[k];
    if (eigvals3)
      for (int k = 0; k < 3; ++k) eigvals3[3 * (size_t)i + k] = s[9 + k];
    const int cnt = __builtin_bit_cast(int, s[15]);
    if (means4) {
      means4[4 * (size_t)i] = s[12];
      means4[4 * (size_t)i + 1] = s[13];
      means4[4 * (size_t)i + 2] = s[14];
      // Stats::mean() is a Point: w = 1 (pointaccumulator.h:66-70, homogeneousvector4f.h:19)
      means4[4 * (size_t)i + 3] = 1.0f;
    }
    if (npts) npts[i] = cnt;
  }
  return NICP_OK;
}

extern "C" int nicp_cloud_index_image(nicp_cloud *cl, int *index_image) {
  if (!cl || !index_image) return set_err(NICP_ERR_INVALID, "null argument");
  if (!cl->has_index) return set_err(NICP_ERR_STATE, "the cloud has no index image");
  HIPCHK(hipSetDevice(cl->ctx->device));
  HIPCHK(hipMemcpyAsync(index_image, cl->fd.index, (size_t)cl->fd.rows * cl->fd.cols * sizeof(int),
                        hipMemcpyDeviceToHost, cl->ctx->stream));
  HIPCHK(hipStreamSynchronize(cl->ctx->stream));
  return NICP_OK;
}

// ---------------------------------------------------------------------------------------
// single-pair operations
// ---------------------------------------------------------------------------------------
namespace {

AlignBuffers ctx_buffers(nicp_ctx *c) {
  return AlignBuffers{&c->pairs_tbl, &c->maps, &c->zbuf, &c->cur_index, &c->pairs_buf, &c->state, &c->partials,
                      &c->result};
}

// uploads a frame table of the given clouds
int upload_frames(nicp_ctx *c, std::vector<FrameDev> &frames, FrameDev **d) {
  CHK(c->frames_tbl.reserve(frames.size() * sizeof(FrameDev)));
  *d = c->frames_tbl.as<FrameDev>();
  HIPCHK(hipMemcpyAsync(*d, frames.data(), frames.size() * sizeof(FrameDev), hipMemcpyHostToDevice, c->stream));
  return NICP_OK;
}

}  // namespace

namespace {
// PinholePointProjector::project (pinholepointprojector.cpp:33-66) of a cloud seen from
// `transform`; leaves the index image in ctx->cur_index and the depth image (FLT_MAX where
// empty) in ctx->depth_tmp, both rows x cols on the device, enqueued on ctx->stream.
int project_device(nicp_ctx *ctx, const nicp_cloud *cloud, const float transform[16], const nicp_camera *cam,
                   int rows, int cols) {
  std::vector<FrameDev> frames(1, cloud->fd);
  frames[0].rows = rows;  // projection capacity follows the cloud size
  frames[0].cols = std::max(1, cloud->n);
  FrameDev *d_frames;
  CHK(upload_frames(ctx, frames, &d_frames));
  int zero = 0;
  const nicp_camera *cams[1] = {cam};
  PairPlan plan;
  // a pair whose "current" frame is the cloud itself and whose image is rows x cols
  frames[0].rows = 1;
  frames[0].cols = std::max(1, cloud->n);
  make_plan(frames, 1, &zero, &zero, cams, plan);
  PairDev &pd = plan.pairs[0];
  pd.rows = rows;
  pd.cols = cols;
  pd.sph = sph_params(cam, rows, cols);
  const int P = rows * cols;
  plan_pixel_tiles(plan);
  plan.total_pix = align_up((size_t)P, 64);
  plan.blk_ref.clear();
  plan.total_rtiles = 0;
  pd.rtile_base = 0;
  pd.n_rtiles = 0;
  AlignBuffers B = ctx_buffers(ctx);
  int *d_bp, *d_br, *d_bc;
  PairDev *d_pairs;
  CHK(prepare_align(ctx->stream, plan, B, transform, &d_bp, &d_br, &d_bc, &d_pairs));
  // the current-projection slot must hold the transform's KRt (PinholePointProjector::setTransform)
  {
    float T[16], KRt[16];
    memcpy(T, transform, sizeof(T));
    set_row3(T);
    projector_state(cam->model, cam->K, T, KRt);
    HIPCHK(hipMemcpyAsync(ctx->state.as<float>() + kSlotKRtCur, KRt, sizeof(KRt), hipMemcpyHostToDevice, ctx->stream));
  }
  CHK(ctx->depth_tmp.reserve((size_t)P * sizeof(float)));
  const dim3 tb(kThreads);
  hipLaunchKernelGGL(k_zfill, dim3(pd.n_tiles), tb, 0, ctx->stream, d_pairs, d_bp);
  if (plan.total_ctiles) {
    if (cam->model == NICP_SPHERICAL)
      hipLaunchKernelGGL(k_zproject_sph, dim3(plan.total_ctiles), tb, 0, ctx->stream, d_pairs, d_frames, d_bc, 1, 0);
    else
      hipLaunchKernelGGL(k_zproject, dim3(plan.total_ctiles), tb, 0, ctx->stream, d_pairs, d_frames, d_bc, 1, 0);
  }
  hipLaunchKernelGGL(k_zresolve, dim3(pd.n_tiles), tb, 0, ctx->stream, d_pairs, d_bp, ctx->depth_tmp.as<float>());
  HIPCHK(hipGetLastError());
  return NICP_OK;
}
}  // namespace

extern "C" int nicp_project(nicp_ctx *ctx, const nicp_cloud *cloud, const float transform[16], const nicp_camera *cam,
                            int rows, int cols, int *index_image, float *depth_image) {
  if (!ctx || !cloud || !transform || !index_image) return set_err(NICP_ERR_INVALID, "null argument");
  if (rows <= 0 || cols <= 0) return set_err(NICP_ERR_INVALID, "bad image size");
  CHK(check_camera(cam));
  HIPCHK(hipSetDevice(ctx->device));
  CHK(project_device(ctx, cloud, transform, cam, rows, cols));
  const size_t P = (size_t)rows * cols;
  HIPCHK(hipMemcpyAsync(index_image, ctx->cur_index.p, P * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  if (depth_image)
    HIPCHK(hipMemcpyAsync(depth_image, ctx->depth_tmp.p, P * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return NICP_OK;
}

extern "C" int nicp_correspondences(nicp_ctx *ctx, const nicp_cloud *ref, const nicp_cloud *cur, const int *ref_index,
                                    const int *cur_index, int rows, int cols, const float T[16],
                                    const nicp_align_params *ap, int *pairs, int *num_pairs) {
  if (!ctx || !ref || !cur || !ref_index || !cur_index || !T || !pairs || !num_pairs)
    return set_err(NICP_ERR_INVALID, "null argument");
  if (rows <= 0 || cols <= 0) return set_err(NICP_ERR_INVALID, "bad image size");
  CHK(check_align_params(ap));
  const int P = rows * cols;
  for (int i = 0; i < P; ++i)
    if (ref_index[i] >= ref->n || cur_index[i] >= cur->n)
      return set_err(NICP_ERR_INVALID, "index image entry out of range at pixel %d", i);
  HIPCHK(hipSetDevice(ctx->device));
  std::vector<FrameDev> frames = {ref->fd, cur->fd};
  frames[0].rows = 1; frames[0].cols = std::max(1, ref->n);
  frames[1].rows = rows; frames[1].cols = cols;
  FrameDev *d_frames;
  CHK(upload_frames(ctx, frames, &d_frames));
  nicp_camera cam{};
  cam.K[0] = cam.K[4] = cam.K[8] = 1.0f;
  cam.min_distance = 1.0f;
  cam.max_distance = 1.0f;
  const nicp_camera *cams[1] = {&cam};
  int r0 = 0, c1 = 1;
  PairPlan plan;
  make_plan(frames, 1, &r0, &c1, cams, plan);
  AlignBuffers B = ctx_buffers(ctx);
  int *d_bp, *d_br, *d_bc;
  PairDev *d_pairs;
  CHK(prepare_align(ctx->stream, plan, B, nullptr, &d_bp, &d_br, &d_bc, &d_pairs));
  // the finder's transform goes into the Tinv slot; index images into z-buffer / current index
  HIPCHK(hipMemcpyAsync(ctx->state.as<float>() + kSlotTinv, T, 16 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  std::vector<unsigned long long> keys(P);
  for (int i = 0; i < P; ++i) keys[i] = ref_index[i] < 0 ? kZInit : (unsigned long long)(unsigned int)ref_index[i];
  HIPCHK(hipMemcpyAsync(ctx->zbuf.p, keys.data(), (size_t)P * sizeof(unsigned long long), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->cur_index.p, cur_index, (size_t)P * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_corr_lin, dim3(plan.total_dtiles), dim3(kThreads), 0, ctx->stream, d_pairs, d_frames,
                     (const int *)plan.d_dpair, to_dev(ap), 1);
  HIPCHK(hipGetLastError());
  std::vector<int2> pp(P);
  HIPCHK(hipMemcpyAsync(pp.data(), ctx->pairs_buf.p, (size_t)P * sizeof(int2), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  int n = 0;
  for (int i = 0; i < P; ++i)
    if (pp[i].x >= 0) {
      pairs[2 * n] = pp[i].x;
      pairs[2 * n + 1] = pp[i].y;
      ++n;
    }
  *num_pairs = n;
  return NICP_OK;
}

extern "C" int nicp_linearize(nicp_ctx *ctx, const nicp_cloud *ref, const nicp_cloud *cur, const int *pairs,
                              int num_pairs, const float T[16], const nicp_align_params *ap, double H[36],
                              double b[6], double *error, int *inliers) {
  if (!ctx || !ref || !cur || (!pairs && num_pairs > 0) || !T || num_pairs < 0)
    return set_err(NICP_ERR_INVALID, "bad arguments");
  CHK(check_align_params(ap));
  for (int i = 0; i < num_pairs; ++i)
    if (pairs[2 * i] < 0 || pairs[2 * i] >= ref->n || pairs[2 * i + 1] < 0 || pairs[2 * i + 1] >= cur->n)
      return set_err(NICP_ERR_INVALID, "pair %d out of range", i);
  HIPCHK(hipSetDevice(ctx->device));
  std::vector<FrameDev> frames = {ref->fd, cur->fd};
  const int npx = std::max(num_pairs, 1);
  frames[0].rows = 1; frames[0].cols = std::max(1, ref->n);
  frames[1].rows = 1; frames[1].cols = npx;
  FrameDev *d_frames;
  CHK(upload_frames(ctx, frames, &d_frames));
  nicp_camera cam{};
  cam.K[0] = cam.K[4] = cam.K[8] = 1.0f;
  cam.min_distance = cam.max_distance = 1.0f;
  const nicp_camera *cams[1] = {&cam};
  int r0 = 0, c1 = 1;
  PairPlan plan;
  make_plan(frames, 1, &r0, &c1, cams, plan);
  AlignBuffers B = ctx_buffers(ctx);
  int *d_bp, *d_br, *d_bc;
  PairDev *d_pairs;
  CHK(prepare_align(ctx->stream, plan, B, nullptr, &d_bp, &d_br, &d_bc, &d_pairs));
  HIPCHK(hipMemcpyAsync(ctx->state.as<float>() + kSlotTinv, T, 16 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  std::vector<int2> pp(npx, make_int2(-1, -1));
  for (int i = 0; i < num_pairs; ++i) pp[i] = make_int2(pairs[2 * i], pairs[2 * i + 1]);
  HIPCHK(hipMemcpyAsync(ctx->pairs_buf.p, pp.data(), (size_t)npx * sizeof(int2), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_lin_pairs, dim3(plan.total_pix_tiles), dim3(kThreads), 0, ctx->stream, d_pairs, d_frames, d_bp,
                     to_dev(ap), kSlotTinv, npx);
  hipLaunchKernelGGL(k_solve, dim3(1), dim3(kSolveThreads), 0, ctx->stream, d_pairs, 0.0f, 2, 0, 0);
  HIPCHK(hipGetLastError());
  double r[2 * kPartial];
  HIPCHK(hipMemcpyAsync(r, ctx->result.p, sizeof(r), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const double *t = r + kPartial;
  int e = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 6; ++j) {
      if (H) H[j * 6 + i] = H[i * 6 + j] = t[e];
      ++e;
    }
  if (b)
    for (int i = 0; i < 6; ++i) b[i] = t[21 + i];
  if (error) *error = t[27];
  if (inliers) *inliers = (int)llround(t[28]);
  return NICP_OK;
}

namespace {
// AlignerProjective::align of one pair with the given workspace (the context's growable
// buffers, or fixed views of the no-allocation workspace)
int align_single(nicp_ctx *ctx, DevBuf &frames_tbl, AlignBuffers B, const nicp_cloud *ref, const nicp_cloud *cur,
                 const nicp_camera *cam, int rows, int cols, const float init_guess[16], const nicp_align_params *ap,
                 float T_out[16], nicp_align_result *result, const float *ref_offset = nullptr,
                 const float *cur_offset = nullptr) {
  if (!ctx || !ref || !cur) return set_err(NICP_ERR_INVALID, "null argument");
  if (rows <= 0 || cols <= 0) return set_err(NICP_ERR_INVALID, "bad image size");
  CHK(check_camera(cam));
  CHK(check_align_params(ap));
  HIPCHK(hipSetDevice(ctx->device));
  FrameDev frames[2] = {ref->fd, cur->fd};
  // projections cover the clouds' capacities and use the projector image rows x cols
  frames[0].rows = 1; frames[0].cols = std::max(1, ref->n);
  frames[1].rows = 1; frames[1].cols = std::max(1, cur->n);
  CHK(frames_tbl.reserve(sizeof(frames)));
  FrameDev *d_frames = frames_tbl.as<FrameDev>();
  HIPCHK(hipMemcpyAsync(d_frames, frames, sizeof(frames), hipMemcpyHostToDevice, ctx->stream));
  const nicp_camera *cams[1] = {cam};
  int r0 = 0, c1 = 1;
  PairPlan plan;
  make_plan(std::vector<FrameDev>(frames, frames + 2), 1, &r0, &c1, cams, plan);
  PairDev &pd = plan.pairs[0];
  pd.rows = rows;
  pd.cols = cols;
  pd.sph = sph_params(cam, rows, cols);
  plan_pixel_tiles(plan);
  plan.total_pix = align_up((size_t)rows * cols, 64);
  plan.ref_offset = ref_offset;
  plan.cur_offset = cur_offset;
  int *d_bp, *d_br, *d_bc;
  PairDev *d_pairs;
  float I4[16];
  set_identity4(I4);
  CHK(prepare_align(ctx->stream, plan, B, init_guess ? init_guess : I4, &d_bp, &d_br, &d_bc, &d_pairs));
  HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
  ctx->prof.sink = &ctx->prof.eager;
  CHK(launch_align(&ctx->prof, ctx->stream, plan, d_pairs, d_frames, d_bp, d_br, d_bc, ap));
  HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
  float T[16];
  double r[kResultDoubles];
  HIPCHK(hipMemcpyAsync(T, B.state->p, sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(r, B.result->p, sizeof(r), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  if (ap->outer_iterations == 0) memset(r, 0, sizeof(r));
  if (T_out) memcpy(T_out, T, sizeof(T));
  if (result) fill_result(r, ap, result, ms);
  return NICP_OK;
}
}  // namespace

extern "C" int nicp_align(nicp_ctx *ctx, const nicp_cloud *ref, const nicp_cloud *cur, const nicp_camera *cam,
                          int rows, int cols, const float init_guess[16], const nicp_align_params *ap, float T_out[16],
                          nicp_align_result *result) {
  return nicp_align_offsets(ctx, ref, cur, cam, rows, cols, init_guess, nullptr, nullptr, ap, T_out, result);
}

extern "C" int nicp_align_offsets(nicp_ctx *ctx, const nicp_cloud *ref, const nicp_cloud *cur, const nicp_camera *cam,
                                  int rows, int cols, const float init_guess[16], const float reference_offset[16],
                                  const float current_offset[16], const nicp_align_params *ap, float T_out[16],
                                  nicp_align_result *result) {
  if (!ctx) return set_err(NICP_ERR_INVALID, "null argument");
  ctx->last_valid = false;
  CHK(align_single(ctx, ctx->frames_tbl, ctx_buffers(ctx), ref, cur, cam, rows, cols, init_guess, ap, T_out, result,
                   reference_offset, current_offset));
  ctx->last_rows = rows;
  ctx->last_cols = cols;
  ctx->last_valid = ap->outer_iterations > 0;
  return NICP_OK;
}

extern "C" int nicp_align_noalloc(nicp_ctx *ctx, const nicp_cloud *ref, const nicp_cloud *cur, const nicp_camera *cam,
                                  int rows, int cols, const float init_guess[16], const nicp_align_params *ap,
                                  float T_out[16], nicp_align_result *result) {
  if (!ctx || !ref || !cur) return set_err(NICP_ERR_INVALID, "null argument");
  if ((long long)rows * cols > kNoAllocMaxPixels || rows > kNoAllocMaxRows || cols > kNoAllocMaxCols ||
      ref->n > kNoAllocMaxPixels || cur->n > kNoAllocMaxPixels)
    return set_err(NICP_ERR_INVALID,
                   "nicp_align_noalloc handles at most %d projector pixels (%d rows, %d columns) and %d points "
                   "per cloud (got %dx%d, %d and %d points)",
                   kNoAllocMaxPixels, kNoAllocMaxRows, kNoAllocMaxCols, kNoAllocMaxPixels, rows, cols, ref->n, cur->n);
  if (!ctx->noalloc_ws) return set_err(NICP_ERR_STATE, "the context has no fixed workspace");
  // views of the fixed workspace, sized for the largest pair the entry point accepts; every
  // region is rewritten or re-initialised by each call's launches (k_zfill clears the z-buffer)
  // bounds within these limits: 1024-pixel tiles <= 300 per image and cloud; kDTH x kDTW tiles
  // <= P / 1024 + rows / 16 + cols / 64 + 1 <= 381; source-tile blocks <= 38; fused tiles <= 111
  const size_t P = align_up((size_t)kNoAllocMaxPixels, 64);
  const size_t parts = 400, nmaps = 2048;
  Carver cv;
  const size_t o_frames = cv.take<FrameDev>(2), o_pairs = cv.take<PairDev>(1), o_maps = cv.take<int>(nmaps),
               o_z = cv.take<unsigned long long>(P), o_ci = cv.take<int>(P), o_pp = cv.take<int2>(P),
               o_state = cv.take<float>(kStateFloats), o_part = cv.take<double>(parts * kPartial),
               o_res = cv.take<double>(kResultDoubles);
  static_assert(3 * (kNoAllocMaxPixels + kTile - 1) / kTile + 400 + 40 + 120 <= 2048, "no-allocation maps");
  if (cv.off > kNoAllocScratchBytes) return set_err(NICP_ERR_STATE, "fixed workspace layout exceeds its size");
  char *base = (char *)ctx->noalloc_ws;
  DevBuf frames_tbl, pairs_tbl, maps, zbuf, cur_index, pairs_buf, state, partials, res;
  const struct {
    DevBuf *b;
    size_t off, bytes;
  } views[] = {{&frames_tbl, o_frames, 2 * sizeof(FrameDev)},
               {&pairs_tbl, o_pairs, sizeof(PairDev)},
               {&maps, o_maps, nmaps * sizeof(int)},
               {&zbuf, o_z, P * sizeof(unsigned long long)},
               {&cur_index, o_ci, P * sizeof(int)},
               {&pairs_buf, o_pp, P * sizeof(int2)},
               {&state, o_state, kStateFloats * sizeof(float)},
               {&partials, o_part, parts * kPartial * sizeof(double)},
               {&res, o_res, kResultDoubles * sizeof(double)}};
  for (const auto &v : views) {
    v.b->p = base + v.off;
    v.b->bytes = v.bytes;
    v.b->fixed = true;
  }
  AlignBuffers B{&pairs_tbl, &maps, &zbuf, &cur_index, &pairs_buf, &state, &partials, &res};
  ctx->last_valid = false;  // the correspondences of this call stay in the fixed workspace
  return align_single(ctx, frames_tbl, B, ref, cur, cam, rows, cols, init_guess, ap, T_out, result);
}

extern "C" int nicp_solution_statistics(nicp_ctx *ctx, const double H[36], const float T[16],
                                        const nicp_align_params *ap, nicp_align_result *result) {
  if (!ctx || !H || !T || !ap || !result) return set_err(NICP_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  CHK(ctx->pairs_tbl.reserve(sizeof(PairDev)));
  CHK(ctx->state.reserve(kStateFloats * sizeof(float)));
  CHK(ctx->result.reserve(kResultDoubles * sizeof(double)));
  PairDev pd{};
  pd.state = ctx->state.as<float>();
  pd.result = ctx->result.as<double>();
  double r[kResultDoubles] = {0};
  int e = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 6; ++j) r[kPartial + e++] = H[j * 6 + i];
  HIPCHK(hipMemcpyAsync(ctx->pairs_tbl.p, &pd, sizeof(pd), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(pd.state + kSlotT, T, 16 * sizeof(float), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(pd.result, r, sizeof(r), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_solution_stats, dim3(1), dim3(64), 0, s, ctx->pairs_tbl.as<PairDev>(),
                     ap->translational_min_eigen_ratio, ap->rotational_min_eigen_ratio);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(r, pd.result, sizeof(r), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const double *st = r + 2 * kPartial;
  memcpy(result->omega, st + nicp_stats::kSolOmega, sizeof(result->omega));
  memcpy(result->mean, st + nicp_stats::kSolMean, sizeof(result->mean));
  result->translational_ratio = st[nicp_stats::kSolTrans];
  result->rotational_ratio = st[nicp_stats::kSolRot];
  result->solution_valid = st[nicp_stats::kSolValid] != 0.0;
  return NICP_OK;
}

extern "C" int nicp_last_correspondences(nicp_ctx *ctx, int *pairs, int capacity, int *num_pairs) {
  if (!ctx || !num_pairs) return set_err(NICP_ERR_INVALID, "null argument");
  if (!ctx->last_valid) return set_err(NICP_ERR_STATE, "no alignment has been run on this context");
  const int P = ctx->last_rows * ctx->last_cols;
  std::vector<int2> pp(P);
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipMemcpyAsync(pp.data(), ctx->pairs_buf.p, (size_t)P * sizeof(int2), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  int n = 0;
  for (int i = 0; i < P; ++i)
    if (pp[i].x >= 0) {
      if (pairs && n < capacity) {
        pairs[2 * n] = pp[i].x;
        pairs[2 * n + 1] = pp[i].y;
      }
      ++n;
    }
  *num_pairs = n;
  if (pairs && n > capacity) return set_err(NICP_ERR_INVALID, "capacity %d < %d correspondences", capacity, n);
  return NICP_OK;
}

// ---------------------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------------------
extern "C" int nicp_batch_create(nicp_ctx *ctx, int num_frames, const int *rows, const int *cols,
                                 const nicp_camera *cams, const nicp_stats_params *sp, nicp_batch **out) {
  if (!ctx || num_frames <= 0 || !rows || !cols || !cams || !sp || !out)
    return set_err(NICP_ERR_INVALID, "bad arguments");
  for (int i = 0; i < num_frames; ++i) {
    if (rows[i] <= 0 || cols[i] <= 0 || (long long)rows[i] * cols[i] > (1ll << 26))
      return set_err(NICP_ERR_INVALID, "frame %d: bad size %dx%d", i, rows[i], cols[i]);
    CHK(check_camera(&cams[i]));
    if (cams[i].model != cams[0].model) return set_err(NICP_ERR_INVALID, "the frames of a batch share one projector model");
  }
  HIPCHK(hipSetDevice(ctx->device));
  nicp_batch *b = new (std::nothrow) nicp_batch();
  if (!b) return set_err(NICP_ERR_OOM, "out of host memory");
  b->ctx = ctx;
  b->num_frames = num_frames;
  b->frames.resize(num_frames);
  b->npoints_host.assign(num_frames, 0);
  b->sp = to_dev(sp, 0);
  std::vector<FrameLayout> lay(num_frames);
  size_t off = 0;
  int tiles = 0, xb = 0, yb = 0;
  std::vector<int> blk, xblk, yblk;
  for (int i = 0; i < num_frames; ++i) {
    lay[i] = frame_layout(rows[i], cols[i], false, true, off, false, spack_frame(cams[i].model, rows[i], cols[i]));
    off = align_up(lay[i].total, 256);
    FrameDev &f = b->frames[i];
    frame_geometry(f, rows[i], cols[i], &cams[i], sp->world_radius, 0.001f);
    set_omega_model(f, b->sp);
    f.tile_base = tiles;
    f.xblk_base = xb;
    f.yblk_base = yb;
    tiles += f.n_tiles;
    xb += xblocks(rows[i]);
    yb += yblocks(cols[i]);
    blk.insert(blk.end(), f.n_tiles, i);
    xblk.insert(xblk.end(), xblocks(rows[i]), i);
    yblk.insert(yblk.end(), yblocks(cols[i]), i);
  }
  Carver cv;
  cv.off = off;
  const size_t o_tbl = cv.take<FrameDev>(num_frames);
  const size_t o_blk = cv.take<int>(blk.size());
  const size_t o_x = cv.take<int>(xblk.size());
  const size_t o_y = cv.take<int>(yblk.size());
  const size_t o_ord = cv.take<int>(num_frames);
  const size_t o_pb = cv.take<long long>(num_frames);
  std::vector<long long> pix_base(num_frames);
  long long total_pix = 0;
  for (int i = 0; i < num_frames; ++i) {
    pix_base[i] = total_pix;
    total_pix += (long long)rows[i] * cols[i];
  }
  std::vector<int> order(num_frames);
  for (int i = 0; i < num_frames; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return (long long)rows[x] * cols[x] > (long long)rows[y] * cols[y];
  });
  hipError_t e = hipMalloc(&b->arena, cv.off);
  if (e != hipSuccess) {
    delete b;
    return set_err(NICP_ERR_OOM, "hipMalloc(%zu): %s", cv.off, hipGetErrorString(e));
  }
  char *base = (char *)b->arena;
  for (int i = 0; i < num_frames; ++i)
    frame_pointers(b->frames[i], base, lay[i], false, true, false, spack_frame(cams[i].model, rows[i], cols[i]));
  b->d_frames = (FrameDev *)(base + o_tbl);
  b->d_blk_frame = (int *)(base + o_blk);
  b->d_xblk_frame = (int *)(base + o_x);
  b->d_yblk_frame = (int *)(base + o_y);
  b->d_frame_order = (int *)(base + o_ord);
  b->d_pix_base = (long long *)(base + o_pb);
  b->total_pix = total_pix;
  b->total_tiles = tiles;
  b->total_xblk = xb;
  b->total_yblk = yb;
  hipStream_t s = ctx->stream;
  e = hipMemcpyAsync(b->d_frames, b->frames.data(), num_frames * sizeof(FrameDev), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(b->d_blk_frame, blk.data(), blk.size() * sizeof(int), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(b->d_xblk_frame, xblk.data(), xblk.size() * sizeof(int), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(b->d_yblk_frame, yblk.data(), yblk.size() * sizeof(int), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(b->d_frame_order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->d_pix_base, pix_base.data(), pix_base.size() * sizeof(long long), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(b->arena);
    delete b;
    return set_err(NICP_ERR_HIP, "batch upload: %s", hipGetErrorString(e));
  }
  *out = b;
  return NICP_OK;
}

extern "C" int nicp_batch_destroy(nicp_batch *b) {
  if (!b) return NICP_OK;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  std::vector<std::vector<ProfEntry> *> &pg = b->ctx->pending_graphs;
  for (std::vector<ProfEntry> *ev : {&b->graph.events, &b->learned_graph.events, &b->conv_graph.events})
    pg.erase(std::remove(pg.begin(), pg.end(), ev), pg.end());
  b->graph.reset();
  b->conv_graph.reset();
  b->learned_graph.reset();
  DevBuf *bufs[] = {&b->pairs_tbl,    &b->maps,         &b->zbuf,     &b->cur_index,   &b->pairs_buf,
                    &b->state,        &b->partials,     &b->result,   &b->sbox,        &b->synth_buf,
                    &b->pr_pairs_tbl, &b->pr_maps,      &b->pr_zbuf,  &b->pr_cur_index, &b->pr_pairs_buf,
                    &b->pr_state,     &b->pr_partials,  &b->pr_result, &b->learned_aux};
  for (DevBuf *x : bufs) x->release();
  b->stage.release();
  (void)hipFree(b->arena);
  delete b;
  return NICP_OK;
}

extern "C" int nicp_batch_upload_depth(nicp_batch *b, int frame, const uint16_t *depth, float depth_scale) {
  if (!b || !depth || frame < 0 || frame >= b->num_frames) return set_err(NICP_ERR_INVALID, "bad arguments");
  FrameDev &f = b->frames[frame];
  HIPCHK(hipSetDevice(b->ctx->device));
  if (f.depth_scale != depth_scale) {
    f.depth_scale = depth_scale;
    HIPCHK(hipMemcpyAsync(&b->d_frames[frame], &f, sizeof(FrameDev), hipMemcpyHostToDevice, b->ctx->stream));
  }
  HIPCHK(hipMemcpyAsync((void *)f.depth, depth, (size_t)f.rows * f.cols * sizeof(uint16_t), hipMemcpyHostToDevice,
                        b->ctx->stream));
  HIPCHK(hipStreamSynchronize(b->ctx->stream));
  return NICP_OK;
}

extern "C" int nicp_batch_upload_depth_device(nicp_batch *b, const uint16_t *depth, float depth_scale) {
  if (!b || !depth) return set_err(NICP_ERR_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(b->ctx->device));
  bool scale_changed = false;
  for (FrameDev &f : b->frames) {
    scale_changed = scale_changed || f.depth_scale != depth_scale;
    f.depth_scale = depth_scale;
  }
  if (scale_changed)
    HIPCHK(hipMemcpyAsync(b->d_frames, b->frames.data(), b->frames.size() * sizeof(FrameDev), hipMemcpyHostToDevice,
                          b->ctx->stream));
  LAUNCH((Prof *)nullptr, 0, k_batch_depth_copy, dim3(kDepthCopyBlocks, std::min(b->num_frames, 65535)), dim3(kThreads),
         b->ctx->stream, b->d_frames, b->d_pix_base, b->num_frames, depth, (uint16_t *)nullptr, 1);
  // the host copy of the frame table is pageable: wait for it before the caller may change it
  if (scale_changed) HIPCHK(hipStreamSynchronize(b->ctx->stream));
  return NICP_OK;
}

extern "C" int nicp_batch_download_depth_device(nicp_batch *b, uint16_t *depth) {
  if (!b || !depth) return set_err(NICP_ERR_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(b->ctx->device));
  LAUNCH((Prof *)nullptr, 0, k_batch_depth_copy, dim3(kDepthCopyBlocks, std::min(b->num_frames, 65535)), dim3(kThreads),
         b->ctx->stream, b->d_frames, b->d_pix_base, b->num_frames, (const uint16_t *)nullptr, depth, 0);
  return NICP_OK;
}

extern "C" int nicp_batch_num_pixels(const nicp_batch *b, long long *num_pixels) {
  if (!b || !num_pixels) return set_err(NICP_ERR_INVALID, "bad arguments");
  *num_pixels = b->total_pix;
  return NICP_OK;
}

extern "C" int nicp_batch_convert(nicp_batch *b) {
  if (!b) return set_err(NICP_ERR_INVALID, "batch is null");
  HIPCHK(hipSetDevice(b->ctx->device));
  b->ctx->prof.sink = &b->ctx->prof.eager;
  StatsParamsDev spd = b->sp;
  spd.fp64_gap = b->ctx->fp64_eigen_gap;
  // the linearisers form closed-form information matrices from the normals, so the conversion
  // does not store them (frames share the batch's calculators, so one frame's model is all's)
  spd.skip_omega = b->frames[0].omega_closed ? 1 : 0;
  auto enqueue = [&]() {
    return launch_convert(&b->ctx->prof, b->ctx->stream, b->d_frames, b->num_frames, b->d_blk_frame, b->total_tiles,
                          b->d_xblk_frame, b->total_xblk, b->d_yblk_frame, b->total_yblk, spd, b->frames[0].model,
                          b->d_frame_order);
  };
  if (b->ctx->prof.on) {
    // per-launch timing runs eagerly
    CHK(enqueue());
  } else {
    // the conversion's launches depend only on the batch and the FP64 eigen setting: one graph
    std::vector<int> key = {0};
    memcpy(&key[0], &spd.fp64_gap, sizeof(float));
    CHK(graph_replay(b->ctx, b->conv_graph, key, enqueue));
  }
  b->converted = true;
  b->omega_deferred = spd.skip_omega != 0;
  b->omega_filled.assign(b->num_frames, 0);
  return NICP_OK;
}

namespace {
// One batched AlignerProjective::align call in three stages, so that several calls can be
// prepared before any of them runs (the learned aligner prepares its start probe and its
// refinement before the network): batch_setup validates, plans, uploads the tables and the
// initial state into the buffers B and picks the projection strategy; batch_run launches
// (eagerly, timing the strategies, or from the batch's graph); batch_collect copies the
// transforms and results back. with_stats = false skips the statistics linearisation (results
// then hold the in-loop linearisation only and no solution statistics); raw, when not null,
// receives the kResultDoubles result slots per pair; sub > 1 linearises over every sub-th row and
// column only (the start probe).
struct BatchCall {
  PairPlan plan;
  AlignBuffers B{};
  int *d_bp = nullptr, *d_br = nullptr, *d_bc = nullptr;
  PairDev *d_pairs = nullptr;
  nicp_align_params ap{};
  bool with_stats = false, fused = false, tuning = false;
  int num_pairs = 0, sub = 1, zstrategy = 0;
};

AlignBuffers batch_buffers(nicp_batch *b) {
  return AlignBuffers{&b->pairs_tbl, &b->maps, &b->zbuf, &b->cur_index, &b->pairs_buf, &b->state, &b->partials,
                      &b->result, &b->sbox};
}
// the start probe's (subsampled, so never tile-binned: no source-tile boxes)
AlignBuffers probe_buffers(nicp_batch *b) {
  return AlignBuffers{&b->pr_pairs_tbl, &b->pr_maps,     &b->pr_zbuf,   &b->pr_cur_index, &b->pr_pairs_buf,
                      &b->pr_state,     &b->pr_partials, &b->pr_result, nullptr};
}

int batch_plan(nicp_batch *b, const AlignBuffers &B, int num_pairs, const int *ref_frame, const int *cur_frame,
               const nicp_align_params *ap, bool with_stats, int sub, int twin, BatchCall &c) {
  if (!b || num_pairs <= 0 || !ref_frame || !cur_frame) return set_err(NICP_ERR_INVALID, "bad arguments");
  if (!b->converted) return set_err(NICP_ERR_STATE, "the batch's frames have not been converted (nicp_batch_convert)");
  CHK(check_align_params(ap));
  for (int i = 0; i < num_pairs; ++i)
    if (ref_frame[i] < 0 || ref_frame[i] >= b->num_frames || cur_frame[i] < 0 || cur_frame[i] >= b->num_frames)
      return set_err(NICP_ERR_INVALID, "pair %d: frame out of range", i);
  nicp_ctx *ctx = b->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  std::vector<nicp_camera> camv(num_pairs);
  std::vector<const nicp_camera *> cams(num_pairs);
  for (int i = 0; i < num_pairs; ++i) {
    const FrameDev &fc = b->frames[cur_frame[i]];
    // the projector of a pair is the current frame's camera
    nicp_camera &c = camv[i];
    for (int k = 0; k < 9; ++k) c.K[k] = fc.K[k];
    c.min_distance = fc.min_d;
    c.max_distance = fc.max_d;
    c.model = fc.model;
    c.horizontal_fov = fc.sph.hfov;
    c.vertical_fov = fc.sph.vfov;
    c.horizontal_center = fc.sph.hc;
    c.vertical_center = fc.sph.vc;
    cams[i] = &c;
  }
  if (sub < 1 || (sub > 1 && (ap->inner_iterations > 1 || with_stats)))
    return set_err(NICP_ERR_INVALID, "a subsampled linearisation runs one inner iteration and no statistics");
  PairPlan &plan = c.plan;
  make_plan(b->frames, num_pairs, ref_frame, cur_frame, cams.data(), plan);
  plan_batch_frames(b->frames, plan);
  if (sub > 1) {
    // per-pixel blocks cover the subsampled grid only (k_zfill clears its cells only)
    for (int i = 0; i < num_pairs; ++i) plan.pairs[i].sub = sub;
    plan_pixel_tiles(plan);
    // the twin pass needs pairs q and q + twin on the same frames (pinhole)
    bool twins = twin > 0 && num_pairs == 2 * twin && plan.pairs[0].model != NICP_SPHERICAL;
    for (int i = 0; twins && i < twin; ++i)
      twins = ref_frame[i] == ref_frame[i + twin] && cur_frame[i] == cur_frame[i + twin];
    plan.twin = twins ? twin : 0;
  }
  // the tile-binned projection needs pinhole pairs whose current index image is the frame's own,
  // reference frames with source-tile depth ranges and the full projector image
  plan.fused_ok = sub == 1;
  for (int i = 0; plan.fused_ok && i < num_pairs; ++i) {
    const PairDev &pd = plan.pairs[i];
    plan.fused_ok = pd.model == NICP_PINHOLE && pd.cur_own && b->frames[pd.ref].stile_d && b->frames[pd.ref].spix &&
                    b->frames[pd.ref].index;
  }
  const bool fused = plan.fused_ok && ap->row_search_region == 0 && ap->col_search_region == 0;
  plan.list_cap = b->list_cap;
  c.B = B;
  // strategy: forced, or the autotuner's (the first call of a shape times the applicable
  // strategies on alternate iterations, eagerly; later calls use the fastest), or by default
  // the tile-binned projection where it applies and its fused tiles fill the device (one
  // workgroup per 4096 pixels: a single 640x480 pair has 80, and then the per-launch latency
  // of the global z-buffer kernels is lower)
  int zstrategy = fused && plan.total_ftiles >= kFusedMinTiles ? 2 : 0;
  bool tuning = false;
  if (b->projection >= 0) {
    zstrategy = b->projection == 2 && !fused ? 0 : b->projection;
  } else if (b->autotune) {
    const std::vector<int> skey = {num_pairs, plan.total_pix_tiles, plan.total_rtiles, plan.total_ctiles, sub,
                                   with_stats ? 1 : 0, fused ? 1 : 0};
    size_t shape = 0;
    while (shape < b->tuned_keys.size() && b->tuned_keys[shape] != skey) ++shape;
    if (shape == b->tuned_keys.size()) {
      tuning = ap->outer_iterations >= 2;
      if (!tuning) {  // nothing to compare: keep the default
        b->tuned_keys.push_back(skey);
        b->tuned_choice.push_back(zstrategy);
        for (int k = 0; k < kNumStrategies; ++k) b->tuned_ms.push_back(-1.0f);
      }
    } else {
      zstrategy = b->tuned_choice[shape];
    }
  }
  c.ap = *ap;
  c.with_stats = with_stats;
  c.fused = fused;
  c.tuning = tuning;
  c.num_pairs = num_pairs;
  c.sub = sub;
  c.zstrategy = zstrategy;
  return NICP_OK;
}

// the buffers, tables and initial state of a planned call (staged: asynchronous copies)
int batch_prepare(nicp_batch *b, BatchCall &c, const float *init_guesses16, PinnedStage *stage) {
  return prepare_align(b->ctx->stream, c.plan, c.B, init_guesses16, &c.d_bp, &c.d_br, &c.d_bc, &c.d_pairs, stage);
}

void key_ptr(std::vector<int> &key, const void *p) {
  key.push_back((int)(uint32_t)((uintptr_t)p));
  key.push_back((int)(uint32_t)((uint64_t)(uintptr_t)p >> 32));
}

// The graph key of a planned call: the launch geometry, the parameters baked into the kernel
// arguments, and every device address a captured launch takes as an argument (the tables'
// contents are uploaded before each launch, so only their addresses are baked into a graph).
std::vector<int> call_key(const nicp_batch *b, const BatchCall &c) {
  const PairPlan &plan = c.plan;
  std::vector<int> key = {c.num_pairs, plan.total_pix_tiles, plan.total_rtiles, plan.total_ctiles, plan.total_dtiles};
  const int *apw = (const int *)&c.ap;
  key.insert(key.end(), apw, apw + sizeof(nicp_align_params) / sizeof(int));
  key.push_back(plan.total_sblk);
  key.push_back(plan.total_ftiles);
  key.push_back(plan.list_cap);
  key.push_back(plan.twin);
  for (const void *ptr : {(const void *)c.d_pairs, (const void *)c.d_bp, (const void *)c.d_br, (const void *)c.d_bc,
                          (const void *)plan.d_dpair, (const void *)plan.d_spair, (const void *)plan.d_fpair,
                          (const void *)b->d_frames})
    key_ptr(key, ptr);
  key.push_back(b->ctx->prof.on ? 1 : 0);
  key.push_back(c.with_stats ? 1 : 0);
  key.push_back(c.sub);
  key.push_back(c.zstrategy);
  return key;
}

int batch_run(nicp_batch *b, BatchCall &c, int use_graph, GraphCache &graph) {
  nicp_ctx *ctx = b->ctx;
  PairPlan &plan = c.plan;
  const nicp_align_params *ap = &c.ap;
  const bool with_stats = c.with_stats, fused = c.fused;
  const int num_pairs = c.num_pairs, sub = c.sub, zstrategy = c.zstrategy;
  int *d_bp = c.d_bp, *d_br = c.d_br, *d_bc = c.d_bc;
  PairDev *d_pairs = c.d_pairs;
  if (c.tuning) {
    ZTune tune;
    ctx->prof.sink = &ctx->prof.eager;
    int rc = launch_align(&ctx->prof, ctx->stream, plan, d_pairs, b->d_frames, d_bp, d_br, d_bc, ap, with_stats, 0,
                          &tune);
    hipError_t e = rc == NICP_OK ? hipStreamSynchronize(ctx->stream) : hipSuccess;
    double sum[kNumStrategies] = {0.0, 0.0, 0.0};
    int cnt[kNumStrategies] = {0, 0, 0};
    for (size_t i = 0; rc == NICP_OK && e == hipSuccess && i < tune.strategy.size(); ++i) {
      float ms = 0.0f;
      e = hipEventElapsedTime(&ms, tune.ev[2 * i], tune.ev[2 * i + 1]);
      // the first iteration also warms caches and code: it does not vote
      if (i > 0) {
        sum[tune.strategy[i]] += ms;
        ++cnt[tune.strategy[i]];
      }
    }
    tune.release();
    if (rc != NICP_OK) return rc;
    if (e != hipSuccess) return set_err(NICP_ERR_HIP, "autotune timing: %s", hipGetErrorString(e));
    b->tuned_keys.push_back({num_pairs, plan.total_pix_tiles, plan.total_rtiles, plan.total_ctiles, sub,
                             with_stats ? 1 : 0, fused ? 1 : 0});
    int best = -1;
    float best_ms = 0.0f;
    for (int k = 0; k < kNumStrategies; ++k) {
      const float m = cnt[k] ? (float)(sum[k] / cnt[k]) : -1.0f;
      b->tuned_ms.push_back(m);
      if (m >= 0.0f && (best < 0 || m < best_ms)) {
        best = k;
        best_ms = m;
      }
    }
    b->tuned_choice.push_back(best >= 0 ? best : zstrategy);
  } else if (use_graph && !ctx->prof.on) {
    CHK(graph_replay(ctx, graph, call_key(b, c), [&]() {
      return launch_align(&ctx->prof, ctx->stream, plan, d_pairs, b->d_frames, d_bp, d_br, d_bc, ap, with_stats,
                          zstrategy);
    }));
  } else {
    ctx->prof.sink = &ctx->prof.eager;
    CHK(launch_align(&ctx->prof, ctx->stream, plan, d_pairs, b->d_frames, d_bp, d_br, d_bc, ap, with_stats, zstrategy));
  }
  return NICP_OK;
}

int batch_collect(nicp_batch *b, BatchCall &c, float *T_out16, nicp_align_result *results, std::vector<double> *raw) {
  nicp_ctx *ctx = b->ctx;
  const nicp_align_params *ap = &c.ap;
  const int num_pairs = c.num_pairs;
  const bool with_stats = c.with_stats;
  if (!T_out16 && !results && !raw) return NICP_OK;
  std::vector<float> st((size_t)num_pairs * kStateFloats);
  std::vector<double> local;
  std::vector<double> &rr = raw ? *raw : local;
  rr.assign((size_t)num_pairs * kResultDoubles, 0.0);
  HIPCHK(hipMemcpyAsync(st.data(), c.B.state->p, st.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  if ((results || raw) && ap->outer_iterations > 0)
    HIPCHK(hipMemcpyAsync(rr.data(), c.B.result->p, rr.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  if (ctx->prof.on) {
    ctx->prof.read(ctx->prof.eager, true);
    ctx->read_pending();
  }
  nicp_align_params ap_res = *ap;
  if (!with_stats) ap_res.outer_iterations = 0;  // no statistics were computed
  for (int i = 0; i < num_pairs; ++i) {
    const float *T = &st[(size_t)i * kStateFloats];
    if (T_out16) memcpy(T_out16 + 16 * (size_t)i, T, 16 * sizeof(float));
    if (results) fill_result(&rr[(size_t)i * kResultDoubles], &ap_res, &results[i], ms);
  }
  return NICP_OK;
}
}  // namespace


int batch_align(nicp_batch *b, int num_pairs, const int *ref_frame, const int *cur_frame, const float *init_guesses16,
                const nicp_align_params *ap, int use_graph, float *T_out16, nicp_align_result *results,
                bool with_stats, std::vector<double> *raw, int sub = 1, int twin = 0) {
  BatchCall c;
  CHK(batch_plan(b, batch_buffers(b), num_pairs, ref_frame, cur_frame, ap, with_stats, sub, twin, c));
  CHK(b->stage.begin(prepare_stage_bytes(c.plan)));
  CHK(batch_prepare(b, c, init_guesses16, &b->stage));
  CHK(b->stage.end(b->ctx->stream));
  HIPCHK(hipEventRecord(b->ctx->ev0, b->ctx->stream));
  CHK(batch_run(b, c, use_graph, b->graph));
  HIPCHK(hipEventRecord(b->ctx->ev1, b->ctx->stream));
  return batch_collect(b, c, T_out16, results, raw);
}

extern "C" int nicp_batch_align(nicp_batch *b, int num_pairs, const int *ref_frame, const int *cur_frame,
                                const float *init_guesses16, const nicp_align_params *ap, int use_graph, float *T_out16,
                                nicp_align_result *results) {
  return batch_align(b, num_pairs, ref_frame, cur_frame, init_guesses16, ap, use_graph, T_out16, results, true,
                     nullptr);
}

extern "C" int nicp_batch_set_autotune(nicp_batch *b, int enable) {
  if (!b) return set_err(NICP_ERR_INVALID, "batch is null");
  b->autotune = enable != 0;
  b->tuned_keys.clear();
  b->tuned_choice.clear();
  b->tuned_ms.clear();
  return NICP_OK;
}

extern "C" int nicp_batch_autotune_read(nicp_batch *b, int capacity, int *num_shapes, int *choice, float *ms) {
  if (!b || !num_shapes) return set_err(NICP_ERR_INVALID, "null argument");
  const int n = (int)b->tuned_choice.size();
  *num_shapes = n;
  for (int i = 0; i < n && i < capacity; ++i) {
    if (choice) choice[i] = b->tuned_choice[i];
    if (ms)
      for (int k = 0; k < kNumStrategies; ++k) ms[kNumStrategies * i + k] = b->tuned_ms[kNumStrategies * i + k];
  }
  return NICP_OK;
}

extern "C" int nicp_batch_set_bin_capacity(nicp_batch *b, int capacity) {
  if (!b) return set_err(NICP_ERR_INVALID, "batch is null");
  if (capacity < 0 || capacity > kListCap) return set_err(NICP_ERR_INVALID, "capacity must lie in [0, %d]", kListCap);
  b->list_cap = capacity;
  return NICP_OK;
}

extern "C" int nicp_batch_set_projection(nicp_batch *b, int strategy) {
  if (!b) return set_err(NICP_ERR_INVALID, "batch is null");
  if (strategy < -1 || strategy >= kNumStrategies) return set_err(NICP_ERR_INVALID, "unknown strategy %d", strategy);
  b->projection = strategy;
  return NICP_OK;
}

extern "C" int nicp_batch_frame_cloud(nicp_batch *b, int frame, nicp_cloud **out) {
  if (!b || !out || frame < 0 || frame >= b->num_frames) return set_err(NICP_ERR_INVALID, "bad arguments");
  if (!b->converted) return set_err(NICP_ERR_STATE, "the batch's frames have not been converted (nicp_batch_convert)");
  nicp_cloud *cl = new (std::nothrow) nicp_cloud();
  if (!cl) return set_err(NICP_ERR_OOM, "out of host memory");
  HIPCHK(hipSetDevice(b->ctx->device));
  if (b->omega_deferred && !b->omega_filled[frame]) {
    // the frame's stored information matrices, from its normals (the closed form the batch's
    // linearisers use)
    const int cap = b->frames[frame].rows * b->frames[frame].cols;
    LAUNCH((Prof *)nullptr, 0, k_omega_fill, dim3((cap + kThreads - 1) / kThreads), dim3(kThreads), b->ctx->stream,
           b->d_frames + frame);
    b->omega_filled[frame] = 1;
  }
  int n = 0;
  HIPCHK(hipMemcpyAsync(&n, b->frames[frame].npoints, sizeof(int), hipMemcpyDeviceToHost, b->ctx->stream));
  HIPCHK(hipStreamSynchronize(b->ctx->stream));
  cl->ctx = b->ctx;
  cl->fd = b->frames[frame];
  cl->n = n;
  cl->has_index = 1;
  cl->has_stats = 0;
  cl->alloc = nullptr;
  cl->cap = b->frames[frame].rows * b->frames[frame].cols;
  *out = cl;
  return NICP_OK;
}

// ---------------------------------------------------------------------------------------
// per-kernel timing
// ---------------------------------------------------------------------------------------
extern "C" int nicp_kernel_count(void) { return KID_N; }

extern "C" const char *nicp_kernel_name(int id) { return (id >= 0 && id < KID_N) ? kKernelNames[id] : ""; }

extern "C" int nicp_ctx_profile(nicp_ctx *ctx, int enable) {
  if (!ctx) return set_err(NICP_ERR_INVALID, "ctx is null");
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->prof.read(ctx->prof.eager, true);
  ctx->prof.on = enable != 0;
  ctx->prof.err = hipSuccess;
  ctx->prof.err_where = "";
  for (int i = 0; i < KID_N; ++i) {
    ctx->prof.total_ms[i] = 0.0;
    ctx->prof.launches[i] = 0;
  }
  ctx->pending_graphs.clear();
  return NICP_OK;
}

extern "C" int nicp_ctx_profile_read(nicp_ctx *ctx, double *total_ms, long long *launches) {
  if (!ctx) return set_err(NICP_ERR_INVALID, "ctx is null");
  if (ctx->prof.err != hipSuccess)
    return set_err(NICP_ERR_HIP, "kernel timing failed in %s: %s", ctx->prof.err_where, hipGetErrorString(ctx->prof.err));
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->prof.read(ctx->prof.eager, true);
  ctx->read_pending();
  for (int i = 0; i < KID_N; ++i) {
    if (total_ms) total_ms[i] = ctx->prof.total_ms[i];
    if (launches) launches[i] = ctx->prof.launches[i];
  }
  return NICP_OK;
}

// ---------------------------------------------------------------------------------------
// kNN correspondences and AlignerNN (correspondencefindernn.cpp, alignernn.cpp)
// ---------------------------------------------------------------------------------------
namespace {

constexpr float kDefaultCell = 0.02f;
constexpr long long kMaxCells = 1ll << 25;        // cells of one pair's grid
constexpr long long kMaxArenaCells = 1ll << 26;   // cells (and sentinels) of all pairs: the scan's reach
constexpr int kMaxNNPairs = 1 << 16;               // leaves every pair a share of >= 1023 cells

// decodes k_nn_bbox's order-preserving keys; an empty or non-finite box becomes the origin
void decode_bbox(const unsigned int *bb, int n, float lo[3], float hi[3]) {
  for (int k = 0; k < 3; ++k) {
    unsigned int a = bb[k], b = bb[3 + k];
    a = (a & 0x80000000u) ? (a & 0x7fffffffu) : ~a;
    b = (b & 0x80000000u) ? (b & 0x7fffffffu) : ~b;
    lo[k] = __builtin_bit_cast(float, a);
    hi[k] = __builtin_bit_cast(float, b);
    if (n == 0 || !std::isfinite(lo[k]) || !std::isfinite(hi[k])) { lo[k] = 0.0f; hi[k] = 0.0f; }
  }
}

// exclusive scan of the m ints at cells (m <= kMaxArenaCells), copied to cursor as well
int scan_arena(nicp_ctx *ctx, int *cells, int m, int *cursor, int *sums, int *total) {
  const int sb = (m + 4095) / 4096;
  if (sb > 16384) return set_err(NICP_ERR_INVALID, "grid too large");
  hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(k_scan_local, dim3(sb), dim3(1024), 0, s, cells, m, sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, sums, sb, total);
  hipLaunchKernelGGL(k_scan_add, dim3(sb), dim3(1024), 0, s, cells, m, (const int *)sums, cursor);
  return NICP_OK;
}

// Uniform grids over the current clouds of np pairs (pd.cur of the plan's pairs): one cell
// arena holding each pair's cell starts plus a sentinel (an exclusive scan of counts over
// the whole arena, so every pair's sorted points follow the previous pair's), and the clouds
// sorted by cell. The nearest neighbour the query finds is exact at any cell size, so a pair
// whose grid would exceed its share of the arena gets coarser cells and the same result, and
// a call may search with either of two grids: the fine one (cell) and, when asked for, a
// coarse one (kCoarseCells x cell) for searches over a wide radius (the first iteration
// from an initial guess). Also the query order: each pair's reference indices sorted by the
// Morton code of a coarse cubic grid over the reference cloud (ctx->nn_order, pair p from
// grids[p].order_base).
constexpr float kCoarseCells = 3.0f;

struct NNGrids {
  const GridDev *fine = nullptr, *coarse = nullptr;  // device tables of np grids
  const int *cells_fine = nullptr, *cells_coarse = nullptr;
  const float4 *sorted_fine = nullptr, *sorted_coarse = nullptr;
};

int build_grids(nicp_ctx *ctx, int np, const nicp_cloud *const *refs, const nicp_cloud *const *curs,
                const PairDev *d_pairs, const FrameDev *d_frames, float ns, float cell, bool with_coarse,
                NNGrids &out) {
  hipStream_t s = ctx->stream;
  const size_t bbox_bytes = align_up((size_t)np * 12 * sizeof(unsigned int), 256);
  const size_t grid_bytes = align_up((size_t)np * sizeof(GridDev), 256);
  CHK(ctx->grid_aux.reserve(bbox_bytes + 3 * grid_bytes + 64 * sizeof(int) + 16384 * sizeof(int)));
  unsigned int *d_bbox = ctx->grid_aux.as<unsigned int>();
  GridDev *grids_d = (GridDev *)(ctx->grid_aux.as<char>() + bbox_bytes);
  GridDev *cgrids_d = (GridDev *)(ctx->grid_aux.as<char>() + bbox_bytes + grid_bytes);
  GridDev *ogrids_d = (GridDev *)(ctx->grid_aux.as<char>() + bbox_bytes + 2 * grid_bytes);
  int *total = (int *)(ctx->grid_aux.as<char>() + bbox_bytes + 3 * grid_bytes);
  int *sums = total + 64;
  std::vector<unsigned int> bb((size_t)np * 12);
  int nmax = 1;
  size_t ncur = 0, nref = 0;
  for (int i = 0; i < np; ++i) {
    for (int w = 0; w < 2; ++w)
      for (int k = 0; k < 3; ++k) {
        bb[12 * i + 6 * w + k] = 0xffffffffu;
        bb[12 * i + 6 * w + 3 + k] = 0u;
      }
    nmax = std::max(nmax, std::max(curs[i]->n, refs[i]->n));
    ncur += (size_t)curs[i]->n;
    nref += (size_t)refs[i]->n;
  }
  HIPCHK(hipMemcpyAsync(d_bbox, bb.data(), bb.size() * sizeof(unsigned int), hipMemcpyHostToDevice, s));
  const int pb = (nmax + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(k_nn_bbox, dim3(std::min(64, pb), np, 2), dim3(kThreads), 0, s, d_pairs, d_frames, d_bbox);
  HIPCHK(hipMemcpyAsync(bb.data(), d_bbox, bb.size() * sizeof(unsigned int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const long long share = std::min(kMaxCells, kMaxArenaCells / np - 1);
  // Morton bits per axis of the order grids: 2^(3 bits) cells per pair within the arena
  int obits = 7;
  while (obits > 1 && (long long)np << (3 * obits) > kMaxArenaCells - 1) --obits;
  std::vector<GridDev> ogrids(np);
  long long oarena = 0;
  size_t obase = 0;
  std::vector<int> order_base(np);
  for (int i = 0; i < np; ++i) {
    order_base[i] = (int)obase;
    obase += (size_t)refs[i]->n;
    // order grid: cubic cells over the reference box
    float lo[3], hi[3];
    decode_bbox(&bb[12 * i + 6], refs[i]->n, lo, hi);
    double ext = 0.0;
    for (int k = 0; k < 3; ++k) ext = std::max(ext, (double)hi[k] - (double)lo[k]);
    GridDev &o = ogrids[i];
    o.ox = lo[0]; o.oy = lo[1]; o.oz = lo[2];
    const double oh = std::max(ext, 1e-30) / (double)(1 << obits) * (1.0 + 1e-6);
    o.h = (float)oh;
    o.inv_h = std::isfinite(1.0 / oh) ? (float)(1.0 / oh) : 0.0f;
    o.nx = obits;
    o.ny = o.nz = 0;
    o.cell_base = (int)oarena;
    o.order_base = 0;
    oarena += 1ll << (3 * obits);
  }
  // grids of the current clouds at cell edge h0 into (cells, sorted, table)
  auto cur_grid = [&](float h0, DevBuf &cells_buf, DevBuf &sorted_buf, GridDev *table) -> int {
    std::vector<GridDev> grids(np);
    long long arena = 0;
    for (int i = 0; i < np; ++i) {
      float lo[3], hi[3];
      decode_bbox(&bb[12 * i], curs[i]->n, lo, hi);
      float h = h0;
      long long cells = 0;
      int dims[3];
      // extents in double: a float difference of extreme coordinates can overflow
      for (;;) {
        double c = 1.0;
        for (int k = 0; k < 3; ++k) {
          const double d = std::floor(((double)hi[k] - (double)lo[k]) / (double)h) + 1.0;
          dims[k] = (int)std::min(d, 2e9);
          c *= d;
        }
        if (c <= (double)share) {
          cells = (long long)c;
          break;
        }
        h *= 1.25f;
      }
      GridDev &g = grids[i];
      g.ox = lo[0]; g.oy = lo[1]; g.oz = lo[2];
      g.h = h;
      g.inv_h = 1.0f / h;
      g.nx = dims[0]; g.ny = dims[1]; g.nz = dims[2];
      g.cell_base = (int)arena;
      g.order_base = order_base[i];
      arena += cells + 1;
    }
    const int m = (int)arena;
    CHK(cells_buf.reserve((size_t)m * sizeof(int)));
    CHK(ctx->grid_cursor.reserve((size_t)m * sizeof(int)));
    CHK(sorted_buf.reserve(std::max<size_t>(ncur, 1) * 2 * sizeof(float4)));
    HIPCHK(hipMemcpyAsync(table, grids.data(), (size_t)np * sizeof(GridDev), hipMemcpyHostToDevice, s));
    int *cells_d = cells_buf.as<int>();
    HIPCHK(hipMemsetAsync(cells_d, 0, (size_t)m * sizeof(int), s));
    hipLaunchKernelGGL(k_nn_count, dim3(pb, np), dim3(kThreads), 0, s, d_pairs, d_frames, (const GridDev *)table,
                       cells_d);
    CHK(scan_arena(ctx, cells_d, m, ctx->grid_cursor.as<int>(), sums, total));
    hipLaunchKernelGGL(k_nn_scatter, dim3(pb, np), dim3(kThreads), 0, s, d_pairs, d_frames, (const GridDev *)table,
                       ns, ctx->grid_cursor.as<int>(), sorted_buf.as<float4>());
    return NICP_OK;
  };
  // query order first (its arena is then reused by the grids' scans)
  const int mo = (int)oarena + 1;
  CHK(ctx->grid_cells.reserve((size_t)mo * sizeof(int)));
  CHK(ctx->grid_cursor.reserve((size_t)mo * sizeof(int)));
  CHK(ctx->nn_order.reserve(std::max<size_t>(nref, 1) * sizeof(int)));
  HIPCHK(hipMemcpyAsync(ogrids_d, ogrids.data(), (size_t)np * sizeof(GridDev), hipMemcpyHostToDevice, s));
  int *ocells = ctx->grid_cells.as<int>();
  HIPCHK(hipMemsetAsync(ocells, 0, (size_t)mo * sizeof(int), s));
  hipLaunchKernelGGL(k_nn_order_count, dim3(pb, np), dim3(kThreads), 0, s, d_pairs, d_frames,
                     (const GridDev *)ogrids_d, ocells);
  CHK(scan_arena(ctx, ocells, mo, ctx->grid_cursor.as<int>(), sums, total));
  hipLaunchKernelGGL(k_nn_order_scatter, dim3(pb, np), dim3(kThreads), 0, s, d_pairs, d_frames,
                     (const GridDev *)ogrids_d, ctx->grid_cursor.as<int>(), ctx->nn_order.as<int>());
  const float h0 = cell > 0.0f ? cell : kDefaultCell;
  CHK(cur_grid(h0, ctx->grid_cells, ctx->grid_sorted, grids_d));
  out.fine = grids_d;
  out.cells_fine = ctx->grid_cells.as<int>();
  out.sorted_fine = ctx->grid_sorted.as<float4>();
  out.coarse = grids_d;
  out.cells_coarse = out.cells_fine;
  out.sorted_coarse = out.sorted_fine;
  if (with_coarse) {
    CHK(cur_grid(h0 * kCoarseCells, ctx->grid_cells_c, ctx->grid_sorted_c, cgrids_d));
    out.coarse = cgrids_d;
    out.cells_coarse = ctx->grid_cells_c.as<int>();
    out.sorted_coarse = ctx->grid_sorted_c.as<float4>();
  }
  HIPCHK(hipGetLastError());
  return NICP_OK;
}

// plan of np NN pairs: pair i's "image" is 1 x N_ref(i); frames 2i and 2i+1 are its clouds
int nn_plan(nicp_ctx *ctx, int np, const nicp_cloud *const *refs, const nicp_cloud *const *curs, const float *init16,
            PairPlan &plan, FrameDev **d_frames, PairDev **d_pairs, int **d_bp) {
  std::vector<FrameDev> frames(2 * (size_t)np);
  std::vector<int> ri(np), ci(np);
  nicp_camera cam{};
  cam.K[0] = cam.K[4] = cam.K[8] = 1.0f;
  cam.min_distance = cam.max_distance = 1.0f;
  std::vector<const nicp_camera *> cams(np, &cam);
  for (int i = 0; i < np; ++i) {
    frames[2 * i] = refs[i]->fd;
    frames[2 * i + 1] = curs[i]->fd;
    frames[2 * i].rows = frames[2 * i + 1].rows = 1;
    frames[2 * i].cols = frames[2 * i + 1].cols = std::max(1, refs[i]->n);
    ri[i] = 2 * i;
    ci[i] = 2 * i + 1;
  }
  CHK(upload_frames(ctx, frames, d_frames));
  make_plan(frames, np, ri.data(), ci.data(), cams.data(), plan);
  plan.blk_ref.clear();
  plan.blk_cur.clear();
  plan.total_rtiles = plan.total_ctiles = 0;
  AlignBuffers B = ctx_buffers(ctx);
  int *d_br, *d_bc;
  return prepare_align(ctx->stream, plan, B, init16, d_bp, &d_br, &d_bc, d_pairs);
}

int check_nn_pairs(nicp_ctx *ctx, int np, const nicp_cloud *const *refs, const nicp_cloud *const *curs) {
  if (!ctx || !refs || !curs) return set_err(NICP_ERR_INVALID, "null argument");
  if (np <= 0 || np > kMaxNNPairs) return set_err(NICP_ERR_INVALID, "num_pairs out of range");
  for (int i = 0; i < np; ++i) {
    if (!refs[i] || !curs[i]) return set_err(NICP_ERR_INVALID, "null cloud");
    if (refs[i]->ctx != ctx || curs[i]->ctx != ctx) return set_err(NICP_ERR_INVALID, "cloud of another context");
  }
  return NICP_OK;
}

// AlignerNN::align (alignernn.cpp:21-118) for np pairs in the same launches; T_out and the
// results hold one entry per pair, the time is the whole call's
int align_nn_pairs(nicp_ctx *ctx, int np, const nicp_cloud *const *refs, const nicp_cloud *const *curs,
                   const float *init16, const nicp_align_params *ap, float normal_scaling, float cell_size,
                   float *T_out, nicp_align_result *results) {
  CHK(check_nn_pairs(ctx, np, refs, curs));
  CHK(check_align_params(ap));
  if (ap->inner_iterations != 1) return set_err(NICP_ERR_INVALID, "the kNN aligner supports one inner iteration");
  HIPCHK(hipSetDevice(ctx->device));
  ctx->last_valid = false;  // the pairs buffer is rewritten
  PairPlan plan;
  FrameDev *d_frames;
  PairDev *d_pairs;
  int *d_bp;
  CHK(nn_plan(ctx, np, refs, curs, init16, plan, &d_frames, &d_pairs, &d_bp));
  hipStream_t s = ctx->stream;
  HIPCHK(hipEventRecord(ctx->ev0, s));
  NNGrids G;
  CHK(build_grids(ctx, np, refs, curs, d_pairs, d_frames, normal_scaling, cell_size, ap->outer_iterations > 0, G));
  const AlignParamsDev a = to_dev(ap);
  Prof *pf = &ctx->prof;
  pf->sink = &pf->eager;
  const dim3 tb(kThreads);
  const dim3 ntiles(plan.total_pix_tiles);
  for (int it = 0; it < ap->outer_iterations; ++it) {
    // the first iteration searches from the initial guess, usually over a wide radius
    const bool coarse = it == 0;
    LAUNCH(pf, KID_NN_QUERY, k_nn_query, ntiles, tb, s, d_pairs, d_frames, d_bp, coarse ? G.coarse : G.fine,
           coarse ? G.cells_coarse : G.cells_fine, coarse ? G.sorted_coarse : G.sorted_fine,
           (const int *)ctx->nn_order.as<int>(), a, normal_scaling, (int *)nullptr);
    LAUNCH(pf, KID_LIN_PAIRS, k_lin_pairs, ntiles, tb, s, d_pairs, d_frames, d_bp, a, kSlotTinv, 0);
    LAUNCH(pf, KID_SOLVE, k_solve, dim3(np), dim3(kSolveThreads), s, d_pairs, ap->lambda, 0, 0, 0);
  }
  if (ap->outer_iterations > 0) {
    LAUNCH(pf, KID_LIN_PAIRS, k_lin_pairs, ntiles, tb, s, d_pairs, d_frames, d_bp, a, kSlotTinv, 0);
    LAUNCH(pf, KID_SOLVE_STATS, k_solve, dim3(np), dim3(kSolveThreads), s, d_pairs, 0.0f, 2, 0, 0);
    LAUNCH(pf, KID_SOLUTION_STATS, k_solution_stats, dim3(np), dim3(64), s, d_pairs,
           ap->translational_min_eigen_ratio, ap->rotational_min_eigen_ratio);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ctx->ev1, s));
  std::vector<float> st((size_t)np * kStateFloats);
  std::vector<double> r((size_t)np * kResultDoubles);
  HIPCHK(hipMemcpyAsync(st.data(), ctx->state.p, st.size() * sizeof(float), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(r.data(), ctx->result.p, r.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  if (ap->outer_iterations == 0) std::fill(r.begin(), r.end(), 0.0);
  for (int i = 0; i < np; ++i) {
    const float *T = &st[(size_t)i * kStateFloats + kSlotT];
    if (T_out) memcpy(T_out + (size_t)i * 16, T, 16 * sizeof(float));
    if (results) fill_result(&r[(size_t)i * kResultDoubles], ap, &results[i], ms);
  }
  return NICP_OK;
}

}  // namespace

extern "C" int nicp_nn_correspondences(nicp_ctx *ctx, const nicp_cloud *ref, const nicp_cloud *cur, const float T[16],
                                       const nicp_align_params *ap, float normal_scaling, float cell_size, int *pairs,
                                       int *num_pairs, int *nn_out) {
  CHK(check_nn_pairs(ctx, 1, &ref, &cur));
  if (!T || !num_pairs) return set_err(NICP_ERR_INVALID, "null argument");
  CHK(check_align_params(ap));
  HIPCHK(hipSetDevice(ctx->device));
  ctx->last_valid = false;  // the pairs buffer is rewritten
  PairPlan plan;
  FrameDev *d_frames;
  PairDev *d_pairs;
  int *d_bp;
  CHK(nn_plan(ctx, 1, &ref, &cur, nullptr, plan, &d_frames, &d_pairs, &d_bp));
  HIPCHK(hipMemcpyAsync(ctx->state.as<float>() + kSlotTinv, T, 16 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  NNGrids G;
  CHK(build_grids(ctx, 1, &ref, &cur, d_pairs, d_frames, normal_scaling, cell_size, false, G));
  const int nr = std::max(ref->n, 1);
  CHK(ctx->nn_out.reserve((size_t)nr * sizeof(int)));
  hipLaunchKernelGGL(k_nn_query, dim3(plan.total_pix_tiles), dim3(kThreads), 0, ctx->stream, d_pairs, d_frames, d_bp,
                     G.fine, G.cells_fine, G.sorted_fine, (const int *)ctx->nn_order.as<int>(), to_dev(ap),
                     normal_scaling, ctx->nn_out.as<int>());
  HIPCHK(hipGetLastError());
  std::vector<int2> pp(nr);
  HIPCHK(hipMemcpyAsync(pp.data(), ctx->pairs_buf.p, (size_t)nr * sizeof(int2), hipMemcpyDeviceToHost, ctx->stream));
  if (nn_out && ref->n)
    HIPCHK(hipMemcpyAsync(nn_out, ctx->nn_out.p, (size_t)ref->n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  int n = 0;
  for (int i = 0; i < ref->n; ++i)
    if (pp[i].x >= 0) {
      if (pairs) {
        pairs[2 * n] = pp[i].x;
        pairs[2 * n + 1] = pp[i].y;
      }
      ++n;
    }
  *num_pairs = n;
  return NICP_OK;
}

extern "C" int nicp_align_nn(nicp_ctx *ctx, const nicp_cloud *ref, const nicp_cloud *cur, const float init_guess[16],
                             const nicp_align_params *ap, float normal_scaling, float cell_size, float T_out[16],
                             nicp_align_result *result) {
  CHK(align_nn_pairs(ctx, 1, &ref, &cur, init_guess, ap, normal_scaling, cell_size, T_out, result));
  // the pair's correspondences stay in the pairs buffer at offset 0 (nicp_last_correspondences)
  ctx->last_rows = 1;
  ctx->last_cols = std::max(1, ref->n);
  ctx->last_valid = ap->outer_iterations > 0;
  return NICP_OK;
}

extern "C" int nicp_align_nn_batch(nicp_ctx *ctx, int num_pairs, const nicp_cloud *const *refs,
                                   const nicp_cloud *const *curs, const float *init_guesses,
                                   const nicp_align_params *ap, float normal_scaling, float cell_size, float *T_out,
                                   nicp_align_result *results) {
  return align_nn_pairs(ctx, num_pairs, refs, curs, init_guesses, ap, normal_scaling, cell_size, T_out, results);
}

// ---------------------------------------------------------------------------------------
// learned initial guess and the synthetic-scene renderer
// ---------------------------------------------------------------------------------------
#include "initnet.hip"
#include "synth_gpu.hip"
// ---------------------------------------------------------------------------------------
// Gaussians, cloud transform / add and the Merger
// ---------------------------------------------------------------------------------------
#include "merge.hip"
