// Device renderer of the seeded synthetic scenes (synth_scene.h) into the depth
// images of a batch: the training-set generator of the learned initial guess and
// the fast path of the benchmark's input generation. No counterpart in the
// reference. The per-pixel arithmetic is the host generator's (synth.cpp), in
// double precision with contraction off, so both produce the same images.
//
// This file is part of the nicp_capi.hip translation unit.
#include "synth_scene.h"

namespace {

struct RenderJob {
  uint16_t *depth;
  int rows, cols;
  double fx, fy, cx, cy;
  double R[9], t[3];
  uint64_t seed, stream;
  int prim_base, tile_base, nprims;
};

constexpr int kRenderTile = 1024;
static_assert(sizeof(synth::Prim) % sizeof(int) == 0, "Prim is copied as ints");

// one block per 1024-pixel tile of a job's image; the job's primitives are staged in LDS
__global__ __launch_bounds__(256) void k_synth_render(const RenderJob *__restrict__ jobs, const int *__restrict__ blk_job,
                                                      const synth::Prim *__restrict__ prims, double noise, double holes) {
  __shared__ synth::Prim sp[synth::kPrims];
  const int j = blk_job[blockIdx.x];
  const RenderJob &J = jobs[j];
  {
    const int *src = (const int *)(prims + J.prim_base);
    int *dst = (int *)sp;
    constexpr int words = synth::kPrims * (int)(sizeof(synth::Prim) / sizeof(int));
    for (int w = threadIdx.x; w < words; w += blockDim.x) dst[w] = src[w];
  }
  __syncthreads();
  const int P = J.rows * J.cols;
  const int base = (blockIdx.x - J.tile_base) * kRenderTile;
  for (int k = threadIdx.x; k < kRenderTile; k += blockDim.x) {
    const int p = base + k;
    if (p >= P) break;
    const int r = p / J.cols, c = p - r * J.cols;
    J.depth[p] = synth::depth_sample(sp, J.nprims, r, c, J.cols, J.fx, J.fy, J.cx, J.cy, J.R, J.t, noise, holes,
                                     J.seed, J.stream);
  }
}

}  // namespace

extern "C" int nicp_batch_render_synthetic(nicp_batch *b, int num_frames, const int *frames,
                                           const uint64_t *scene_seeds, const uint64_t *streams, const float *poses16,
                                           float noise_coeff, float hole_fraction) {
  if (!b || num_frames <= 0 || !frames || !scene_seeds || !streams || !poses16)
    return set_err(NICP_ERR_INVALID, "bad arguments");
  for (int i = 0; i < num_frames; ++i)
    if (frames[i] < 0 || frames[i] >= b->num_frames) return set_err(NICP_ERR_INVALID, "job %d: frame out of range", i);
  nicp_ctx *ctx = b->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  std::vector<RenderJob> jobs(num_frames);
  std::vector<synth::Prim> prims;
  prims.reserve((size_t)num_frames * synth::kPrims);
  std::vector<int> blk;
  std::vector<synth::Prim> one;
  for (int i = 0; i < num_frames; ++i) {
    FrameDev &f = b->frames[frames[i]];
    RenderJob &J = jobs[i];
    J.depth = (uint16_t *)f.depth;
    J.rows = f.rows;
    J.cols = f.cols;
    J.fx = f.K[0];
    J.fy = f.K[4];
    J.cx = f.K[6];
    J.cy = f.K[7];
    const float *T = poses16 + 16 * (size_t)i;
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r) J.R[c * 3 + r] = T[c * 4 + r];
    for (int r = 0; r < 3; ++r) J.t[r] = T[12 + r];
    J.seed = scene_seeds[i];
    J.stream = streams[i];
    synth::make_scene(J.seed, J.fx, J.fy, J.cols, J.rows, one);
    J.prim_base = (int)prims.size();
    J.nprims = (int)one.size();
    if (J.nprims != synth::kPrims) return set_err(NICP_ERR_INVALID, "scene with %d primitives", J.nprims);
    prims.insert(prims.end(), one.begin(), one.end());
    J.tile_base = (int)blk.size();
    const int tiles = (J.rows * J.cols + kRenderTile - 1) / kRenderTile;
    blk.insert(blk.end(), tiles, i);
    // rendered depth is in millimetres
    if (f.depth_scale != 0.001f) {
      f.depth_scale = 0.001f;
      HIPCHK(hipMemcpyAsync(&b->d_frames[frames[i]], &f, sizeof(FrameDev), hipMemcpyHostToDevice, s));
    }
  }
  Carver cv;
  const size_t o_jobs = cv.take<RenderJob>(jobs.size());
  const size_t o_prims = cv.take<synth::Prim>(prims.size());
  const size_t o_blk = cv.take<int>(blk.size());
  CHK(b->synth_buf.reserve(cv.off));
  char *base = b->synth_buf.as<char>();
  HIPCHK(hipMemcpyAsync(base + o_jobs, jobs.data(), jobs.size() * sizeof(RenderJob), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(base + o_prims, prims.data(), prims.size() * sizeof(synth::Prim), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(base + o_blk, blk.data(), blk.size() * sizeof(int), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_synth_render, dim3((unsigned)blk.size()), dim3(256), 0, s, (const RenderJob *)(base + o_jobs),
                     (const int *)(base + o_blk), (const synth::Prim *)(base + o_prims), (double)noise_coeff,
                     (double)hole_fraction);
  HIPCHK(hipGetLastError());
  // the host tables above are read by the copies before this call returns
  HIPCHK(hipStreamSynchronize(s));
  return NICP_OK;
}
