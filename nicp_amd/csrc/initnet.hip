// Learned initial guess for the registration loop (no counterpart in the
// reference, whose Aligner takes the initial guess from the caller,
// nicp/nicp/aligner.h:102-106): a small CNN maps a pair of 60x80 depth images
// (the reference and current frames reduced by DepthImage_scale,
// nicp/nicp/imageutils.cpp:6-37) to the 6-vector (t, q_xyz) of the relative
// pose in the parametrisation of v2t/t2v (nicp/nicp/bm_se3.h:10-51).
//
//   conv 3x3 s2 p1:  2 -> 32   60x80 -> 30x40   ReLU
//   conv 3x3 s2 p1: 32 -> 32   30x40 -> 15x20   ReLU
//   conv 3x3 s2 p1: 32 -> 64   15x20 ->  8x10   ReLU
//   conv 3x3 s2 p1: 64 -> 64    8x10 ->  4x5    ReLU
//   fc 1280 -> 256 ReLU, fc 256 -> 6
//
// Every layer is one GEMM on the matrix cores: bf16 operands, fp32 accumulation
// (gfx950's v_mfma_f32_16x16x32_bf16), with the im2col gather of the input and the bias
// (an extra weight row multiplied by a constant 1) folded into the operand loads.
// Backward: weight gradients are GEMMs over the batch (split over the reduction,
// fp32 atomics), input gradients are GEMMs scattered back by col2im (fp32
// atomics). Activations are NHWC fp32; master weights are fp32 and rounded to
// bf16 when staged into LDS for the matrix cores. Adam updates the master weights.
//
// Parameter layout: layer by layer, a row-major (K + 1) x N matrix, where the
// reduction index K of a convolution is (ky, kx, ci) with ci fastest and the last
// row is the bias.
//
// This file is part of the nicp_capi.hip translation unit (it uses its error
// handling, context and batch objects).

namespace initnet {

constexpr int kLayers = 6;
constexpr int kFcSplits = 16;  // most reduction chunks of a forward GEMM (fwd_splits)
constexpr int kProbeSub = 8;  // pixel stride of the start probe's linearisation
constexpr int kInRows = 60, kInCols = 80, kInCh = 2;

struct LayerShape {
  int ih, iw, ic, oh, ow, oc, conv, relu;
};
__host__ __device__ constexpr LayerShape layer_shape(int l) {
  return l == 0   ? LayerShape{60, 80, 2, 30, 40, 32, 1, 1}
         : l == 1 ? LayerShape{30, 40, 32, 15, 20, 32, 1, 1}
         : l == 2 ? LayerShape{15, 20, 32, 8, 10, 64, 1, 1}
         : l == 3 ? LayerShape{8, 10, 64, 4, 5, 64, 1, 1}
         : l == 4 ? LayerShape{1, 1, 1280, 1, 1, 256, 0, 1}
                  : LayerShape{1, 1, 256, 1, 1, 6, 0, 0};
}
// reduction length of a layer without the bias row
__host__ __device__ constexpr int layer_k(int l) {
  return layer_shape(l).conv ? 9 * layer_shape(l).ic : layer_shape(l).ic;
}
__host__ __device__ constexpr int layer_params(int l) { return (layer_k(l) + 1) * layer_shape(l).oc; }

// fp32 -> bf16, round to nearest even (inf and nan keep their class)
__device__ __forceinline__ unsigned short to_bf16(float f) {
  unsigned int u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

typedef __bf16 __attribute__((ext_vector_type(8))) bf16x8;
typedef float __attribute__((ext_vector_type(4))) f32x4;

// ---------------------------------------------------------------------------------
// operand gathers
// ---------------------------------------------------------------------------------
// A of the forward GEMM: row m = (b, oy, ox), column k = (ky, kx, ci), or the bias column k == K
// inv = 1 (the first layer in the inverse-depth input mode) reads 1 / x, and 0 for x == 0
struct ConvIn {
  const float *x;
  int ih, iw, ic, oh, ow, K;
  int conv, inv;
  __device__ __forceinline__ float feature(float v) const { return inv ? (v != 0.0f ? 1.0f / v : 0.0f) : v; }
  __device__ __forceinline__ float operator()(int m, int k) const {
    if (k == K) return 1.0f;
    if (!conv) return feature(x[(size_t)m * ic + k]);
    const int ox = m % ow, t = m / ow, oy = t % oh, b = t / oh;
    const int ci = k % ic, kk = k / ic, kx = kk % 3, ky = kk / 3;
    const int iy = 2 * oy + ky - 1, ix = 2 * ox + kx - 1;
    if (iy < 0 || iy >= ih || ix < 0 || ix >= iw) return 0.0f;
    return feature(x[(((size_t)b * ih + iy) * iw + ix) * ic + ci]);
  }
};
// B of the forward GEMM: weights (K + 1) x N
struct Weights {
  const float *w;
  int N;
  __device__ __forceinline__ float operator()(int k, int n) const { return w[(size_t)k * N + n]; }
};
// A'(k, m) = A(m, k): the weight gradient is A^T dY
struct ConvInT {
  ConvIn a;
  __device__ __forceinline__ float operator()(int k, int m) const { return a(m, k); }
};
// a row-major M x N matrix, read as either operand
struct Rows {
  const float *y;
  int N;
  __device__ __forceinline__ float operator()(int m, int n) const { return y[(size_t)m * N + n]; }
};
// B'(n, k) = W(k, n): the input gradient dcol = dY W^T (the bias row is not part of it)
struct WeightsT {
  const float *w;
  int N;
  __device__ __forceinline__ float operator()(int n, int k) const { return w[(size_t)k * N + n]; }
};

// ---------------------------------------------------------------------------------
// result stores
// ---------------------------------------------------------------------------------
struct StoreAct {  // forward: y = relu?(v)
  float *y;
  int N, relu;
  __device__ __forceinline__ void operator()(int m, int n, float v) const {
    y[(size_t)m * N + n] = relu ? fmaxf(v, 0.0f) : v;
  }
};
struct StorePartial {  // one chunk of a split reduction: p[chunk][m][n] = v (summed by k_sum_chunks)
  float *p;
  int M, N;
  __device__ __forceinline__ void operator()(int m, int n, float v) const {
    p[((size_t)blockIdx.z * M + m) * N + n] = v;
  }
};
struct StoreAtomic {  // weight gradient, split over the reduction: dW += v
  float *g;
  int N;
  __device__ __forceinline__ void operator()(int m, int n, float v) const { atomicAdd(g + (size_t)m * N + n, v); }
};
struct StoreCol2Im {  // input gradient: scatter dcol(m, k) into dX (col2im)
  float *dx;
  int ih, iw, ic, oh, ow, conv;
  __device__ __forceinline__ void operator()(int m, int k, float v) const {
    if (!conv) {
      dx[(size_t)m * ic + k] = v;  // one writer per element
      return;
    }
    const int ox = m % ow, t = m / ow, oy = t % oh, b = t / oh;
    const int ci = k % ic, kk = k / ic, kx = kk % 3, ky = kk / 3;
    const int iy = 2 * oy + ky - 1, ix = 2 * ox + kx - 1;
    if (iy < 0 || iy >= ih || ix < 0 || ix >= iw) return;
    atomicAdd(dx + (((size_t)b * ih + iy) * iw + ix) * ic + ci, v);
  }
};

// ---------------------------------------------------------------------------------
// C[M][N] = sum_k A(m, k) B(k, n) over k in this block's chunk. A 64 x 64 output tile
// per block of 4 wavefronts, each wavefront a 32 x 32 quadrant of 2 x 2 tiles of gfx950's
// v_mfma_f32_16x16x32_bf16; the reduction advances 32 at a time through LDS in bf16.
// blockIdx.z selects the chunk [z * kchunk, (z + 1) * kchunk) of the reduction.
// ---------------------------------------------------------------------------------
constexpr int kGemmK = 32;      // reduction depth of one MFMA step
constexpr int kGemmLds = 40;    // LDS row pitch in bf16: 80 bytes, so a lane's 8 values are one 16-byte read
template <class LA, class LB, class ST>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, int kchunk, LA la, LB lb, ST st) {
  __shared__ __attribute__((aligned(16))) unsigned short sA[64][kGemmLds];  // [row][k]
  __shared__ __attribute__((aligned(16))) unsigned short sB[64][kGemmLds];  // [col][k]
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;  // output tiles: rows on x (unbounded), columns on y
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int kbeg = blockIdx.z * kchunk, kend = min(K, kbeg + kchunk);
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  // staging: thread t fills row/col (t >> 2) of both tiles at k offsets (t & 3) * 8 .. + 7
  const int sr = tid >> 2, sk = (tid & 3) * 8;
  // operands: lane holds A[lane % 16][8 (lane / 16) + j] and B[8 (lane / 16) + j][lane % 16],
  // j = 0..7. The matrix core pairs A's and B's values of one lane slot and register element,
  // so this assignment of the 32 reduction indices gives the full dot product.
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  for (int k0 = kbeg; k0 < kend; k0 += kGemmK) {
    unsigned short va[8], vb[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = k0 + sk + j, m = m0 + sr, n = n0 + sr;
      const bool kin = k < kend;
      va[j] = to_bf16(kin && m < M ? la(m, k) : 0.0f);
      vb[j] = to_bf16(kin && n < N ? lb(k, n) : 0.0f);
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      sA[sr][sk + j] = va[j];
      sB[sr][sk + j] = vb[j];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      bf16x8 a;
      const unsigned short *ra = &sA[wm + i * 16 + fr][fk];
#pragma unroll
      for (int j = 0; j < 8; ++j) a[j] = __builtin_bit_cast(__bf16, ra[j]);
#pragma unroll
      for (int jj = 0; jj < 2; ++jj) {
        bf16x8 b;
        const unsigned short *rb = &sB[wn + jj * 16 + fr][fk];
#pragma unroll
        for (int j = 0; j < 8; ++j) b[j] = __builtin_bit_cast(__bf16, rb[j]);
        acc[i][jj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i][jj], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // result layout: lane holds C[4 (lane / 16) + r][lane % 16], r = 0..3
  const int dc = lane & 15, dr = (lane >> 4) * 4;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + i * 16 + dr + r, n = n0 + wn + jj * 16 + dc;
        if (m < M && n < N) st(m, n, acc[i][jj][r]);
      }
}

// the reduction chunk of a GEMM split `splits` ways (a multiple of kGemmK) and the chunk count
inline int gemm_kchunk(int K, int splits) { return ((K + splits - 1) / splits + kGemmK - 1) / kGemmK * kGemmK; }
inline int gemm_chunks(int K, int splits) { return (K + gemm_kchunk(K, splits) - 1) / gemm_kchunk(K, splits); }

template <class LA, class LB, class ST>
hipError_t gemm(hipStream_t s, int M, int N, int K, int splits, LA la, LB lb, ST st) {
  const int kchunk = gemm_kchunk(K, splits);
  const int z = gemm_chunks(K, splits);
  dim3 grid((M + 63) / 64, (N + 63) / 64, z);
  hipLaunchKernelGGL((k_gemm<LA, LB, ST>), grid, dim3(256), 0, s, M, N, K, kchunk, la, lb, st);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// elementwise kernels
// ---------------------------------------------------------------------------------
// y[m][n] = relu?(sum over chunks c = 0, 1, .. of p[c][m][n]), the chunks added in order
__global__ void k_sum_chunks(const float *__restrict__ p, int chunks, int M, int N, float *__restrict__ y, int relu) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x, mn = (size_t)M * N;
  if (i >= mn) return;
  float v = p[i];
  for (int c = 1; c < chunks; ++c) v += p[c * mn + i];
  y[i] = relu ? fmaxf(v, 0.0f) : v;
}

__global__ void k_fill(float *p, size_t n, float v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
// gradient through a ReLU: g = 0 where the output was not positive
__global__ void k_relu_mask(float *g, const float *y, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(y[i] > 0.0f)) g[i] = 0.0f;
}
// weighted squared error: loss = 1/B sum_b sum_j w_j (y - t)^2, dY = 2 w_j (y - t) / B
__global__ void k_mse(const float *y, const float *t, const float *w, int B, float *dy, float *loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < B * 6) {
    const float d = y[i] - t[i];
    const float wi = w[i % 6];
    dy[i] = 2.0f * wi * d / (float)B;
    l = wi * d * d / (float)B;
  }
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
}
__global__ void k_adam(float *w, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// minibatch assembly from a resident set of network inputs: dst[b] = sym_b(src[rows[b]]),
// where the symmetry flags of sample b (flags[b]) are bit 0: swap the two frames, bit 1:
// mirror the columns, bit 2: mirror the rows
__global__ void k_gather_inputs(const float *__restrict__ src, const int *__restrict__ rows,
                                const int *__restrict__ flags, int B, float *__restrict__ dst) {
  constexpr int W = kInRows * kInCols * kInCh;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * W) return;
  const int b = i / W, j = i - b * W;
  const int ch = j % kInCh, cell = j / kInCh, x = cell % kInCols, y = cell / kInCols;
  const int f = flags[b];
  const int sch = (f & 1) ? 1 - ch : ch, sx = (f & 2) ? kInCols - 1 - x : x, sy = (f & 4) ? kInRows - 1 - y : y;
  dst[i] = src[(size_t)rows[b] * W + (sy * kInCols + sx) * kInCh + sch];
}

// Network input of pair p: channel 0 = reference frame, channel 1 = current frame, each
// reduced by DepthImage_scale with step = cols / 80 (the block mean of the non-zero
// depths, 0 when the block's variance exceeds max_cov). One thread per output cell.
struct InputJob {
  const uint16_t *depth;
  float scale;
  int rows, cols, step;
};
__global__ __launch_bounds__(256) void k_net_input(const InputJob *__restrict__ jobs, int num_pairs, float max_cov,
                                                   float *__restrict__ out) {
  const int o = blockIdx.x * 256 + threadIdx.x;
  const int cells = kInRows * kInCols * kInCh;
  if (o >= num_pairs * cells) return;
  const int p = o / cells, rem = o - p * cells;
  const int ch = rem % kInCh, cell = rem / kInCh, x = cell % kInCols, y = cell / kInCols;
  const InputJob J = jobs[2 * p + ch];
  float acc = 0.0f, acc2 = 0.0f;
  int np = 0;
  const int sr = y * J.step, sc = x * J.step;
  for (int i = 0; i < J.step; ++i)
    for (int j = 0; j < J.step; ++j) {
      const float f = ::depth_of(J.depth[(size_t)(sr + i) * J.cols + sc + j], J.scale);
      acc += f;
      acc2 += f * f;
      np += f > 0.0f;
    }
  float v = 0.0f;
  if (np) {
    const float mu = acc / (float)np;
    const float sigma = acc2 / (float)np - mu * mu;
    if (!(sigma > max_cov)) v = mu;
  }
  out[o] = v;
}

}  // namespace initnet

// ---------------------------------------------------------------------------------
// network object
// ---------------------------------------------------------------------------------
struct nicp_initnet {
  nicp_ctx *ctx = nullptr;
  int cap = 0;  // batch capacity of the activation buffers
  float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  size_t nparams = 0;
  size_t poff[initnet::kLayers + 1] = {0};
  float *act[initnet::kLayers + 1] = {nullptr};   // act[0] = input, act[l + 1] = output of layer l
  float *dact[initnet::kLayers + 1] = {nullptr};  // gradients of the loss w.r.t. act[]
  float *target = nullptr, *lossbuf = nullptr, *wloss = nullptr;
  void *jobs = nullptr;
  int jobs_cap = 0;
  int *rows_dev = nullptr;  // minibatch row indices and symmetry flags
  int rows_cap = 0;
  int input_mode = 0;       // 0: depth (m) into the first layer, 1: inverse depth
  int adam_t = 0;
  float *chunks = nullptr;  // partial sums of the fully connected layers' split reductions
};

namespace {

using namespace initnet;

size_t act_count(int l, int B) {
  if (l == 0) return (size_t)B * kInRows * kInCols * kInCh;
  const LayerShape s = layer_shape(l - 1);
  return (size_t)B * s.oh * s.ow * s.oc;
}

// reduction chunks of a layer's forward GEMM: a layer with fewer than 256 output tiles of 64 x 64
// (the fully connected ones and the last convolutions) splits its reduction over more
// workgroups, at most kFcSplits ways; the chunks are then added in order
int fwd_splits(int M, int N) {
  const int tiles = ((M + 63) / 64) * ((N + 63) / 64);
  return tiles >= 256 ? 1 : std::min(kFcSplits, 256 / tiles);
}

int net_reserve(nicp_initnet *n, int B) {
  if (B <= n->cap) return NICP_OK;
  (void)hipStreamSynchronize(n->ctx->stream);
  for (int l = 0; l <= kLayers; ++l) {
    if (n->act[l]) (void)hipFree(n->act[l]);
    if (n->dact[l]) (void)hipFree(n->dact[l]);
    n->act[l] = n->dact[l] = nullptr;
  }
  if (n->target) (void)hipFree(n->target);
  n->target = nullptr;
  n->cap = 0;
  for (int l = 0; l <= kLayers; ++l) {
    HIPCHK(hipMalloc(&n->act[l], act_count(l, B) * sizeof(float)));
    HIPCHK(hipMalloc(&n->dact[l], act_count(l, B) * sizeof(float)));
  }
  HIPCHK(hipMalloc(&n->target, (size_t)B * 6 * sizeof(float)));
  if (n->chunks) (void)hipFree(n->chunks);
  n->chunks = nullptr;
  // the split reductions' chunks for every batch size up to B: a later call with fewer rows
  // (which reuses these buffers) splits its reductions further, and can need more room than B
  size_t most = 1;
  for (int b = 1; b <= B; ++b)
    for (int l = 0; l < kLayers; ++l) {
      const LayerShape sh = layer_shape(l);
      const int M = b * sh.oh * sh.ow, sp = fwd_splits(M, sh.oc);
      if (sp > 1) most = std::max(most, (size_t)gemm_chunks(layer_k(l) + 1, sp) * M * sh.oc);
    }
  HIPCHK(hipMalloc(&n->chunks, most * sizeof(float)));
  n->cap = B;
  return NICP_OK;
}

int net_fill(hipStream_t s, float *p, size_t cnt) {
  if (!cnt) return NICP_OK;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, p, cnt, 0.0f);
  HIPCHK(hipGetLastError());
  return NICP_OK;
}

int net_forward(nicp_initnet *n, int B) {
  hipStream_t s = n->ctx->stream;
  for (int l = 0; l < kLayers; ++l) {
    const LayerShape sh = layer_shape(l);
    const int M = B * sh.oh * sh.ow, K = layer_k(l);
    ConvIn a{n->act[l], sh.ih, sh.iw, sh.ic, sh.oh, sh.ow, K, sh.conv, l == 0 ? n->input_mode : 0};
    const int sp = fwd_splits(M, sh.oc);
    if (sp == 1) {
      HIPCHK(gemm(s, M, sh.oc, K + 1, 1, a, Weights{n->params + n->poff[l], sh.oc},
                  StoreAct{n->act[l + 1], sh.oc, sh.relu}));
    } else {
      HIPCHK(gemm(s, M, sh.oc, K + 1, sp, a, Weights{n->params + n->poff[l], sh.oc}, StorePartial{n->chunks, M, sh.oc}));
      const size_t mn = (size_t)M * sh.oc;
      hipLaunchKernelGGL(k_sum_chunks, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, s, (const float *)n->chunks,
                         gemm_chunks(K + 1, sp), M, sh.oc, n->act[l + 1], sh.relu);
      HIPCHK(hipGetLastError());
    }
  }
  return NICP_OK;
}

// gradient of the loss in n->grads; n->dact[kLayers] must hold dLoss/dOutput
int net_backward(nicp_initnet *n, int B) {
  hipStream_t s = n->ctx->stream;
  CHK(net_fill(s, n->grads, n->nparams));
  for (int l = kLayers - 1; l >= 0; --l) {
    const LayerShape sh = layer_shape(l);
    const int M = B * sh.oh * sh.ow, K = layer_k(l);
    if (sh.relu) {
      const size_t cnt = act_count(l + 1, B);
      hipLaunchKernelGGL(k_relu_mask, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, n->dact[l + 1],
                         n->act[l + 1], cnt);
      HIPCHK(hipGetLastError());
    }
    // dW[k][n] = sum_m A(m, k) dY(m, n), the sum over m split in chunks of >= 256
    ConvIn a{n->act[l], sh.ih, sh.iw, sh.ic, sh.oh, sh.ow, K, sh.conv, l == 0 ? n->input_mode : 0};
    const int splits = std::max(1, std::min(128, M / 256));
    HIPCHK(gemm(s, K + 1, sh.oc, M, splits, ConvInT{a}, Rows{n->dact[l + 1], sh.oc},
                StoreAtomic{n->grads + n->poff[l], sh.oc}));
    if (l > 0) {
      // dX = col2im(dY W^T)
      CHK(net_fill(s, n->dact[l], act_count(l, B)));
      HIPCHK(gemm(s, M, K, sh.oc, 1, Rows{n->dact[l + 1], sh.oc}, WeightsT{n->params + n->poff[l], sh.oc},
                  StoreCol2Im{n->dact[l], sh.ih, sh.iw, sh.ic, sh.oh, sh.ow, sh.conv}));
    }
  }
  return NICP_OK;
}

int check_net(const nicp_initnet *n) {
  if (!n) return set_err(NICP_ERR_INVALID, "network is null");
  return NICP_OK;
}

}  // namespace

extern "C" int nicp_initnet_num_params(void) {
  int n = 0;
  for (int l = 0; l < kLayers; ++l) n += layer_params(l);
  return n;
}

extern "C" int nicp_initnet_destroy(nicp_initnet *n);

extern "C" int nicp_initnet_create(nicp_ctx *ctx, uint64_t seed, nicp_initnet **out) {
  if (!ctx || !out) return set_err(NICP_ERR_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(ctx->device));
  nicp_initnet *n = new (std::nothrow) nicp_initnet();
  if (!n) return set_err(NICP_ERR_OOM, "out of host memory");
  n->ctx = ctx;
  size_t off = 0;
  for (int l = 0; l < kLayers; ++l) {
    n->poff[l] = off;
    off += layer_params(l);
  }
  n->poff[kLayers] = off;
  n->nparams = off;
  hipError_t e = hipMalloc(&n->params, off * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&n->grads, off * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&n->adam_m, off * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&n->adam_v, off * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&n->lossbuf, 4 * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&n->wloss, 6 * sizeof(float));
  if (e != hipSuccess) {
    nicp_initnet_destroy(n);
    return set_err(NICP_ERR_OOM, "network allocation: %s", hipGetErrorString(e));
  }
  // He-normal weights from a seeded xorshift (the output layer scaled by 0.1), zero biases
  std::vector<float> w(off, 0.0f);
  uint64_t state = seed * 0x9e3779b97f4a7c15ull + 0x2545F4914F6CDD1Dull;
  auto uni = [&]() {
    state ^= state >> 12;
    state ^= state << 25;
    state ^= state >> 27;
    return (double)((state * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int l = 0; l < kLayers; ++l) {
    const int K = layer_k(l), N = layer_shape(l).oc;
    const double sd = std::sqrt(2.0 / K) * (l == kLayers - 1 ? 0.1 : 1.0);
    for (int k = 0; k < K; ++k)
      for (int c = 0; c < N; ++c) {
        const double u1 = std::max(uni(), 1e-300), u2 = uni();
        w[n->poff[l] + (size_t)k * N + c] =
            (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
      }
  }
  const float lw[6] = {1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f};
  hipStream_t s = ctx->stream;
  e = hipMemcpyAsync(n->params, w.data(), off * sizeof(float), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(n->wloss, lw, sizeof(lw), hipMemcpyHostToDevice, s);
  if (e == hipSuccess && (net_fill(s, n->adam_m, off) != NICP_OK || net_fill(s, n->adam_v, off) != NICP_OK))
    e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    nicp_initnet_destroy(n);
    return set_err(NICP_ERR_HIP, "network initialisation: %s", hipGetErrorString(e));
  }
  *out = n;
  return NICP_OK;
}

extern "C" int nicp_initnet_destroy(nicp_initnet *n) {
  if (!n) return NICP_OK;
  (void)hipSetDevice(n->ctx->device);
  (void)hipStreamSynchronize(n->ctx->stream);
  for (int l = 0; l <= kLayers; ++l) {
    if (n->act[l]) (void)hipFree(n->act[l]);
    if (n->dact[l]) (void)hipFree(n->dact[l]);
  }
  float *bufs[] = {n->params, n->grads, n->adam_m, n->adam_v, n->target, n->lossbuf, n->wloss};
  for (float *b : bufs)
    if (b) (void)hipFree(b);
  if (n->jobs) (void)hipFree(n->jobs);
  if (n->rows_dev) (void)hipFree(n->rows_dev);
  if (n->chunks) (void)hipFree(n->chunks);
  delete n;
  return NICP_OK;
}

extern "C" int nicp_initnet_get_params(nicp_initnet *n, float *out) {
  CHK(check_net(n));
  if (!out) return set_err(NICP_ERR_INVALID, "output is null");
  HIPCHK(hipSetDevice(n->ctx->device));
  HIPCHK(hipMemcpyAsync(out, n->params, n->nparams * sizeof(float), hipMemcpyDeviceToHost, n->ctx->stream));
  HIPCHK(hipStreamSynchronize(n->ctx->stream));
  return NICP_OK;
}

extern "C" int nicp_initnet_set_params(nicp_initnet *n, const float *in) {
  CHK(check_net(n));
  if (!in) return set_err(NICP_ERR_INVALID, "input is null");
  HIPCHK(hipSetDevice(n->ctx->device));
  HIPCHK(hipMemcpyAsync(n->params, in, n->nparams * sizeof(float), hipMemcpyHostToDevice, n->ctx->stream));
  n->adam_t = 0;
  CHK(net_fill(n->ctx->stream, n->adam_m, n->nparams));
  CHK(net_fill(n->ctx->stream, n->adam_v, n->nparams));
  HIPCHK(hipStreamSynchronize(n->ctx->stream));
  return NICP_OK;
}

extern "C" int nicp_initnet_set_loss_weights(nicp_initnet *n, const float w6[6]) {
  CHK(check_net(n));
  if (!w6) return set_err(NICP_ERR_INVALID, "weights are null");
  HIPCHK(hipSetDevice(n->ctx->device));
  HIPCHK(hipMemcpyAsync(n->wloss, w6, 6 * sizeof(float), hipMemcpyHostToDevice, n->ctx->stream));
  HIPCHK(hipStreamSynchronize(n->ctx->stream));
  return NICP_OK;
}

extern "C" int nicp_initnet_forward(nicp_initnet *n, const float *input, int batch, int on_device, float *out6) {
  CHK(check_net(n));
  if (!input || batch <= 0) return set_err(NICP_ERR_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(n->ctx->device));
  CHK(net_reserve(n, batch));
  hipStream_t s = n->ctx->stream;
  HIPCHK(hipMemcpyAsync(n->act[0], input, act_count(0, batch) * sizeof(float),
                        on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
  CHK(net_forward(n, batch));
  if (out6)
    HIPCHK(hipMemcpyAsync(out6, n->act[kLayers], (size_t)batch * 6 * sizeof(float),
                          on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return NICP_OK;
}

namespace {
// forward + loss + backward (+ Adam when lr > 0) on act[0] / target already in place
int net_train(nicp_initnet *n, int batch, float lr, float *loss_out) {
  hipStream_t s = n->ctx->stream;
  CHK(net_forward(n, batch));
  CHK(net_fill(s, n->lossbuf, 4));
  hipLaunchKernelGGL(k_mse, dim3((batch * 6 + 255) / 256), dim3(256), 0, s, n->act[kLayers], n->target, n->wloss,
                     batch, n->dact[kLayers], n->lossbuf);
  HIPCHK(hipGetLastError());
  CHK(net_backward(n, batch));
  if (lr > 0.0f) {
    ++n->adam_t;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)n->adam_t), c2 = 1.0f - std::pow(b2, (float)n->adam_t);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n->nparams + 255) / 256)), dim3(256), 0, s, n->params, n->grads,
                       n->adam_m, n->adam_v, n->nparams, lr, b1, b2, 1e-8f, c1, c2);
    HIPCHK(hipGetLastError());
  }
  float loss = 0.0f;
  HIPCHK(hipMemcpyAsync(&loss, n->lossbuf, sizeof(float), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (loss_out) *loss_out = loss;
  return NICP_OK;
}
}  // namespace

extern "C" int nicp_initnet_train_step(nicp_initnet *n, const float *input, const float *target6, int batch,
                                       int on_device, float lr, float *loss_out) {
  CHK(check_net(n));
  if (!input || !target6 || batch <= 0) return set_err(NICP_ERR_INVALID, "bad arguments");
  HIPCHK(hipSetDevice(n->ctx->device));
  CHK(net_reserve(n, batch));
  hipStream_t s = n->ctx->stream;
  const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  HIPCHK(hipMemcpyAsync(n->act[0], input, act_count(0, batch) * sizeof(float), kind, s));
  HIPCHK(hipMemcpyAsync(n->target, target6, (size_t)batch * 6 * sizeof(float), kind, s));
  return net_train(n, batch, lr, loss_out);
}

extern "C" int nicp_initnet_train_step_rows(nicp_initnet *n, const float *inputs, int num_rows, const int *rows,
                                            const int *flags, const float *targets6, int batch, float lr,
                                            float *loss_out) {
  CHK(check_net(n));
  if (!inputs || !targets6 || !rows || batch <= 0 || num_rows <= 0) return set_err(NICP_ERR_INVALID, "bad arguments");
  for (int i = 0; i < batch; ++i) {
    if (rows[i] < 0 || rows[i] >= num_rows) return set_err(NICP_ERR_INVALID, "row %d out of range", rows[i]);
    if (flags && (flags[i] < 0 || flags[i] > 7)) return set_err(NICP_ERR_INVALID, "bad symmetry flags %d", flags[i]);
  }
  HIPCHK(hipSetDevice(n->ctx->device));
  CHK(net_reserve(n, batch));
  hipStream_t s = n->ctx->stream;
  if (2 * batch > n->rows_cap) {
    HIPCHK(hipStreamSynchronize(s));
    if (n->rows_dev) (void)hipFree(n->rows_dev);
    n->rows_dev = nullptr;
    n->rows_cap = 0;
    HIPCHK(hipMalloc(&n->rows_dev, (size_t)2 * batch * sizeof(int)));
    n->rows_cap = 2 * batch;
  }
  std::vector<int> rf(2 * (size_t)batch, 0);
  memcpy(rf.data(), rows, (size_t)batch * sizeof(int));
  if (flags) memcpy(rf.data() + batch, flags, (size_t)batch * sizeof(int));
  HIPCHK(hipMemcpyAsync(n->rows_dev, rf.data(), rf.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(n->target, targets6, (size_t)batch * 6 * sizeof(float), hipMemcpyHostToDevice, s));
  const size_t cnt = act_count(0, batch);
  hipLaunchKernelGGL(k_gather_inputs, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, inputs, n->rows_dev,
                     n->rows_dev + batch, batch, n->act[0]);
  HIPCHK(hipGetLastError());
  // net_train ends with a stream synchronisation, after which the host tables may go
  return net_train(n, batch, lr, loss_out);
}

extern "C" int nicp_initnet_set_input_mode(nicp_initnet *n, int mode) {
  CHK(check_net(n));
  if (mode < 0 || mode > 1) return set_err(NICP_ERR_INVALID, "input mode must be 0 or 1");
  n->input_mode = mode;
  return NICP_OK;
}

extern "C" int nicp_initnet_read(nicp_initnet *n, int what, int index, float *out, size_t count) {
  CHK(check_net(n));
  if (!out || what < 0 || what > 2 || (what < 2 && (index < 0 || index > kLayers)))
    return set_err(NICP_ERR_INVALID, "bad arguments");
  const float *src = what == 2 ? n->grads : what == 0 ? n->act[index] : n->dact[index];
  const size_t avail = what == 2 ? n->nparams : (src ? act_count(index, n->cap) : 0);
  if (count > avail) return set_err(NICP_ERR_INVALID, "read of %zu floats exceeds the %zu held", count, avail);
  HIPCHK(hipSetDevice(n->ctx->device));
  HIPCHK(hipMemcpyAsync(out, src, count * sizeof(float), hipMemcpyDeviceToHost, n->ctx->stream));
  HIPCHK(hipStreamSynchronize(n->ctx->stream));
  return NICP_OK;
}

// ---------------------------------------------------------------------------------
// batch integration
// ---------------------------------------------------------------------------------
namespace {

int check_net_frames(const nicp_batch *b, int num_pairs, const int *ref, const int *cur) {
  for (int i = 0; i < num_pairs; ++i) {
    const int f2[2] = {ref[i], cur[i]};
    for (int f : f2) {
      if (f < 0 || f >= b->num_frames) return set_err(NICP_ERR_INVALID, "pair %d: frame out of range", i);
      const FrameDev &fd = b->frames[f];
      if (fd.cols % kInCols != 0 || fd.rows * kInCols != fd.cols * kInRows)
        return set_err(NICP_ERR_INVALID, "frame %d: %dx%d is not a multiple of 80x60", f, fd.cols, fd.rows);
    }
  }
  return NICP_OK;
}

// the decision rule of nicp_batch_align_learned (chi2 and inlier count of one linearisation
// at each start)
NICP_HD bool fallback_wins(double chi2_net, double inl_net, double chi2_fb, double inl_fb) {
  if (!(inl_net > 0.0)) return inl_fb > 0.0;
  if (!(inl_fb >= 0.5 * inl_net)) return false;
  return chi2_fb / inl_fb < chi2_net / inl_net;
}

// The network input's job table, uploaded ahead of the kernels that read it: from the batch's
// pinned staging when given (asynchronous), else from a pageable vector, waited for.
int upload_net_jobs(nicp_initnet *n, nicp_batch *b, int num_pairs, const int *ref, const int *cur,
                    PinnedStage *stage = nullptr) {
  hipStream_t s = b->ctx->stream;
  if (2 * num_pairs > n->jobs_cap) {
    HIPCHK(hipStreamSynchronize(s));
    if (n->jobs) (void)hipFree(n->jobs);
    n->jobs = nullptr;
    n->jobs_cap = 0;
    HIPCHK(hipMalloc(&n->jobs, (size_t)2 * num_pairs * sizeof(InputJob)));
    n->jobs_cap = 2 * num_pairs;
  }
  std::vector<InputJob> jobs(2 * (size_t)num_pairs);
  for (int i = 0; i < num_pairs; ++i)
    for (int ch = 0; ch < 2; ++ch) {
      const FrameDev &fd = b->frames[ch ? cur[i] : ref[i]];
      jobs[2 * i + ch] = InputJob{fd.depth, fd.depth_scale, fd.rows, fd.cols, fd.cols / kInCols};
    }
  const void *src = stage ? stage->put(jobs.data(), jobs.size() * sizeof(InputJob)) : jobs.data();
  HIPCHK(hipMemcpyAsync(n->jobs, src, jobs.size() * sizeof(InputJob), hipMemcpyHostToDevice, s));
  // a pageable vector must outlive the copy
  if (!stage) HIPCHK(hipStreamSynchronize(s));
  return NICP_OK;
}

int enqueue_net_input(nicp_initnet *n, nicp_batch *b, int num_pairs, float max_cov, float *dst) {
  const size_t cnt = (size_t)num_pairs * kInRows * kInCols * kInCh;
  hipLaunchKernelGGL(k_net_input, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, b->ctx->stream,
                     (const InputJob *)n->jobs, num_pairs, max_cov, dst);
  HIPCHK(hipGetLastError());
  return NICP_OK;
}

int launch_net_input(nicp_initnet *n, nicp_batch *b, int num_pairs, const int *ref, const int *cur, float max_cov,
                     float *dst) {
  CHK(upload_net_jobs(n, b, num_pairs, ref, cur));
  CHK(enqueue_net_input(n, b, num_pairs, max_cov, dst));
  // the device table is rewritten by the next call
  HIPCHK(hipStreamSynchronize(b->ctx->stream));
  return NICP_OK;
}

// A pair's loop state at start T (AlignerProjective::align, alignerprojective.cpp:44-51): T, its
// inverse and the reference projector K (T * reference offset)^-1, with the offset already in
// the state; the same arithmetic as the host's initial state.
__device__ void set_start_state(const PairDev &pd, const float *Tin) {
  float T[16], Ti[16], RO[16], TR[16], KRt[16];
  for (int k = 0; k < 16; ++k) T[k] = Tin[k];
  set_row3(T);
  iso_inverse(T, Ti);
  for (int k = 0; k < 16; ++k) RO[k] = pd.state[kSlotRefOffset + k];
  iso_mul(T, RO, TR);
  projector_state(pd.model, pd.K, TR, KRt);
  for (int k = 0; k < 16; ++k) {
    pd.state[kSlotT + k] = T[k];
    pd.state[kSlotTinv + k] = Ti[k];
    pd.state[kSlotKRt + k] = KRt[k];
  }
}

// The two candidate starts of each pair: pair i of the probe starts at the network's prediction
// v2t(x) (non-finite entries set to 0, the quaternion's vector part scaled into the ball of
// radius sqrt(0.99)), pair n + i at the fallback guess its state already holds; starts16 keeps
// both (2n x 16) for the choice, which runs after the probe has moved the probe's state.
__global__ __launch_bounds__(64) void k_probe_starts(const float *__restrict__ pred, int n, const PairDev *__restrict__ Pp,
                                                     float *__restrict__ starts16) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  float x[6];
  for (int k = 0; k < 6; ++k) {
    const float v = pred[6 * i + k];
    x[k] = isfinite(v) ? v : 0.0f;
  }
  const float qn = x[3] * x[3] + x[4] * x[4] + x[5] * x[5];
  if (qn > 0.99f) {
    const float sc = sqrtf(0.99f / qn);
    x[3] *= sc;
    x[4] *= sc;
    x[5] *= sc;
  }
  float T[16];
  v2t<float>(x, T);
  set_start_state(Pp[i], T);
  for (int k = 0; k < 16; ++k) {
    starts16[16 * i + k] = Pp[i].state[kSlotT + k];
    starts16[16 * (n + i) + k] = Pp[n + i].state[kSlotT + k];
  }
}

// The choice: the prediction is kept unless the fallback has a lower chi2 per correspondence with
// at least half the prediction's correspondences (result slots 27 = chi2, 28 = inliers of the
// probe's linearisation), or the prediction has none; the refinement's pair i starts there.
__global__ __launch_bounds__(64) void k_choose_starts(const PairDev *__restrict__ probe, const PairDev *__restrict__ refine,
                                                      int n, const float *__restrict__ starts16,
                                                      float *__restrict__ chosen16, int *__restrict__ used) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  const double *a = probe[i].result, *f = probe[n + i].result;
  const bool net = !fallback_wins(a[27], a[28], f[27], f[28]);
  const float *T = starts16 + 16 * (size_t)(net ? i : n + i);
  for (int k = 0; k < 16; ++k) chosen16[16 * i + k] = T[k];
  used[i] = net ? 1 : 0;
  set_start_state(refine[i], T);
}

}  // namespace

extern "C" int nicp_batch_net_input(nicp_batch *b, nicp_initnet *n, int num_pairs, const int *ref_frame,
                                    const int *cur_frame, float max_depth_cov, float *dev_out) {
  CHK(check_net(n));
  if (!b || num_pairs <= 0 || !ref_frame || !cur_frame || !dev_out) return set_err(NICP_ERR_INVALID, "bad arguments");
  if (b->ctx != n->ctx) return set_err(NICP_ERR_INVALID, "batch and network belong to different contexts");
  CHK(check_net_frames(b, num_pairs, ref_frame, cur_frame));
  HIPCHK(hipSetDevice(b->ctx->device));
  return launch_net_input(n, b, num_pairs, ref_frame, cur_frame, max_depth_cov, dev_out);
}

extern "C" int nicp_batch_probe(nicp_batch *b, int num_pairs, const int *ref_frame, const int *cur_frame,
                                const float *starts16, const nicp_align_params *ap, int sub, int twin,
                                nicp_align_result *results) {
  if (!b || num_pairs <= 0 || !ref_frame || !cur_frame || !starts16 || !ap || !results || sub < 1)
    return set_err(NICP_ERR_INVALID, "bad arguments");
  if (twin && (num_pairs % 2 != 0)) return set_err(NICP_ERR_INVALID, "a twin probe needs an even number of pairs");
  for (int i = 0; twin && i < num_pairs / 2; ++i)
    if (ref_frame[i] != ref_frame[i + num_pairs / 2] || cur_frame[i] != cur_frame[i + num_pairs / 2])
      return set_err(NICP_ERR_INVALID, "twin pairs %d and %d have different frames", i, i + num_pairs / 2);
  nicp_align_params probe = *ap;
  probe.outer_iterations = 1;
  probe.inner_iterations = 1;
  std::vector<double> raw;
  CHK(batch_align(b, num_pairs, ref_frame, cur_frame, starts16, &probe, 0, nullptr, nullptr, false, &raw, sub,
                  twin ? num_pairs / 2 : 0));
  nicp_align_params none = probe;
  none.outer_iterations = 0;  // no statistics were computed
  for (int i = 0; i < num_pairs; ++i) fill_result(&raw[(size_t)i * kResultDoubles], &none, &results[i], 0.0f);
  return NICP_OK;
}

extern "C" int nicp_batch_align_learned(nicp_batch *b, nicp_initnet *n, int num_pairs, const int *ref_frame,
                                        const int *cur_frame, const float *fallback16, const nicp_align_params *ap,
                                        int use_graph, float *T_out16, nicp_align_result *results, float *start16,
                                        int *used_network) {
  CHK(check_net(n));
  if (!b || num_pairs <= 0 || !ref_frame || !cur_frame) return set_err(NICP_ERR_INVALID, "bad arguments");
  if (b->ctx != n->ctx) return set_err(NICP_ERR_INVALID, "batch and network belong to different contexts");
  CHK(check_align_params(ap));
  CHK(check_net_frames(b, num_pairs, ref_frame, cur_frame));
  nicp_ctx *ctx = b->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  Prof &pf = ctx->prof;
  pf.sink = &pf.eager;
  // Everything the device needs is prepared first, so that the network, the start probe, the
  // choice and the refinement then run back to back with no host round trip:
  // - the probe: pairs [0, n) start at the prediction (set on the device), [n, 2n) at the
  //   fallback guess; one linearisation over every kProbeSub-th row and column, the whole
  //   reference cloud projected (pairs i and n + i share their frames, so one pass reads each
  //   reference point for both starts);
  // - the refinement of the n pairs, whose starts the choice sets on the device.
  CHK(net_reserve(n, num_pairs));
  std::vector<float> init2((size_t)2 * num_pairs * 16);
  std::vector<int> ref2(2 * (size_t)num_pairs), cur2(2 * (size_t)num_pairs);
  for (int i = 0; i < num_pairs; ++i) {
    for (int k = 0; k < 16; ++k) {
      const float id = k % 5 == 0 ? 1.0f : 0.0f;
      init2[16 * (size_t)i + k] = id;
      init2[16 * ((size_t)num_pairs + i) + k] = fallback16 ? fallback16[16 * (size_t)i + k] : id;
    }
    ref2[i] = ref2[num_pairs + i] = ref_frame[i];
    cur2[i] = cur2[num_pairs + i] = cur_frame[i];
  }
  nicp_align_params probe = *ap;
  probe.outer_iterations = 1;
  probe.inner_iterations = 1;
  BatchCall pc, rc;
  CHK(batch_plan(b, probe_buffers(b), 2 * num_pairs, ref2.data(), cur2.data(), &probe, false, kProbeSub, num_pairs,
                 pc));
  CHK(batch_plan(b, batch_buffers(b), num_pairs, ref_frame, cur_frame, ap, true, 1, 0, rc));
  CHK(b->learned_aux.reserve(((size_t)3 * num_pairs * 16 + num_pairs) * sizeof(float)));
  float *d_starts = b->learned_aux.as<float>();
  float *d_chosen = d_starts + (size_t)2 * num_pairs * 16;
  int *d_used = (int *)(d_chosen + (size_t)num_pairs * 16);
  // all uploads through the batch's pinned staging: none of them waits for earlier work
  CHK(b->stage.begin(prepare_stage_bytes(pc.plan) + prepare_stage_bytes(rc.plan) +
                     PinnedStage::slot((size_t)2 * num_pairs * sizeof(InputJob))));
  CHK(batch_prepare(b, pc, init2.data(), &b->stage));
  CHK(batch_prepare(b, rc, nullptr, &b->stage));
  CHK(upload_net_jobs(n, b, num_pairs, ref_frame, cur_frame, &b->stage));
  CHK(b->stage.end(s));
  HIPCHK(hipEventRecord(ctx->ev0, s));
  // 1. prediction, 2. both starts judged by the probe, 3. the choice (the probe and the choice
  //    are timed as one stage: their kernels are not profiled one by one)
  auto enqueue_front = [&]() -> int {
    {
      ProfEntry en{};
      const bool p = pf.on;
      if (p) pf.begin(KID_NET_INPUT, s, en);
      CHK(enqueue_net_input(n, b, num_pairs, 0.01f, n->act[0]));
      if (p) pf.end(s, en);
    }
    {
      ProfEntry en{};
      const bool p = pf.on;
      if (p) pf.begin(KID_NET_FORWARD, s, en);
      CHK(net_forward(n, num_pairs));
      if (p) pf.end(s, en);
    }
    ProfEntry en{};
    const bool p = pf.on;
    if (p) pf.begin(KID_START_PROBE, s, en);
    pf.on = false;
    const unsigned g = (unsigned)((num_pairs + 63) / 64);
    hipLaunchKernelGGL(k_probe_starts, dim3(g), dim3(64), 0, s, (const float *)n->act[kLayers], num_pairs,
                       (const PairDev *)pc.d_pairs, d_starts);
    int r = hipGetLastError() == hipSuccess ? NICP_OK : set_err(NICP_ERR_HIP, "k_probe_starts launch failed");
    if (r == NICP_OK)
      r = launch_align(&pf, s, pc.plan, pc.d_pairs, b->d_frames, pc.d_bp, pc.d_br, pc.d_bc, &pc.ap, false,
                       pc.zstrategy);
    if (r == NICP_OK) {
      hipLaunchKernelGGL(k_choose_starts, dim3(g), dim3(64), 0, s, (const PairDev *)pc.d_pairs,
                         (const PairDev *)rc.d_pairs, num_pairs, (const float *)d_starts, d_chosen, d_used);
      if (hipGetLastError() != hipSuccess) r = set_err(NICP_ERR_HIP, "k_choose_starts launch failed");
    }
    pf.on = p;
    CHK(r);
    if (p) pf.end(s, en);
    return NICP_OK;
  };
  // 4. the refinement from the chosen starts. With use_graph the whole sequence is one graph,
  //    unless this call's refinement shape is being autotuned (eagerly, with its own timing) or
  //    the per-launch timing is on (then only the refinement is a graph, as nicp_batch_align's).
  if (use_graph && !rc.tuning && !pf.on) {
    std::vector<int> key = call_key(b, pc);
    const std::vector<int> rkey = call_key(b, rc);
    key.insert(key.end(), rkey.begin(), rkey.end());
    key.push_back(num_pairs);
    key.push_back(n->input_mode);
    for (int l = 0; l <= kLayers; ++l) key_ptr(key, n->act[l]);
    for (const void *ptr : {(const void *)n->jobs, (const void *)n->params, (const void *)n->chunks,
                            (const void *)d_starts})
      key_ptr(key, ptr);
    CHK(graph_replay(ctx, b->learned_graph, key, [&]() -> int {
      CHK(enqueue_front());
      return launch_align(&pf, s, rc.plan, rc.d_pairs, b->d_frames, rc.d_bp, rc.d_br, rc.d_bc, &rc.ap, true,
                          rc.zstrategy);
    }));
  } else {
    CHK(enqueue_front());
    CHK(batch_run(b, rc, use_graph, b->graph));
  }
  HIPCHK(hipEventRecord(ctx->ev1, s));
  if (start16) HIPCHK(hipMemcpyAsync(start16, d_chosen, (size_t)num_pairs * 16 * sizeof(float), hipMemcpyDeviceToHost, s));
  if (used_network) HIPCHK(hipMemcpyAsync(used_network, d_used, (size_t)num_pairs * sizeof(int), hipMemcpyDeviceToHost, s));
  CHK(batch_collect(b, rc, T_out16, results, nullptr));
  HIPCHK(hipStreamSynchronize(s));
  return NICP_OK;
}
