/*
 * nicp_amd — C ABI of the MI355X (gfx950) implementation of NICP's
 * registration loop. Plain pointers and sizes only; every function returns
 * an int status (NICP_OK = 0, negative on error) and never throws.
 * nicp_last_error() gives a message for the last failure on the calling thread.
 *
 * Conventions (as in yorsh87/nicp):
 *  - images are row-major rows x cols (cv::Mat_ layout, nicp/nicp/definitions.h:37-74);
 *  - 4x4 transforms are float[16] column-major (Eigen::Isometry3f::matrix());
 *  - points are float4 (x, y, z, 1), normals float4 (x, y, z, 0)
 *    (nicp/nicp/homogeneousvector4f.h:73-88);
 *  - information matrices are float[16] column-major 4x4 with row/col 3 zero
 *    (nicp/nicp/informationmatrix.h:13-84);
 *  - an index image holds a point index or -1.
 *  - T brings the current cloud onto the reference (nicp/nicp/alignerprojective.cpp:51,61).
 *
 * Threading: a context owns one HIP stream and its scratch memory; it is not
 * thread-safe. Separate contexts are independent. Host buffers passed in are
 * borrowed for the duration of the call only.
 */
#ifndef NICP_AMD_H
#define NICP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NICP_OK 0
#define NICP_ERR_INVALID (-1)
#define NICP_ERR_HIP (-2)
#define NICP_ERR_OOM (-3)
#define NICP_ERR_STATE (-4)

typedef struct nicp_ctx nicp_ctx;
typedef struct nicp_cloud nicp_cloud;
typedef struct nicp_batch nicp_batch;
typedef struct nicp_initnet nicp_initnet;

#define NICP_PINHOLE 0
#define NICP_SPHERICAL 1

/* Projector state. model NICP_PINHOLE: PinholePointProjector (nicp/nicp/pinholepointprojector.h:67-70,
 * nicp/nicp/pointprojector.cpp:9-10), K the (already scaled) camera matrix. model NICP_SPHERICAL:
 * SphericalPointProjector (nicp/nicp/sphericalpointprojector.h:252-313, .cpp:9-27): half fields of
 * view in radians and the image centre in pixels; the angular resolutions follow from the image size
 * of each call (cols / (2 horizontal_fov), rows / (2 vertical_fov)). A zero-initialised tail is the
 * pinhole model. */
typedef struct {
  float K[9];               /* column-major 3x3 (pinhole) */
  float min_distance;       /* must be > 0 */
  float max_distance;
  int   model;              /* NICP_PINHOLE or NICP_SPHERICAL */
  float horizontal_fov;     /* spherical: reference defaults pi, pi/4 */
  float vertical_fov;
  float horizontal_center;  /* spherical: reference defaults cols/2, rows/2 of its 640x480 default */
  float vertical_center;
} nicp_camera;

/* StatsCalculatorIntegralImage (nicp/nicp/statscalculatorintegralimage.cpp:12-18) and
 * Point/NormalInformationMatrixCalculator (nicp/nicp/informationmatrixcalculator.h:115-155). */
typedef struct {
  float world_radius;
  int   min_image_radius;
  int   max_image_radius;
  int   min_points;
  float curvature_threshold;
  float point_flat[3], point_nonflat[3];
  float normal_flat[3], normal_nonflat[3];
  float point_curvature_threshold;
  float normal_curvature_threshold;
} nicp_stats_params;

/* CorrespondenceFinder (nicp/nicp/correspondencefinder.cpp:9-18,
 * nicp/nicp/correspondencefinderprojective.cpp:9-14), Linearizer
 * (nicp/nicp/linearizer.cpp:9-19) and Aligner (nicp/nicp/aligner.cpp:21-41). */
typedef struct {
  float inlier_distance_threshold;
  float inlier_normal_angular_threshold;   /* a cosine, compared with n_cur . n_ref */
  float flat_curvature_threshold;
  float inlier_curvature_ratio_threshold;
  int   row_search_region;                 /* 0,0 = single-pixel search */
  int   col_search_region;
  int   demoted_to_gicp;
  float inlier_max_chi2;
  int   robust_kernel;
  float normal_omega_scale;                /* Linearizer::_scale */
  int   outer_iterations;
  int   inner_iterations;
  float lambda;
  float translational_min_eigen_ratio;
  float rotational_min_eigen_ratio;
} nicp_align_params;

typedef struct {
  float  error;                /* Aligner::error(): chi2 of the last linearisation */
  int    inliers;              /* Aligner::inliers() */
  int    num_correspondences;  /* CorrespondenceFinder::numCorrespondences() of the last iteration */
  double H[36];                /* last linearisation (no lambda), column-major 6x6 */
  double b[6];
  double mean[6];              /* Aligner::_computeStatistics outputs */
  double omega[36];
  double translational_ratio;
  double rotational_ratio;
  int    solution_valid;       /* Aligner::solutionValid(); 0 also when H is zero */
  float  total_time_ms;        /* Aligner::totalTime(): device time of the call */
  int    overlap_pixels;       /* pixels of the last iteration where both index images hold a point */
} nicp_align_result;

const char *nicp_last_error(void);

/* context: one device, one stream (Aligner / converter object state) */
int nicp_ctx_create(int device, nicp_ctx **out);
int nicp_ctx_destroy(nicp_ctx *ctx);
int nicp_ctx_synchronize(nicp_ctx *ctx);
/* the HIP stream the context launches on (hipStream_t) */
void *nicp_ctx_stream(nicp_ctx *ctx);
/* FP64 eigensolve of near-degenerate windows (no reference counterpart: the reference's
 * StatsCalculatorIntegralImage::compute, nicp/nicp/statscalculatorintegralimage.cpp:60-61,
 * always solves in FP32). gap > 0: every later conversion of this context solves a pixel's
 * window covariance again in FP64 when its FP32 eigenvalues have a relative gap
 * (l1 - l0) / l2 below gap, and takes those eigenvalues, eigenvectors, normal and curvature;
 * the normals are then the eigenvectors of the same FP32 covariance to FP32 rounding, which
 * near degeneracy departs from the reference's FP32 normals by up to ~1e-3. gap = 0 (the
 * default) keeps the reference's FP32 solve everywhere. */
int nicp_ctx_set_fp64_eigen_gap(nicp_ctx *ctx, float gap);
/* plain device memory on the context's device (no reference counterpart; for callers that
 * keep datasets or network inputs resident). kind: 0 host->device, 1 device->host,
 * 2 device->device; copies complete before the call returns. */
int nicp_device_alloc(nicp_ctx *ctx, size_t bytes, void **ptr);
int nicp_device_free(nicp_ctx *ctx, void *ptr);
int nicp_device_copy(nicp_ctx *ctx, void *dst, const void *src, size_t bytes, int kind);

/* Replaces DepthImage_convert_16UC1_to_32FC1 (nicp/nicp/imageutils.cpp:85-99) +
 * DepthImageConverterIntegralImage::compute (nicp/nicp/depthimageconverterintegralimage.cpp:18-64)
 * with an identity sensor offset. keep_stats also keeps the per-point Stats
 * (eigenvectors, eigenvalues, mean, n) for download. */
int nicp_cloud_from_depth_u16(nicp_ctx *ctx, const uint16_t *depth, int rows, int cols, float depth_scale,
                              const nicp_camera *cam, const nicp_stats_params *sp, int keep_stats,
                              nicp_cloud **out);
/* The same conversion from a float depth image in metres (DepthImage, 32FC1). */
int nicp_cloud_from_depth_f32(nicp_ctx *ctx, const float *depth, int rows, int cols, const nicp_camera *cam,
                              const nicp_stats_params *sp, int keep_stats, nicp_cloud **out);
/* Replaces DepthImage_scale (nicp/nicp/imageutils.cpp:6-37): (rows/step) x (cols/step) block means
 * of the non-zero depths, 0 where the block variance exceeds max_depth_cov (reference default 0.01).
 * Input is uint16 times depth_scale, or float metres when depthf is given. */
int nicp_depth_scale(nicp_ctx *ctx, const uint16_t *depth, const float *depthf, int rows, int cols, float depth_scale,
                     int step, float max_depth_cov, float *out);
/* Replaces filling a Cloud's vectors (nicp/nicp/cloud.h:20-193): points, normals, the curvature
 * of Stats, and the two information matrices; any of normals/curvature/omegas may be NULL (zero). */
int nicp_cloud_from_host(nicp_ctx *ctx, int n, const float *points4, const float *normals4, const float *curvature,
                         const float *omega_p16, const float *omega_n16, nicp_cloud **out);
int nicp_cloud_destroy(nicp_cloud *cloud);
int nicp_cloud_size(const nicp_cloud *cloud, int *n);
/* Cloud::points/normals/stats().curvature()/pointInformationMatrix/normalInformationMatrix */
int nicp_cloud_download(nicp_cloud *cloud, float *points4, float *normals4, float *curvature, float *omega_p16,
                        float *omega_n16);
/* Cloud::stats(): eigenvectors (9 floats, column-major), eigenvalues (3), mean (4), n */
int nicp_cloud_download_stats(nicp_cloud *cloud, float *eigvecs9, float *eigvals3, float *means4, int *npts);
/* the index image of DepthImageConverter::_indexImage (rows*cols ints) */
int nicp_cloud_index_image(nicp_cloud *cloud, int *index_image);

/* Gaussian3f (nicp/nicp/gaussian.h:8-133) with the state of its lazy conversions between
 * the moments form (mean, covariance) and the information form; 112 bytes, the layout of
 * the device records. A form whose flag is clear holds stale values. */
typedef struct {
  float mean[3];
  int   flags;                   /* bit 0: mean/covariance current; bit 1: information form current */
  float covariance[9];           /* column-major */
  float information[9];          /* column-major */
  float information_vector[3];
  float pad[3];
} nicp_gaussian;

/* Merger (nicp/nicp/merger.cpp:5-14) */
typedef struct {
  float distance_threshold;      /* _distanceThreshold, default 0.1 */
  float normal_threshold;        /* _normalThreshold, default cos(10 deg); unused by the reference's merge */
  float max_point_depth;         /* _maxPointDepth, default 10 */
} nicp_merger_params;

/* Replaces the Gaussian half of PinholePointProjector::unProject (nicp/nicp/pinholepointprojector.cpp:93-134),
 * which DepthImageConverter::compute (nicp/nicp/depthimageconverterintegralimage.cpp:42) runs on every
 * conversion: per point, mean = the point and covariance J diag(3, 3, zVar) J^T with
 * J = K^-1 [[z,0,c],[0,z,r],[0,0,1]], zVar = alpha z^2 / (baseline K(0,0) + z alpha). The cloud must come
 * from a depth conversion (it needs the index and depth images). baseline / alpha are the projector's
 * (pinholepointprojector.cpp:11-12 defaults 0.075, 0.1). */
int nicp_cloud_gaussians(nicp_cloud *cloud, float baseline, float alpha);
/* Cloud::gaussians() of a host-built cloud (n records), and their download */
int nicp_cloud_set_gaussians(nicp_cloud *cloud, const nicp_gaussian *gaussians);
int nicp_cloud_download_gaussians(nicp_cloud *cloud, nicp_gaussian *gaussians);
/* Replaces Cloud::transformInPlace (nicp/nicp/cloud.cpp:193-205): points, normals, Stats, Gaussians
 * and both information matrices; nothing changes when T (with its last row set to 0 0 0 1) is the
 * identity. Information matrices are kept as their upper triangles. */
int nicp_cloud_transform(nicp_cloud *cloud, const float T[16]);
/* Replaces Cloud::add (nicp/nicp/cloud.cpp:163-190): transforms src in place by T, then appends it to
 * dst. dst becomes a plain point cloud (no index image); it keeps Stats when both clouds have them
 * and Gaussians when both have them (an error when only dst has them). */
int nicp_cloud_add(nicp_cloud *dst, nicp_cloud *src, const float T[16]);
/* Replaces Merger::merge (nicp/nicp/merger.cpp:15-109): projects the cloud from `transform`
 * with cam into a rows x cols z-buffer, folds the information of every point whose depth lies
 * within distance_threshold of its pixel's winner into the winner's Gaussian (in point order),
 * moves every winner that kept its own pixel to its Gaussian's mean, and compacts the cloud in
 * order. num_suppressed (optional) receives the number of removed points; collapsed (optional,
 * the old size) receives per point -1 (kept as is), its own index (kept, moved to the mean) or
 * the winner it was folded into. The cloud must have Gaussians. */
int nicp_merge(nicp_cloud *cloud, const float transform[16], const nicp_camera *cam, int rows, int cols,
               const nicp_merger_params *params, int *num_suppressed, int *collapsed);

/* Replaces PinholePointProjector::project (nicp/nicp/pinholepointprojector.cpp:33-66):
 * z-buffer projection of the cloud seen from `transform`; depth_image may be NULL. */
int nicp_project(nicp_ctx *ctx, const nicp_cloud *cloud, const float transform[16], const nicp_camera *cam,
                 int rows, int cols, int *index_image, float *depth_image);

/* Replaces CorrespondenceFinderProjective::compute (nicp/nicp/correspondencefinderprojective.cpp:16-21)
 * for given index images; T is the finder's transform (the aligner passes T^-1). pairs receives
 * (referenceIndex, currentIndex) in row-major pixel order; capacity rows*cols pairs. */
int nicp_correspondences(nicp_ctx *ctx, const nicp_cloud *ref, const nicp_cloud *cur, const int *ref_index,
                         const int *cur_index, int rows, int cols, const float T[16], const nicp_align_params *ap,
                         int *pairs, int *num_pairs);

/* Replaces Linearizer::update (nicp/nicp/linearizer.cpp:21-134) over an explicit pair list. */
int nicp_linearize(nicp_ctx *ctx, const nicp_cloud *ref, const nicp_cloud *cur, const int *pairs, int num_pairs,
                   const float T[16], const nicp_align_params *ap, double H[36], double b[6], double *error,
                   int *inliers);

/* Replaces AlignerProjective::align (nicp/nicp/alignerprojective.cpp:21-136) with identity
 * sensor offsets and no priors; the image size and camera are the projector's. */
int nicp_align(nicp_ctx *ctx, const nicp_cloud *ref, const nicp_cloud *cur, const nicp_camera *cam, int rows,
               int cols, const float init_guess[16], const nicp_align_params *ap, float T_out[16],
               nicp_align_result *result);
/* AlignerProjective::align with the aligner's sensor offsets (Aligner::setReferenceSensorOffset /
 * setCurrentSensorOffset, nicp/nicp/aligner.h:157-196): the current cloud is projected from
 * current_offset (nicp/nicp/alignerprojective.cpp:36-39) and the reference cloud from
 * T * reference_offset (nicp/nicp/alignerprojective.cpp:51). Column-major 4x4 each, the last row
 * taken as 0 0 0 1; NULL is the identity. */
int nicp_align_offsets(nicp_ctx *ctx, const nicp_cloud *ref, const nicp_cloud *cur, const nicp_camera *cam, int rows,
                       int cols, const float init_guess[16], const float reference_offset[16],
                       const float current_offset[16], const nicp_align_params *ap, float T_out[16],
                       nicp_align_result *result);
/* Replaces Aligner::_computeStatistics (nicp/nicp/aligner.cpp:82-135, nicp/nicp/unscented.h:11-65)
 * for a given statistics linearisation H (column-major 6x6) at the final transform T: fills
 * result->mean, omega, translational_ratio, rotational_ratio and solution_valid (the other fields
 * are left unchanged). It runs the device computation every align call runs after its statistics
 * linearisation (FP64, pseudo-inverse of H, unscented sigma points through t2v(T v2t(s)^-1)). */
int nicp_solution_statistics(nicp_ctx *ctx, const double H[36], const float T[16], const nicp_align_params *ap,
                             nicp_align_result *result);
/* No counterpart in the reference: nicp_align for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). It
 * calls no hipMalloc/hipFree: its z-buffer, index image, correspondences, partial sums, loop
 * state and tables live in one fixed workspace, a 7 MiB array in device memory that each GPU
 * the library runs on holds until the process ends. Limits: rows * cols <= 307200 (640x480),
 * rows <= 960, cols <= 1280 and at most 307200 points in each cloud; larger inputs return
 * NICP_ERR_INVALID. All
 * contexts on one GPU share that workspace, so calls that overlap on one GPU (other threads or
 * streams) race; each call rewrites what it reads. Results equal nicp_align's bit for bit;
 * nicp_last_correspondences has nothing to return after it. */
int nicp_align_noalloc(nicp_ctx *ctx, const nicp_cloud *ref, const nicp_cloud *cur, const nicp_camera *cam, int rows,
                       int cols, const float init_guess[16], const nicp_align_params *ap, float T_out[16],
                       nicp_align_result *result);
/* CorrespondenceFinder::correspondences() of the last nicp_align on ctx (capacity in pairs) */
int nicp_last_correspondences(nicp_ctx *ctx, int *pairs, int capacity, int *num_pairs);

/* Replaces CorrespondenceFinderNN::init + compute (nicp/nicp/correspondencefindernn.cpp:44-97):
 * the nearest neighbour of each transformed reference point [T p; s R n] among the current
 * points [p; s n] within the squared inlier distance (exact, on a uniform grid of cell_size
 * metres, <= 0 for the default of 2 cm, which changes the speed only; FLANN's approximate
 * kd-tree is not reproduced; ties go to the lower index), then the reference's gates. pairs (capacity: reference points) are in
 * reference-index order; nn_out (optional, one int per reference point) receives the raw
 * neighbour or -1. */
int nicp_nn_correspondences(nicp_ctx *ctx, const nicp_cloud *ref, const nicp_cloud *cur, const float T[16],
                            const nicp_align_params *ap, float normal_scaling, float cell_size, int *pairs,
                            int *num_pairs, int *nn_out);
/* Replaces AlignerNN::align (nicp/nicp/alignernn.cpp:21-118); one inner iteration. */
int nicp_align_nn(nicp_ctx *ctx, const nicp_cloud *ref, const nicp_cloud *cur, const float init_guess[16],
                  const nicp_align_params *ap, float normal_scaling, float cell_size, float T_out[16],
                  nicp_align_result *result);
/* Batched form of nicp_align_nn (no reference counterpart; it replaces one AlignerNN::align
 * per cloud pair on host threads): num_pairs (1..65536) pairs of clouds of any sizes are
 * aligned in the same launches, each pair's transform and result equal to nicp_align_nn on
 * that pair alone. init_guesses: 16 floats per pair (column-major) or NULL for identities;
 * T_out: 16 floats per pair; results: one per pair, total_time_ms being the whole call's.
 * nicp_last_correspondences has nothing to return after it. */
int nicp_align_nn_batch(nicp_ctx *ctx, int num_pairs, const nicp_cloud *const *refs, const nicp_cloud *const *curs,
                        const float *init_guesses, const nicp_align_params *ap, float normal_scaling,
                        float cell_size, float *T_out, nicp_align_result *results);

/* Batched form (no reference counterpart; it replaces running one converter/aligner per
 * frame pair on host threads): frames of uneven size live in one device arena, depth
 * is uploaded once, then converted and aligned pair-wise in single launches per stage. */
int nicp_batch_create(nicp_ctx *ctx, int num_frames, const int *rows, const int *cols, const nicp_camera *cams,
                      const nicp_stats_params *sp, nicp_batch **out);
int nicp_batch_destroy(nicp_batch *batch);
int nicp_batch_upload_depth(nicp_batch *batch, int frame, const uint16_t *depth, float depth_scale);
/* Depth already in device memory (no reference counterpart; it serves callers whose frames
 * arrive in HBM, and pipelines that keep several batches in flight): every frame's uint16 image
 * back to back in frame order (frame i starts after the rows*cols pixels of the frames before
 * it; nicp_batch_num_pixels gives the total). The copy is enqueued on the batch's stream and
 * the call returns without waiting: the source must stay valid until the stream passes it. */
int nicp_batch_upload_depth_device(nicp_batch *batch, const uint16_t *depth_device, float depth_scale);
/* the frames' depth images into device memory, in the same packed layout (enqueued, no wait) */
int nicp_batch_download_depth_device(nicp_batch *batch, uint16_t *depth_device);
int nicp_batch_num_pixels(const nicp_batch *batch, long long *num_pixels);
/* DepthImageConverterIntegralImage::compute for every frame of the batch */
int nicp_batch_convert(nicp_batch *batch);
/* AlignerProjective::align for pairs (ref_frame[i], cur_frame[i]); T_out and results may be NULL.
 * use_graph != 0 captures the iteration chain in a hipGraph and replays it. */
int nicp_batch_align(nicp_batch *batch, int num_pairs, const int *ref_frame, const int *cur_frame,
                     const float *init_guesses16, const nicp_align_params *ap, int use_graph, float *T_out16,
                     nicp_align_result *results);
/* number of points of a batch frame, and a cloud view of it (downloads only) */
int nicp_batch_frame_cloud(nicp_batch *batch, int frame, nicp_cloud **out);
/* Projection strategies of a batch iteration (no reference counterpart; they replace the serial
 * z-buffer loop of PinholePointProjector::project, nicp/nicp/pinholepointprojector.cpp:33-66, and
 * give the same z-buffer, correspondences and sums): 0 wave-merged 64-bit atomicMin into a global
 * z-buffer, then the correspondence kernel; 1 the same, reading each cell first and skipping atomics
 * that cannot lower it; 2 tile-binned: the reference frame's 16x32 pixel tiles are listed for the
 * 64x64 projector tiles their frustums may reach, and one workgroup per projector tile z-buffers its
 * listed points in LDS and runs the correspondences and the linearisation on them (pinhole batches
 * without window search; elsewhere 0 is used instead). */
/* Autotuner: with enable, the first nicp_batch_align call of each plan shape (pairs, image tiles,
 * cloud tiles, subsampling, statistics, tile-binned applicable) runs eagerly with the applicable
 * strategies on alternate iterations, and later calls of the shape use the fastest. Enabling or
 * disabling forgets earlier choices. Without the autotuner a batch uses strategy 2 where it
 * applies and the call has at least 1024 projector tiles of 64x64 pixels (13 pairs of 640x480),
 * else 0. */
int nicp_batch_set_autotune(nicp_batch *batch, int enable);
/* forces strategy 0, 1 or 2 on every later align call of the batch (2 where it applies, else 0);
 * -1 returns to the autotuner's or the default choice */
int nicp_batch_set_projection(nicp_batch *batch, int strategy);
/* list capacity of the tile-binned projection, 0..256 source tiles per projector tile (default
 * 256); a projector tile reached by more scans every source tile of its reference frame (same
 * results, slower). */
int nicp_batch_set_bin_capacity(nicp_batch *batch, int capacity);
/* the tuned shapes: choice per shape, ms = 3 per shape (average ms per iteration of the projection
 * and correspondence launches of strategies 0, 1, 2; -1 where unmeasured) */
int nicp_batch_autotune_read(nicp_batch *batch, int capacity, int *num_shapes, int *choice, float *ms);

/* ---- Synthetic scenes (no reference counterpart: the training-set and benchmark input
 * generator). Renders the seeded 20-primitive scene scene_seeds[i] into the depth of batch
 * frame frames[i] from the camera pose poses16[i] (column-major, camera in scene frame),
 * with the frame's intrinsics, depth noise sigma = noise_coeff * z^2 m, quantised to mm
 * (the frame's depth scale becomes 0.001) and hole_fraction random holes. streams[i] salts
 * the noise/hole draws. Same images as the host generator. */
int nicp_batch_render_synthetic(nicp_batch *batch, int num_frames, const int *frames, const uint64_t *scene_seeds,
                                const uint64_t *streams, const float *poses16, float noise_coeff, float hole_fraction);

/* ---- Learned initial guess (no reference counterpart: the reference's Aligner takes its
 * initial guess from the caller, nicp/nicp/aligner.h:102-106). A CNN maps the 60x80
 * DepthImage_scale reductions (nicp/nicp/imageutils.cpp:6-37) of a reference/current pair
 * to the pose 6-vector (t, qx, qy, qz) of v2t (nicp/nicp/bm_se3.h:38-45). Weights are fp32,
 * the layers run as bf16 matrix-core GEMMs with fp32 accumulation. Input tensors are
 * NHWC float [batch][60][80][2] (channel 0 = reference depth in metres, 1 = current). */
int nicp_initnet_num_params(void);
int nicp_initnet_create(nicp_ctx *ctx, uint64_t seed, nicp_initnet **out);
int nicp_initnet_destroy(nicp_initnet *net);
int nicp_initnet_get_params(nicp_initnet *net, float *params_out);
int nicp_initnet_set_params(nicp_initnet *net, const float *params); /* also resets the optimiser */
int nicp_initnet_set_loss_weights(nicp_initnet *net, const float weights6[6]);
/* input/out6 are device pointers when on_device != 0, host pointers otherwise */
int nicp_initnet_forward(nicp_initnet *net, const float *input, int batch, int on_device, float *out6);
/* one step of the weighted squared error against target6 ([batch][6]); lr > 0 applies Adam */
int nicp_initnet_train_step(nicp_initnet *net, const float *input, const float *target6, int batch, int on_device,
                            float lr, float *loss);
/* the same step on the minibatch rows[0..batch) of device-resident inputs ([num_rows][60][80][2]),
 * each transformed by its symmetry flags (NULL = none; bit 0 swaps the two frames, bit 1
 * mirrors the columns, bit 2 the rows), against host targets6 [batch][6] (the transformed
 * poses are the caller's). rows and flags are host arrays. */
int nicp_initnet_train_step_rows(nicp_initnet *net, const float *inputs, int num_rows, const int *rows,
                                 const int *flags, const float *targets6, int batch, float lr, float *loss);
/* 0: the first layer reads depth in metres (default); 1: inverse depth (0 stays 0) */
int nicp_initnet_set_input_mode(nicp_initnet *net, int mode);
/* copies of the last pass: what 0 = activation index (0 = input .. 6 = output), 1 = its
 * gradient, 2 = parameter gradient (index ignored); count floats */
int nicp_initnet_read(nicp_initnet *net, int what, int index, float *out, size_t count);
/* the network input of pairs (ref_frame[i], cur_frame[i]) of a batch into dev_out (device
 * pointer, [num_pairs][60][80][2]); frame sizes must be multiples of 80x60 */
int nicp_batch_net_input(nicp_batch *batch, nicp_initnet *net, int num_pairs, const int *ref_frame,
                         const int *cur_frame, float max_depth_cov, float *dev_out);
/* AlignerProjective::align (nicp/nicp/alignerprojective.cpp:21-136) from a learned start:
 * the network predicts a start per pair; one linearisation at the prediction and at the
 * fallback guess (fallback16, or the identity when NULL), over every 8th row and column of the
 * projector image with the whole reference cloud projected, judges both, and the prediction is
 * kept unless the fallback has a lower chi2 per correspondence with at least half the
 * prediction's correspondences (or the prediction has none). start16 (may be NULL) receives
 * the chosen starts, used_network (may be NULL) 1 where the prediction was kept.
 * T_out16/results as nicp_batch_align. */
int nicp_batch_align_learned(nicp_batch *batch, nicp_initnet *net, int num_pairs, const int *ref_frame,
                             const int *cur_frame, const float *fallback16, const nicp_align_params *ap, int use_graph,
                             float *T_out16, nicp_align_result *results, float *start16, int *used_network);

/* The start probe of nicp_batch_align_learned (no reference counterpart): one correspondence
 * search and linearisation (Linearizer::update, nicp/nicp/linearizer.cpp:21-134) of each pair at
 * its start (starts16: 16 floats per pair), over every sub-th row and column of the projector
 * image (sub >= 1; the reference cloud is projected whole). results receive H, b, error, inliers
 * and the correspondence count (no statistics). twin = 1 projects each reference point once for
 * pairs i and i + num_pairs / 2, which must have the same frames (the learned aligner's two
 * starts); the results are those of twin = 0. */
int nicp_batch_probe(nicp_batch *batch, int num_pairs, const int *ref_frame, const int *cur_frame,
                     const float *starts16, const nicp_align_params *ap, int sub, int twin,
                     nicp_align_result *results);

/* Per-kernel device time (no reference counterpart; the reference only has
 * Aligner::totalTime(), nicp/nicp/aligner.h:348). When enabled, every launch of the
 * context is bracketed by HIP events on its stream, and calls that would replay a hipGraph
 * launch eagerly instead (same results); enabling resets the totals. total_ms / launches
 * receive nicp_kernel_count() entries. */
int nicp_kernel_count(void);
const char *nicp_kernel_name(int kernel_id);
int nicp_ctx_profile(nicp_ctx *ctx, int enable);
int nicp_ctx_profile_read(nicp_ctx *ctx, double *total_ms, long long *launches);

#ifdef __cplusplus
}
#endif

#endif /* NICP_AMD_H */
