/*
 * CPU oracle for the NICP registration loop — TEST INFRASTRUCTURE ONLY.
 *
 * This header declares a plain-C ABI over a C++ restatement of the arithmetic
 * of yorsh87/nicp (files cited per function in nicp_oracle.cpp). Only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark may load
 * it, and only as the checker / the timed CPU reference. The product library
 * never links or calls it.
 *
 * Conventions (same as the reference, nicp/nicp/definitions.h:37-74):
 *  - images are row-major rows x cols;
 *  - 4x4 matrices are float[16] column-major (Eigen's default);
 *  - points / normals are float4 with w = 1 / w = 0;
 *  - 3x3 eigenvector matrices are float[9] column-major (column 0 = normal).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  float K[9];              /* camera matrix, column-major */
  float min_distance;      /* N/pointprojector.cpp:9-10 */
  float max_distance;
  int   model;             /* 0 PinholePointProjector, 1 SphericalPointProjector */
  float horizontal_fov;    /* spherical: half fields of view, image centre (N/sphericalpointprojector.h:304-313) */
  float vertical_fov;
  float horizontal_center;
  float vertical_center;
} orc_camera;

typedef struct {
  float world_radius;      /* N/statscalculatorintegralimage.cpp:13-17 */
  int   min_image_radius;
  int   max_image_radius;
  int   min_points;
  float curvature_threshold;
  /* information matrices, N/informationmatrixcalculator.h:115-155 */
  float point_flat[3], point_nonflat[3];
  float normal_flat[3], normal_nonflat[3];
  float point_curvature_threshold;
  float normal_curvature_threshold;
} orc_stats_params;

typedef struct {
  float inlier_distance_threshold;        /* N/correspondencefinder.cpp:9-18 */
  float inlier_normal_angular_threshold;
  float flat_curvature_threshold;
  float inlier_curvature_ratio_threshold;
  int   row_search_region;                /* N/correspondencefinderprojective.cpp:12-13 */
  int   col_search_region;
  int   demoted_to_gicp;
  float inlier_max_chi2;                  /* N/linearizer.cpp:9-19 */
  int   robust_kernel;
  float normal_omega_scale;               /* Linearizer::_scale */
  int   outer_iterations;                 /* N/aligner.cpp:21-41 */
  int   inner_iterations;
  float lambda;
  float translational_min_eigen_ratio;
  float rotational_min_eigen_ratio;
} orc_align_params;

/* A cloud as the oracle sees it (host arrays, n points). */
typedef struct {
  int n;
  const float *points;     /* n*4 */
  const float *normals;    /* n*4 */
  const float *curvature;  /* n   */
  const float *omega_p;    /* n*16 */
  const float *omega_n;    /* n*16 */
} orc_cloud;

typedef struct {
  float error;
  int   inliers;
  int   num_correspondences;
  float H[36];             /* column-major 6x6 of the last linearisation */
  float b[6];
  double mean[6];
  double omega[36];
  double translational_ratio;
  double rotational_ratio;
  int   solution_valid;
} orc_align_result;

/* A1 */
void orc_depth_u16_to_f32(const uint16_t *raw, int n, float scale, float *out);
/* DepthImage_scale (N/imageutils.cpp:6-37): (rows/step) x (cols/step) output */
void orc_depth_scale(const float *src, int rows, int cols, int step, float max_depth_cov, float *dst);
/* A2 */
void orc_inverse3(const float K[9], float iK[9]);
void orc_projector_matrices(const float K[9], const float transform[16], float KRt[16], float iKRt[16]);
/* A3: returns the point count; points must hold rows*cols*4 floats */
int orc_unproject(const float *depth, int rows, int cols, const float iKRt[16],
                  float min_d, float max_d, float *points, int *index_image);
/* A4: interval image, 2 ints per pixel */
void orc_project_intervals(const float *depth, int rows, int cols, const float K[9],
                           float min_d, float max_d, float world_radius, int *intervals);
/* A6: integral image, 10 floats per pixel (n, sx, sy, sz, sxx, sxy, sxz, syy, syz, szz) */
void orc_integral_image(const int *index_image, int rows, int cols, const float *points, float *ii);
/* A7 */
void orc_get_region(const float *ii, int rows, int cols, int xmin, int xmax, int ymin, int ymax, float out[10]);
/* A8/A9 third-party restatement (Eigen SelfAdjointEigenSolver<Matrix3f>::computeDirect) */
void orc_eigen3_direct(const float cov[9], float evals[3], float evecs[9]);
/* A5/A9/A10/A11: full frame conversion (DepthImageConverterIntegralImage::compute).
 * Outputs sized for rows*cols points; returns the point count. Any output may be NULL. */
int orc_frame_from_depth(const uint16_t *raw, int rows, int cols, float depth_scale,
                         const orc_camera *cam, const orc_stats_params *sp,
                         int *index_image, float *points, float *normals, float *curvature,
                         float *eigvecs, float *eigvals, float *means, int *npts,
                         float *omega_p, float *omega_n);
/* FP64 brute-force companion: per point, the covariance of the same 2R x 2R window
 * evaluated directly in double, eigen-decomposed by cyclic Jacobi. normals64 = n*3,
 * curvature64 = n, valid = n (1 where the FP32 path produced stats). */
void orc_stats_f64_bruteforce(const uint16_t *raw, int rows, int cols, float depth_scale,
                              const orc_camera *cam, const orc_stats_params *sp,
                              const int *index_image, const float *points, int n,
                              double *normals64, double *curvature64, int *valid);
/* A14 */
void orc_project(const float *points, int n, const float KRt[16], int rows, int cols,
                 float min_d, float max_d, int *index_image, float *depth_image);
/* A15/A16: T is the correspondence-finder transform (the aligner passes T^-1). */
int orc_correspondences(const orc_cloud *ref, const orc_cloud *cur,
                        const int *ref_index, const int *cur_index, int rows, int cols,
                        const float T[16], const orc_align_params *ap, int num_threads,
                        int *pairs /* 2 per pair, capacity rows*cols */);
/* A17/A18: FP32 restatement (num_threads mimics the OpenMP partition) */
void orc_linearize(const orc_cloud *ref, const orc_cloud *cur, const int *pairs, int num_pairs,
                   const float T[16], const orc_align_params *ap, int num_threads,
                   float H[36], float b[6], float *error, int *inliers);
/* FP64 twin on identical float inputs */
void orc_linearize_f64(const orc_cloud *ref, const orc_cloud *cur, const int *pairs, int num_pairs,
                       const float T[16], const orc_align_params *ap,
                       double H[36], double b[6], double *error, int *inliers);
/* OpenMP threads of the conversion (omp_set_num_threads) and the current maximum */
void orc_set_num_threads(int n);
int orc_get_max_threads(void);
/* A19 */
void orc_v2t(const float x[6], float T[16]);
void orc_t2v(const float T[16], float x[6]);
void orc_isometry_mul(const float A[16], const float B[16], float C[16]);
void orc_isometry_inverse(const float A[16], float Ai[16]);
int  orc_ldlt_solve_f32(const float H[36], const float rhs[6], float x[6]);
int  orc_ldlt_solve_f64(const double H[36], const double rhs[6], double x[6]);
/* A13 + A20: AlignerProjective::align. The projector uses cam; the reference
 * and current sensor offsets are identity. */
void orc_align(const orc_cloud *ref, const orc_cloud *cur, int rows, int cols,
               const orc_camera *cam, const float init_guess[16], const orc_align_params *ap,
               int num_threads, float T_out[16], orc_align_result *res);
/* mode 0 = orc_align; mode 1 = the same loop with H, b and the LDLT in FP64 (updates in FP32) */
void orc_align_offsets(const orc_cloud *ref, const orc_cloud *cur, int rows, int cols, const orc_camera *cam,
                       const float init_guess[16], const float ref_offset[16], const float cur_offset[16],
                       const orc_align_params *ap, int num_threads, int mode, float T_out[16], orc_align_result *res);
void orc_align_mode(const orc_cloud *ref, const orc_cloud *cur, int rows, int cols,
                    const orc_camera *cam, const float init_guess[16], const orc_align_params *ap,
                    int num_threads, int mode, float T_out[16], orc_align_result *res);
void orc_align_nn_mode(const orc_cloud *ref, const orc_cloud *cur, const float init_guess[16],
                       const orc_align_params *ap, float normal_scaling, int mode, float T_out[16],
                       orc_align_result *res);
/* A22: CorrespondenceFinderNN::compute (exact brute-force 1-NN in 6D within the squared
 * inlier distance; ties -> lower index) followed by the reference's gates. pairs: 2 per ref
 * point capacity; returns the number of pairs, in reference-index order. nn_out (optional,
 * n_ref ints) receives the raw nearest neighbour per reference point (-1 when none). */
int orc_nn_correspondences(const orc_cloud *ref, const orc_cloud *cur, const float T[16], const orc_align_params *ap,
                           float normal_scaling, int *pairs, int *nn_out);
/* A22: AlignerNN::align (nicp/nicp/alignernn.cpp:21-118) with the brute-force finder */
void orc_align_nn(const orc_cloud *ref, const orc_cloud *cur, const float init_guess[16], const orc_align_params *ap,
                  float normal_scaling, float T_out[16], orc_align_result *res);
/* A20 only: statistics of a given H at a given T (FP64). */
void orc_compute_statistics(const double H[36], const float T[16], const orc_align_params *ap,
                            orc_align_result *res);

/* SphericalPointProjector (N/sphericalpointprojector.h:252-313, .cpp:19-141) at an image size:
 * unProject with the projector transform T (the converter passes the identity), the interval
 * image, and the z-buffer projection seen from `transform` (its _iT = transform^-1). */
int orc_spherical_unproject(const float *depth, int rows, int cols, const orc_camera *cam, const float T[16],
                            float *points, int *index_image);
void orc_spherical_intervals(const float *depth, int rows, int cols, const orc_camera *cam, float world_radius,
                             int *intervals);
void orc_spherical_project(const float *points, int n, const float transform[16], const orc_camera *cam, int rows,
                           int cols, int *index_image, float *depth_image);

/* Gaussian3f (N/gaussian.h:8-133) with its lazy-evaluation state: flags bit 0 = mean and
 * covariance current, bit 1 = information form current. Same layout as nicp_gaussian. */
typedef struct {
  float mean[3];
  int   flags;
  float covariance[9];
  float information[9];
  float information_vector[3];
  float pad[3];
} orc_gaussian;

/* PinholePointProjector::unProject with Gaussians (N/pinholepointprojector.cpp:93-134) */
int orc_unproject_gaussians(const float *depth, int rows, int cols, const float K[9], const float iKRt[16],
                            float min_d, float max_d, float baseline, float alpha, float *points,
                            orc_gaussian *gaussians, int *index_image);
/* Cloud::transformInPlace (N/cloud.cpp:193-205); stats16 / omegas are full column-major 4x4 */
void orc_cloud_transform(int n, const float T[16], float *points, float *normals, float *stats16, float *omega_p,
                         float *omega_n, orc_gaussian *gaussians);
/* Merger::merge (N/merger.cpp:15-109) in place; returns the new size */
int orc_merge(int n, float *points, float *normals, float *curvature, float *stats16, float *omega_p, float *omega_n,
              orc_gaussian *gaussians, const float transform[16], const orc_camera *cam, int rows, int cols,
              float distance_threshold, float max_point_depth, int *collapsed);

#ifdef __cplusplus
}
#endif
