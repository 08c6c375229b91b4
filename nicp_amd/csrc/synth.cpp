// Seeded synthetic depth scenes for the NICP registration loop.
//
// A scene is 20 planar rectangles and boxes in front of a pinhole camera
// (a back wall, a floor and a side wall, then 17 random objects with
// extents of 0.5-4 m). Depth images are rendered by exact ray casting in
// double precision, perturbed with gaussian noise of sigma = k * z^2 m,
// quantised to uint16 millimetres and punched with uniform random holes
// (value 0). Every random draw comes from a counter-based hash of
// (seed, stream, index), so the output does not depend on the thread count.
//
// The second frame of a pair is the same scene seen from a camera moved by a
// random rigid motion T_gt (uniform axis, angle <= max_rot, translation
// uniform in a ball of radius max_trans). T_gt is the pose of the second
// camera in the first camera's frame, which is the transform the reference's
// Aligner::T() returns for (reference = first, current = second)
// (nicp/nicp/alignerprojective.cpp:51,61).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "synth_scene.h"

namespace {

using synth::Prim;
using synth::gaussian;
using synth::uniform;

// flat information matrix U diag(a, b, b) U^T = b I + (a - b) n n^T, 4x4 column-major
void flat_omega(const float *n, float a, float b, float *O) {
  for (int k = 0; k < 16; ++k) O[k] = 0.0f;
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) O[c * 4 + r] = (float)((r == c ? (double)b : 0.0) + ((double)a - b) * n[r] * n[c]);
}

void render(const std::vector<Prim> &prims, int rows, int cols, double fx, double fy, double cx, double cy,
            const double *R /*3x3 cm*/, const double *t, double noise, double holes, uint64_t seed,
            uint64_t stream, uint16_t *out) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c)
      out[(size_t)r * cols + c] = synth::depth_sample(prims.data(), (int)prims.size(), r, c, cols, fx, fy, cx, cy, R,
                                                      t, noise, holes, seed, stream);
}

}  // namespace

extern "C" {

// Renders the scene of `scene_seed` from the camera pose `pose` (float[16]
// column-major, pose of the camera in the scene frame). `stream` salts the
// noise and hole draws so that two renders of one scene are independent.
int nicp_synth_render(uint64_t scene_seed, uint64_t stream, int rows, int cols, float fx, float fy, float cx,
                      float cy, const float *pose, float noise_coeff, float hole_fraction, uint16_t *depth) {
  if (rows <= 0 || cols <= 0 || !depth || !pose) return -1;
  std::vector<Prim> prims;
  synth::make_scene(scene_seed, fx, fy, cols, rows, prims);
  double R[9], t[3];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) R[c * 3 + r] = pose[c * 4 + r];
  for (int r = 0; r < 3; ++r) t[r] = pose[12 + r];
  render(prims, rows, cols, fx, fy, cx, cy, R, t, noise_coeff, hole_fraction, scene_seed, stream, depth);
  return 0;
}

// Samples the rigid motion of pair `seed` (see the file comment).
void nicp_synth_motion(uint64_t seed, float max_rot_deg, float max_trans_m, float *T_gt) {
  synth::motion(seed, max_rot_deg, max_trans_m, T_gt);
}

// An unstructured cloud pair (the kNN configuration): n points sampled from the
// visible surfaces of scene `seed` (rendered noise-free at rows x cols), with the
// exact surface normals, curvature 0 and flat information matrices; the target is
// the source moved by T_gt^-1 with gaussian position noise `noise` (m), so that
// p_source = T_gt p_target. Arrays: points/normals n*4, curvature n, omegas n*16.
int nicp_synth_cloud_pair(uint64_t seed, int n, int rows, int cols, float fx, float fy, float cx, float cy,
                          float max_rot_deg, float max_trans_m, float noise, float *src_points, float *src_normals,
                          float *tgt_points, float *tgt_normals, float *curvature, float *omega_p, float *omega_n,
                          float *T_gt) {
  if (n <= 0 || rows <= 0 || cols <= 0) return -1;
  std::vector<Prim> prims;
  synth::make_scene(seed, fx, fy, cols, rows, prims);
  const size_t P = (size_t)rows * cols;
  std::vector<float> hit(P * 6);
  std::vector<char> ok(P, 0);
#pragma omp parallel for schedule(dynamic, 4)
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c) {
      double d[3] = {(c - cx) / fx, (r - cy) / fy, 1.0}, o[3] = {0, 0, 0}, nn[3] = {0, 0, 0};
      double t = synth::cast(prims.data(), (int)prims.size(), o, d, nn);
      size_t i = (size_t)r * cols + c;
      if (t > 0.5 && t < 4.5) {
        ok[i] = 1;
        for (int k = 0; k < 3; ++k) { hit[6 * i + k] = (float)(t * d[k]); hit[6 * i + 3 + k] = (float)nn[k]; }
      }
    }
  std::vector<size_t> valid;
  for (size_t i = 0; i < P; ++i)
    if (ok[i]) valid.push_back(i);
  if ((int)valid.size() < n) return -2;
  // partial Fisher-Yates with the counter-based hash
  for (int i = 0; i < n; ++i) {
    size_t j = i + (size_t)(uniform(seed, 0x3131, i) * (valid.size() - i));
    if (j >= valid.size()) j = valid.size() - 1;
    std::swap(valid[i], valid[j]);
  }
  nicp_synth_motion(seed, max_rot_deg, max_trans_m, T_gt);
  // inverse motion for the target: R^T (p - t)
  double R[9], t[3];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) R[c * 3 + r] = T_gt[c * 4 + r];
  for (int r = 0; r < 3; ++r) t[r] = T_gt[12 + r];
  for (int i = 0; i < n; ++i) {
    const float *h = &hit[6 * valid[i]];
    for (int k = 0; k < 3; ++k) {
      src_points[4 * i + k] = h[k];
      src_normals[4 * i + k] = h[3 + k];
    }
    src_points[4 * i + 3] = 1.0f;
    src_normals[4 * i + 3] = 0.0f;
    for (int k = 0; k < 3; ++k) {
      double pk = 0.0, nk = 0.0;
      for (int m = 0; m < 3; ++m) {
        pk += R[k * 3 + m] * ((double)h[m] - t[m]);
        nk += R[k * 3 + m] * (double)h[3 + m];
      }
      tgt_points[4 * i + k] = (float)(pk + noise * gaussian(seed, 0x4141 + k, i));
      tgt_normals[4 * i + k] = (float)nk;
    }
    tgt_points[4 * i + 3] = 1.0f;
    tgt_normals[4 * i + 3] = 0.0f;
    curvature[i] = 0.0f;
    flat_omega(&tgt_normals[4 * i], 1000.0f, 0.001f, omega_p + 16 * (size_t)i);
    flat_omega(&tgt_normals[4 * i], 1000.0f, 0.001f, omega_n + 16 * (size_t)i);
  }
  return 0;
}

// A registration pair: the reference frame at the identity, the current frame
// at T_gt. Both are uint16 rows x cols, values in millimetres.
int nicp_synth_pair(uint64_t seed, int rows, int cols, float fx, float fy, float cx, float cy, float max_rot_deg,
                    float max_trans_m, float noise_coeff, float hole_fraction, uint16_t *depth_ref,
                    uint16_t *depth_cur, float *T_gt) {
  if (rows <= 0 || cols <= 0 || !depth_ref || !depth_cur || !T_gt) return -1;
  nicp_synth_motion(seed, max_rot_deg, max_trans_m, T_gt);
  float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  nicp_synth_render(seed, 10, rows, cols, fx, fy, cx, cy, I, noise_coeff, hole_fraction, depth_ref);
  nicp_synth_render(seed, 11, rows, cols, fx, fy, cx, cy, T_gt, noise_coeff, hole_fraction, depth_cur);
  return 0;
}

}  // extern "C"
