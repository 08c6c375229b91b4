// nicp::StatsCalculatorIntegralImage (nicp/nicp/statscalculatorintegralimage.h:14-133),
// nicp::Point/NormalInformationMatrixCalculator (nicp/nicp/informationmatrixcalculator.h:14-171)
// and nicp::DepthImageConverterIntegralImage (nicp/nicp/depthimageconverterintegralimage.h:15-63):
// parameter holders with the reference's setters and defaults; compute() runs
// the whole conversion on the device.
#pragma once

#include "cloud.h"
#include "definitions.h"
#include "pinholepointprojector.h"

namespace nicp {

class StatsCalculatorIntegralImage {
 public:
  float worldRadius() const { return _worldRadius; }
  void setWorldRadius(float v) { _worldRadius = v; }
  int maxImageRadius() const { return _maxImageRadius; }
  void setMaxImageRadius(int v) { _maxImageRadius = v; }
  int minImageRadius() const { return _minImageRadius; }
  void setMinImageRadius(int v) { _minImageRadius = v; }
  int minPoints() const { return _minPoints; }
  void setMinPoints(int v) { _minPoints = v; }
  float curvatureThreshold() const { return _curvatureThreshold; }
  void setCurvatureThreshold(float v) { _curvatureThreshold = v; }

 private:
  // nicp/nicp/statscalculatorintegralimage.cpp:12-18
  float _worldRadius = 0.1f;
  int _maxImageRadius = 30, _minImageRadius = 10, _minPoints = 50;
  float _curvatureThreshold = 0.02f;
};

class InformationMatrixCalculator {
 public:
  float curvatureThreshold() const { return _curvatureThreshold; }
  void setCurvatureThreshold(float v) { _curvatureThreshold = v; }
  // diagonals of the flat / non-flat information matrices in the eigenbasis
  const float *flatInformationMatrix() const { return _flat; }
  void setFlatInformationMatrix(float a, float b, float c) { _flat[0] = a; _flat[1] = b; _flat[2] = c; }
  const float *nonFlatInformationMatrix() const { return _nonFlat; }
  void setNonFlatInformationMatrix(float a, float b, float c) { _nonFlat[0] = a; _nonFlat[1] = b; _nonFlat[2] = c; }

 protected:
  float _flat[3] = {1.0f, 1.0f, 1.0f}, _nonFlat[3] = {1.0f, 1.0f, 1.0f};
  float _curvatureThreshold = 0.02f;
};

// nicp/nicp/informationmatrixcalculator.h:115-119
class PointInformationMatrixCalculator : public InformationMatrixCalculator {
 public:
  PointInformationMatrixCalculator() {
    setFlatInformationMatrix(1000.0f, 0.001f, 0.001f);
    setNonFlatInformationMatrix(1.0f, 0.1f, 0.1f);
  }
};

// nicp/nicp/informationmatrixcalculator.h:150-155
class NormalInformationMatrixCalculator : public InformationMatrixCalculator {
 public:
  NormalInformationMatrixCalculator() {
    setFlatInformationMatrix(1000.0f, 0.001f, 0.001f);
    setNonFlatInformationMatrix(0.1f, 1.0f, 1.0f);
  }
};

// DepthImage_convert_16UC1_to_32FC1 (nicp/nicp/imageutils.cpp:85-99)
inline void DepthImage_convert_16UC1_to_32FC1(DepthImage &dest, const RawDepthImage &src, float scale = 0.001f) {
  dest.create(src.rows, src.cols);
  for (size_t i = 0; i < src.data.size(); ++i) dest.data[i] = src.data[i] ? scale * (float)src.data[i] : 0.0f;
}

// DepthImage_scale (nicp/nicp/imageutils.cpp:6-37), on the device
inline void DepthImage_scale(DepthImage &dest, const DepthImage &src, int step, float maxDepthCov = 0.01f) {
  dest.create(src.rows / step, src.cols / step);
  if (dest.data.empty()) return;
  check(nicp_depth_scale(Context::get(), nullptr, src.data.data(), src.rows, src.cols, 1.0f, step, maxDepthCov,
                         dest.data.data()));
}

class DepthImageConverterIntegralImage {
 public:
  DepthImageConverterIntegralImage(PointProjector *projector, StatsCalculatorIntegralImage *statsCalculator,
                                   PointInformationMatrixCalculator *pointInformationMatrixCalculator,
                                   NormalInformationMatrixCalculator *normalInformationMatrixCalculator)
      : _projector(projector),
        _statsCalculator(statsCalculator),
        _pointCalc(pointInformationMatrixCalculator),
        _normalCalc(normalInformationMatrixCalculator) {}

  nicp_stats_params statsParams() const {
    nicp_stats_params s;
    s.world_radius = _statsCalculator->worldRadius();
    s.min_image_radius = _statsCalculator->minImageRadius();
    s.max_image_radius = _statsCalculator->maxImageRadius();
    s.min_points = _statsCalculator->minPoints();
    s.curvature_threshold = _statsCalculator->curvatureThreshold();
    std::memcpy(s.point_flat, _pointCalc->flatInformationMatrix(), sizeof(s.point_flat));
    std::memcpy(s.point_nonflat, _pointCalc->nonFlatInformationMatrix(), sizeof(s.point_nonflat));
    std::memcpy(s.normal_flat, _normalCalc->flatInformationMatrix(), sizeof(s.normal_flat));
    std::memcpy(s.normal_nonflat, _normalCalc->nonFlatInformationMatrix(), sizeof(s.normal_nonflat));
    s.point_curvature_threshold = _pointCalc->curvatureThreshold();
    s.normal_curvature_threshold = _normalCalc->curvatureThreshold();
    return s;
  }

  // DepthImage_convert_16UC1_to_32FC1 + DepthImageConverterIntegralImage::compute
  // (nicp/nicp/depthimageconverterintegralimage.cpp:18-64); the cloud is moved by the
  // sensor offset last (cloud.transformInPlace, line 63).
  void compute(Cloud &cloud, const RawDepthImage &depth, float depthScale,
               const Isometry3f &sensorOffset = Isometry3f::Identity()) {
    _projector->setImageSize(depth.rows, depth.cols);
    nicp_camera cam = _projector->camera();
    nicp_stats_params sp = statsParams();
    nicp_cloud *h = nullptr;
    check(nicp_ctx_set_fp64_eigen_gap(Context::get(), _fp64EigenGap));
    check(nicp_cloud_from_depth_u16(Context::get(), depth.data.data(), depth.rows, depth.cols, depthScale, &cam, &sp,
                                    _keepStats ? 1 : 0, &h));
    cloud.reset(h);
    // the projector's unProject fills Cloud::gaussians() on every conversion
    check(nicp_cloud_gaussians(h, _projector->baseline(), _projector->alpha()));
    _indexImage.create(depth.rows, depth.cols);
    check(nicp_cloud_index_image(h, _indexImage.data.data()));
    applySensorOffset(h, sensorOffset);
  }
  // DepthImageConverterIntegralImage::compute(cloud, depthImage, sensorOffset) for a float image in metres
  void compute(Cloud &cloud, const DepthImage &depth, const Isometry3f &sensorOffset = Isometry3f::Identity()) {
    _projector->setImageSize(depth.rows, depth.cols);
    nicp_camera cam = _projector->camera();
    nicp_stats_params sp = statsParams();
    nicp_cloud *h = nullptr;
    check(nicp_ctx_set_fp64_eigen_gap(Context::get(), _fp64EigenGap));
    check(nicp_cloud_from_depth_f32(Context::get(), depth.data.data(), depth.rows, depth.cols, &cam, &sp,
                                    _keepStats ? 1 : 0, &h));
    cloud.reset(h);
    // the projector's unProject fills Cloud::gaussians() on every conversion
    check(nicp_cloud_gaussians(h, _projector->baseline(), _projector->alpha()));
    _indexImage.create(depth.rows, depth.cols);
    check(nicp_cloud_index_image(h, _indexImage.data.data()));
    applySensorOffset(h, sensorOffset);
  }
  const IntImage &indexImage() const { return _indexImage; }
  // FP64 eigensolve of near-degenerate windows (no reference counterpart; see
  // nicp_ctx_set_fp64_eigen_gap): 0, the default, keeps the reference's FP32 solve
  float fp64EigenGap() const { return _fp64EigenGap; }
  void setFP64EigenGap(float gap) { _fp64EigenGap = gap; }
  PointProjector *projector() const { return _projector; }
  void setProjector(PointProjector *p) { _projector = p; }
  // The reference's clouds always hold their Stats (nicp/nicp/cloud.h:190); here they are
  // kept on request only (64 B per point), e.g. for Cloud::save with real Stats records.
  bool keepStats() const { return _keepStats; }
  void setKeepStats(bool k) { _keepStats = k; }

 private:
  // Cloud::transformInPlace(sensorOffset) (depthimageconverterintegralimage.cpp:63,
  // cloud.cpp:191-204): bottom row zeroed, skipped for the identity
  static void applySensorOffset(nicp_cloud *h, const Isometry3f &sensorOffset) {
    Isometry3f o = sensorOffset;
    o.setBottomRow();
    if (!o.isIdentity()) check(nicp_cloud_transform(h, o.matrix()));
  }
  bool _keepStats = false;
  float _fp64EigenGap = 0.0f;
  PointProjector *_projector;
  StatsCalculatorIntegralImage *_statsCalculator;
  PointInformationMatrixCalculator *_pointCalc;
  NormalInformationMatrixCalculator *_normalCalc;
  IntImage _indexImage;
};

}  // namespace nicp
