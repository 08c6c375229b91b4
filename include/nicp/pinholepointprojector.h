// nicp::PointProjector and nicp::PinholePointProjector (nicp/nicp/pointprojector.h:15-273,
// nicp/nicp/pinholepointprojector.h:13-323, nicp/nicp/pointprojector.cpp:6-112): camera
// matrix, distance range, image size, pose; project() runs the z-buffer projection on the
// device.
#pragma once

#include "cloud.h"
#include "definitions.h"

namespace nicp {

// nicp::PointProjector (nicp/nicp/pointprojector.h:15-273): distance range, image size,
// pose and scaling shared by the projectors; camera() gives the device descriptor.
class PointProjector {
 public:
  virtual ~PointProjector() {}
  float minDistance() const { return _minDistance; }
  void setMinDistance(float d) { _minDistance = d; }
  float maxDistance() const { return _maxDistance; }
  void setMaxDistance(float d) { _maxDistance = d; }
  virtual void setImageSize(int rows, int cols) {
    _originalImageRows = _imageRows = rows;
    _originalImageCols = _imageCols = cols;
  }
  int imageRows() const { return _imageRows; }
  int imageCols() const { return _imageCols; }
  const Isometry3f &transform() const { return _transform; }
  void setTransform(const Isometry3f &t) { _transform = t; }
  // PointProjector::scale (pointprojector.cpp:106-112)
  virtual void scale(float s) {
    _scalingFactor = s;
    _imageRows = (int)(_originalImageRows * s);
    _imageCols = (int)(_originalImageCols * s);
  }
  float scalingFactor() const { return _scalingFactor; }
  // sensor noise model of the pinhole Gaussians (pinholepointprojector.h:97-118, defaults .cpp:11-12)
  float baseline() const { return _baseline; }
  void setBaseline(float b) { _baseline = b; }
  float alpha() const { return _alpha; }
  void setAlpha(float a) { _alpha = a; }

  virtual nicp_camera camera() const = 0;

  // project(indexImage, depthImage, points): the z-buffer projection on the device
  void project(IntImage &indexImage, DepthImage &depthImage, const Cloud &cloud) const {
    indexImage.create(_imageRows, _imageCols);
    depthImage.create(_imageRows, _imageCols);
    nicp_camera c = camera();
    check(nicp_project(Context::get(), cloud.handle(), _transform.matrix(), &c, _imageRows, _imageCols,
                       indexImage.data.data(), depthImage.data.data()));
  }

 protected:
  float _minDistance = 0.01f, _maxDistance = 6.0f, _scalingFactor = 1.0f;
  float _baseline = 0.075f, _alpha = 0.1f;
  int _originalImageRows = 0, _originalImageCols = 0, _imageRows = 0, _imageCols = 0;
  Isometry3f _transform;
};

class PinholePointProjector : public PointProjector {
 public:
  PinholePointProjector() {
    // constructor defaults (nicp/nicp/pinholepointprojector.cpp:5-14, pointprojector.cpp:6-16)
    _originalCameraMatrix(0, 0) = 1.0f;
    _originalCameraMatrix(1, 1) = 1.0f;
    _originalCameraMatrix(0, 2) = 0.5f;
    _originalCameraMatrix(1, 2) = 0.5f;
    _originalCameraMatrix(2, 2) = 1.0f;
    _cameraMatrix = _originalCameraMatrix;
  }
  void setCameraMatrix(const Matrix3f &K) {
    _originalCameraMatrix = K;
    _cameraMatrix = K;
    applyScale();
  }
  const Matrix3f &cameraMatrix() const { return _cameraMatrix; }
  const Matrix3f &originalCameraMatrix() const { return _originalCameraMatrix; }
  // PinholePointProjector::scale (pinholepointprojector.cpp:150-155 + pointprojector.cpp:106-112)
  void scale(float s) override {
    PointProjector::scale(s);
    applyScale();
  }

  nicp_camera camera() const override {
    nicp_camera c{};
    std::memcpy(c.K, _cameraMatrix.m, sizeof(c.K));
    c.min_distance = _minDistance;
    c.max_distance = _maxDistance;
    c.model = NICP_PINHOLE;
    return c;
  }

 private:
  void applyScale() {
    _cameraMatrix = _originalCameraMatrix;
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 2; ++r) _cameraMatrix(r, c) *= _scalingFactor;
  }
  Matrix3f _originalCameraMatrix, _cameraMatrix;
};

}  // namespace nicp
