// nicp::SphericalPointProjector (nicp/nicp/sphericalpointprojector.h:13-313,
// nicp/nicp/sphericalpointprojector.cpp:8-141): the panoramic / laser projector. Pixels are
// azimuth and elevation over half fields of view around an image centre; depth is the
// distance from the vertical axis. The converter, aligner and merger take it wherever they
// take a PinholePointProjector.
#pragma once

#include <cmath>

#include "pinholepointprojector.h"

namespace nicp {

class SphericalPointProjector : public PointProjector {
 public:
  // constructor defaults (sphericalpointprojector.cpp:8-17)
  SphericalPointProjector() {
    _imageRows = _originalImageRows = 480;
    _imageCols = _originalImageCols = 640;
    _horizontalFov = (float)M_PI;
    _verticalFov = (float)M_PI / 4.0f;
    _horizontalCenter = _imageCols / 2.0f;
    _verticalCenter = _imageRows / 2.0f;
  }
  float horizontalFov() const { return _horizontalFov; }
  void setHorizontalFov(float v) { _horizontalFov = v; }
  float verticalFov() const { return _verticalFov; }
  void setVerticalFov(float v) { _verticalFov = v; }
  float horizontalCenter() const { return _horizontalCenter; }
  void setHorizontalCenter(float v) { _horizontalCenter = v; }
  float verticalCenter() const { return _verticalCenter; }
  void setVerticalCenter(float v) { _verticalCenter = v; }

  nicp_camera camera() const override {
    nicp_camera c{};
    c.K[0] = c.K[4] = c.K[8] = 1.0f;
    c.min_distance = _minDistance;
    c.max_distance = _maxDistance;
    c.model = NICP_SPHERICAL;
    c.horizontal_fov = _horizontalFov;
    c.vertical_fov = _verticalFov;
    c.horizontal_center = _horizontalCenter;
    c.vertical_center = _verticalCenter;
    return c;
  }

 private:
  float _horizontalFov, _verticalFov, _horizontalCenter, _verticalCenter;
};

}  // namespace nicp
