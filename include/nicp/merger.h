// nicp::Merger (nicp/nicp/merger.h, merger.cpp:5-109): folds the points of a cloud that
// land within distanceThreshold of the nearest point of their pixel, seen from a given
// pose through the converter's projector, into that point's Gaussian, moves the surviving
// points to their Gaussians' means and drops the folded ones. merge() runs on the device.
#pragma once

#include <cmath>
#include <iostream>

#include "cloud.h"
#include "definitions.h"
#include "depthimageconverterintegralimage.h"

namespace nicp {

class Merger {
 public:
  Merger() {}
  float distanceThreshold() const { return _distanceThreshold; }
  void setDistanceThreshold(float v) { _distanceThreshold = v; }
  // kept for the interface; the reference's merge does not read it (merger.cpp:63-66)
  float normalThreshold() const { return _normalThreshold; }
  void setNormalThreshold(float v) { _normalThreshold = v; }
  float maxPointDepth() const { return _maxPointDepth; }
  void setMaxPointDepth(float v) { _maxPointDepth = v; }
  DepthImageConverterIntegralImage *depthImageConverter() const { return _converter; }
  void setDepthImageConverter(DepthImageConverterIntegralImage *c) { _converter = c; }
  // the index / depth images are made at the projector's image size, as in the reference
  void setImageSize(int rows, int cols) {
    _rows = rows;
    _cols = cols;
  }
  int numSuppressed() const { return _suppressed; }

  // Merger::merge (merger.cpp:15-109); the projector's own transform is not changed
  void merge(Cloud *cloud, const Isometry3f &transform = Isometry3f::Identity()) {
    if (_rows <= 0 || _cols <= 0) throw std::runtime_error("Merger: image size is zero");
    if (!_converter || !_converter->projector()) throw std::runtime_error("Merger: missing depth image converter");
    PointProjector *p = _converter->projector();
    if (p->imageRows() <= 0 || p->imageCols() <= 0) throw std::runtime_error("Merger: the projector has no image size");
    nicp_camera cam = p->camera();
    nicp_merger_params mp;
    mp.distance_threshold = _distanceThreshold;
    mp.normal_threshold = _normalThreshold;
    mp.max_point_depth = _maxPointDepth;
    const size_t before = cloud->size();
    check(nicp_merge(cloud->handle(), transform.matrix(), &cam, p->imageRows(), p->imageCols(), &mp, &_suppressed,
                     nullptr));
    std::cout << "[INFO]: number of suppressed points " << _suppressed << std::endl;
    std::cout << "[INFO]: resized cloud from " << before << " to " << cloud->size() << " points" << std::endl;
  }

 private:
  // merger.cpp:5-14
  float _distanceThreshold = 0.1f;
  float _normalThreshold = std::cos(10.0f * (float)M_PI / 180.0f);
  float _maxPointDepth = 10.0f;
  DepthImageConverterIntegralImage *_converter = nullptr;
  int _rows = 0, _cols = 0;
  int _suppressed = 0;
};

}  // namespace nicp
