// nicp::CorrespondenceFinderNN (nicp/nicp/correspondencefindernn.h:17-78) and nicp::AlignerNN
// (nicp/nicp/alignernn.h:15-48) over nicp_nn_correspondences() and nicp_align_nn(). The
// reference searches FLANN's approximate randomized kd-trees; the device search is the exact
// nearest neighbour within the squared inlier distance (ties to the lower index), on a uniform
// grid whose cell size (cellSize(), no reference counterpart) changes the speed only.
#pragma once

#include "aligner.h"

namespace nicp {

class CorrespondenceFinderNN : public CorrespondenceFinder {
 public:
  // _normal_scaling = 1 (nicp/nicp/correspondencefindernn.cpp:8)
  float normalScaling() const { return _normal_scaling; }
  void setNormalScaling(float v) { _normal_scaling = v; }
  // edge of the search grid's cells in metres; <= 0 picks the default (2 cm)
  float cellSize() const { return _cellSize; }
  void setCellSize(float v) { _cellSize = v; }

  // CorrespondenceFinderNN::init (correspondencefindernn.cpp:44-67): the device rebuilds the
  // grid of the current cloud on every compute, so there is nothing to prepare
  void init(const Cloud &, const Cloud &) {}

  // CorrespondenceFinderNN::compute (correspondencefindernn.cpp:69-97): correspondences of
  // the reference points moved by T, in reference-index order
  void compute(const Cloud &reference, const Cloud &current, const Isometry3f &T) {
    nicp_align_params ap{};
    ap.inlier_distance_threshold = _inlierDistanceThreshold;
    ap.inlier_normal_angular_threshold = _inlierNormalAngularThreshold;
    ap.flat_curvature_threshold = _flatCurvatureThreshold;
    ap.inlier_curvature_ratio_threshold = _inlierCurvatureRatioThreshold;
    ap.demoted_to_gicp = _demotedToGICP;
    ap.inner_iterations = 1;
    std::vector<int> pairs(2 * std::max<size_t>(reference.size(), 1));
    int n = 0;
    check(nicp_nn_correspondences(Context::get(), reference.handle(), current.handle(), T.matrix(), &ap,
                                  _normal_scaling, _cellSize, pairs.data(), &n, nullptr));
    _setPairs(pairs, n);
  }

 private:
  float _normal_scaling = 1.0f;
  float _cellSize = 0.0f;
};

class AlignerNN : public Aligner {
 public:
  AlignerNN() {}
  AlignerNN(PointProjector *p, Linearizer *l, CorrespondenceFinderNN *c) : Aligner(p, l, c) {}
  void setCorrespondenceFinder(CorrespondenceFinderNN *c) { _correspondenceFinder = c; }
  CorrespondenceFinderNN *correspondenceFinder() const {
    return static_cast<CorrespondenceFinderNN *>(_correspondenceFinder);
  }

  // AlignerNN::align (nicp/nicp/alignernn.cpp:21-118); the projector is not used
  void align() override {
    _check("AlignerNN");
    nicp_align_params ap = params();
    nicp_align_result r{};
    CorrespondenceFinderNN *c = correspondenceFinder();
    check(nicp_align_nn(Context::get(), _referenceCloud->handle(), _currentCloud->handle(), _initialGuess.matrix(),
                        &ap, c->normalScaling(), c->cellSize(), _T.matrix(), &r));
    _adopt(r);
    CorrespondenceVector cv;
    _fetchCorrespondences(cv);
    c->setCorrespondences(cv);
  }
};

}  // namespace nicp
