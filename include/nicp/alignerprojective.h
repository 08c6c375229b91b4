// nicp::CorrespondenceFinderProjective (nicp/nicp/correspondencefinderprojective.h:17-167) and
// nicp::AlignerProjective (nicp/nicp/alignerprojective.h): the reference's setters, getters
// and defaults over the device loop of nicp_align().
#pragma once

#include "aligner.h"

namespace nicp {

class CorrespondenceFinderProjective : public CorrespondenceFinder {
 public:
  // defaults of nicp/nicp/correspondencefinderprojective.cpp:9-14
  int rowSearchRegion() const { return _rowSearchRegion; }
  void setRowSearchRegion(int v) { _rowSearchRegion = v; }
  int colSearchRegion() const { return _colSearchRegion; }
  void setColSearchRegion(int v) { _colSearchRegion = v; }
  void setImageSize(int rows, int cols) {
    _rows = rows;
    _cols = cols;
  }
  int imageRows() const { return _rows; }
  int imageCols() const { return _cols; }

  // CorrespondenceFinderProjective::compute for explicit index images
  void compute(const Cloud &reference, const Cloud &current, const IntImage &referenceIndexImage,
               const IntImage &currentIndexImage, const Isometry3f &T, const nicp_align_params &ap) {
    std::vector<int> pairs(2 * (size_t)referenceIndexImage.rows * referenceIndexImage.cols);
    int n = 0;
    check(nicp_correspondences(Context::get(), reference.handle(), current.handle(), referenceIndexImage.data.data(),
                               currentIndexImage.data.data(), referenceIndexImage.rows, referenceIndexImage.cols,
                               T.matrix(), &ap, pairs.data(), &n));
    _setPairs(pairs, n);
  }

 private:
  int _rowSearchRegion = 0, _colSearchRegion = 0;
  int _rows = 0, _cols = 0;
};

class AlignerProjective : public Aligner {
 public:
  AlignerProjective() {}
  AlignerProjective(PointProjector *p, Linearizer *l, CorrespondenceFinderProjective *c) : Aligner(p, l, c) {}
  void setCorrespondenceFinder(CorrespondenceFinderProjective *c) { _correspondenceFinder = c; }
  CorrespondenceFinderProjective *correspondenceFinder() const {
    return static_cast<CorrespondenceFinderProjective *>(_correspondenceFinder);
  }

  nicp_align_params params() const override {
    nicp_align_params a = Aligner::params();
    a.row_search_region = correspondenceFinder()->rowSearchRegion();
    a.col_search_region = correspondenceFinder()->colSearchRegion();
    return a;
  }

  // AlignerProjective::align (nicp/nicp/alignerprojective.cpp:21-136)
  void align() override {
    if (!_projector) throw std::runtime_error("nicp_amd: AlignerProjective is missing a component");
    _check("AlignerProjective");
    nicp_camera cam = _projector->camera();
    nicp_align_params ap = params();
    nicp_align_result r{};
    check(nicp_align_offsets(Context::get(), _referenceCloud->handle(), _currentCloud->handle(), &cam,
                             _projector->imageRows(), _projector->imageCols(), _initialGuess.matrix(),
                             _referenceSensorOffset.matrix(), _currentSensorOffset.matrix(), &ap, _T.matrix(), &r));
    _adopt(r);
    CorrespondenceVector cv;
    _fetchCorrespondences(cv);
    _correspondenceFinder->setCorrespondences(cv);
  }
};

}  // namespace nicp
