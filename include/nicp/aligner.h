// nicp::CorrespondenceFinder (nicp/nicp/correspondencefinder.h:14-208), nicp::Linearizer
// (nicp/nicp/linearizer.h:19-155) and the nicp::Aligner base (nicp/nicp/aligner.h:16-429):
// the reference's setters, getters and defaults shared by AlignerProjective and AlignerNN,
// whose align() runs the device loops of nicp_align() and nicp_align_nn().
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>

#include "cloud.h"
#include "definitions.h"
#include "pinholepointprojector.h"

namespace nicp {

class CorrespondenceFinder {
 public:
  virtual ~CorrespondenceFinder() {}
  // defaults of nicp/nicp/correspondencefinder.cpp:9-18
  float inlierDistanceThreshold() const { return _inlierDistanceThreshold; }
  void setInlierDistanceThreshold(float v) { _inlierDistanceThreshold = v; }
  float squaredThreshold() const { return _inlierDistanceThreshold * _inlierDistanceThreshold; }
  float inlierNormalAngularThreshold() const { return _inlierNormalAngularThreshold; }
  void setInlierNormalAngularThreshold(float v) { _inlierNormalAngularThreshold = v; }
  float flatCurvatureThreshold() const { return _flatCurvatureThreshold; }
  void setFlatCurvatureThreshold(float v) { _flatCurvatureThreshold = v; }
  float inlierCurvatureRatioThreshold() const { return _inlierCurvatureRatioThreshold; }
  void setInlierCurvatureRatioThreshold(float v) { _inlierCurvatureRatioThreshold = v; }
  bool demotedToGICP() const { return _demotedToGICP; }
  void setDemotedToGICP(bool v) { _demotedToGICP = v; }
  int numCorrespondences() const { return (int)_correspondences.size(); }
  const CorrespondenceVector &correspondences() const { return _correspondences; }
  void setCorrespondences(const CorrespondenceVector &c) { _correspondences = c; }

 protected:
  void _setPairs(const std::vector<int> &pairs, int n) {
    _correspondences.resize(n);
    for (int i = 0; i < n; ++i) _correspondences[i] = Correspondence(pairs[2 * i], pairs[2 * i + 1]);
  }

  float _inlierDistanceThreshold = 1.0f;
  float _inlierNormalAngularThreshold = std::cos((float)M_PI / 6.0f);
  float _flatCurvatureThreshold = 0.2f;
  float _inlierCurvatureRatioThreshold = 1.3f;
  bool _demotedToGICP = false;
  CorrespondenceVector _correspondences;
};

class Aligner;

class Linearizer {
 public:
  // defaults of nicp/nicp/linearizer.cpp:9-19
  float inlierMaxChi2() const { return _inlierMaxChi2; }
  void setInlierMaxChi2(float v) { _inlierMaxChi2 = v; }
  bool robustKernel() const { return _robustKernel; }
  void setRobustKernel(bool v) { _robustKernel = v; }
  bool demotedToGeneralizedICP() const { return _demotedToGeneralizedICP; }
  void setDemotedToGeneralizedICP(bool v) { _demotedToGeneralizedICP = v; }
  float scale() const { return _scale; }
  void setScale(float v) { _scale = v; }
  const Isometry3f &T() const { return _T; }
  void setT(const Isometry3f &T) { _T = T; }
  void setAligner(Aligner *a) { _aligner = a; }
  Aligner *aligner() const { return _aligner; }
  const Matrix6d &H() const { return _H; }
  const Vector6d &b() const { return _b; }
  float error() const { return _error; }
  int inliers() const { return _inliers; }
  // Linearizer::update over the aligner's clouds and its finder's correspondences
  inline void update();

  // internal: results of a device alignment
  void setResults(const Matrix6d &H, const Vector6d &b, float e, int inl) {
    _H = H;
    _b = b;
    _error = e;
    _inliers = inl;
  }

 private:
  float _inlierMaxChi2 = 9e3f;
  bool _robustKernel = true, _demotedToGeneralizedICP = false;
  float _scale = 1.0f;
  Isometry3f _T;
  Aligner *_aligner = nullptr;
  Matrix6d _H;
  Vector6d _b;
  float _error = 0.0f;
  int _inliers = 0;
};

class Aligner {
 public:
  Aligner() {}
  Aligner(PointProjector *p, Linearizer *l, CorrespondenceFinder *c)
      : _projector(p), _linearizer(l), _correspondenceFinder(c) {
    if (l) l->setAligner(this);
  }
  virtual ~Aligner() {}
  void setProjector(PointProjector *p) { _projector = p; }
  PointProjector *projector() const { return _projector; }
  // Aligner::setLinearizer writes the back-pointer (nicp/nicp/aligner.h:212-216)
  void setLinearizer(Linearizer *l) {
    _linearizer = l;
    if (l) l->setAligner(this);
  }
  Linearizer *linearizer() const { return _linearizer; }
  void setCorrespondenceFinder(CorrespondenceFinder *c) { _correspondenceFinder = c; }
  CorrespondenceFinder *correspondenceFinder() const { return _correspondenceFinder; }
  void setReferenceCloud(const Cloud *c) { _referenceCloud = c; }
  const Cloud *referenceCloud() const { return _referenceCloud; }
  void setCurrentCloud(const Cloud *c) { _currentCloud = c; }
  const Cloud *currentCloud() const { return _currentCloud; }
  int outerIterations() const { return _outerIterations; }
  void setOuterIterations(int v) { _outerIterations = v; }
  int innerIterations() const { return _innerIterations; }
  void setInnerIterations(int v) { _innerIterations = v; }
  float lambda() const { return _lambda; }
  void setLambda(float v) { _lambda = v; }
  int minInliers() const { return _minInliers; }
  void setMinInliers(int v) { _minInliers = v; }
  float translationalMinEigenRatio() const { return _translationalMinEigenRatio; }
  void setTranslationalMinEigenRatio(float v) { _translationalMinEigenRatio = v; }
  float rotationalMinEigenRatio() const { return _rotationalMinEigenRatio; }
  void setRotationalMinEigenRatio(float v) { _rotationalMinEigenRatio = v; }
  void setDebug(bool d) { _debug = d; }
  // Aligner::setInitialGuess zeroes the bottom row of the guess
  void setInitialGuess(const Isometry3f &g) {
    _initialGuess = g;
    _initialGuess.setBottomRow();
  }
  const Isometry3f &initialGuess() const { return _initialGuess; }
  // Aligner::setSensorOffset sets both offsets; each setter zeroes the bottom row
  // (nicp/nicp/aligner.h:157-198). AlignerProjective applies them (alignerprojective.cpp:36,51);
  // AlignerNN ignores them, as the reference's does.
  void setSensorOffset(const Isometry3f &o) {
    setReferenceSensorOffset(o);
    setCurrentSensorOffset(o);
  }
  const Isometry3f &sensorOffset() const { return _referenceSensorOffset; }
  void setReferenceSensorOffset(const Isometry3f &o) {
    _referenceSensorOffset = o;
    _referenceSensorOffset.setBottomRow();
  }
  const Isometry3f &referenceSensorOffset() const { return _referenceSensorOffset; }
  void setCurrentSensorOffset(const Isometry3f &o) {
    _currentSensorOffset = o;
    _currentSensorOffset.setBottomRow();
  }
  const Isometry3f &currentSensorOffset() const { return _currentSensorOffset; }
  const Isometry3f &T() const { return _T; }
  float error() const { return _error; }
  int inliers() const { return _inliers; }
  float totalTime() const { return _totalTime; }
  bool solutionValid() const { return _solutionValid; }
  const Matrix6d &omega() const { return _omega; }
  double translationalEigenRatio() const { return _result.translational_ratio; }
  double rotationalEigenRatio() const { return _result.rotational_ratio; }

  // the thresholds of the finder and the linearizer, the iteration counts and the solver
  // settings as the device loops take them
  virtual nicp_align_params params() const {
    nicp_align_params a;
    const CorrespondenceFinder *c = _correspondenceFinder;
    a.inlier_distance_threshold = c->inlierDistanceThreshold();
    a.inlier_normal_angular_threshold = c->inlierNormalAngularThreshold();
    a.flat_curvature_threshold = c->flatCurvatureThreshold();
    a.inlier_curvature_ratio_threshold = c->inlierCurvatureRatioThreshold();
    a.row_search_region = 0;
    a.col_search_region = 0;
    a.demoted_to_gicp = c->demotedToGICP();
    a.inlier_max_chi2 = _linearizer->inlierMaxChi2();
    a.robust_kernel = _linearizer->robustKernel();
    a.normal_omega_scale = _linearizer->scale();
    a.outer_iterations = _outerIterations;
    a.inner_iterations = _innerIterations;
    a.lambda = _lambda;
    a.translational_min_eigen_ratio = _translationalMinEigenRatio;
    a.rotational_min_eigen_ratio = _rotationalMinEigenRatio;
    return a;
  }

  virtual void align() = 0;

 protected:
  void _check(const char *who) const {
    if (!_linearizer || !_correspondenceFinder || !_referenceCloud || !_currentCloud)
      throw std::runtime_error(std::string("nicp_amd: ") + who + " is missing a component");
  }
  // results of a device alignment into the aligner and its linearizer
  void _adopt(const nicp_align_result &r) {
    _result = r;
    _error = r.error;
    _inliers = r.inliers;
    _totalTime = r.total_time_ms;
    _solutionValid = r.solution_valid != 0;
    std::memcpy(_omega.m, r.omega, sizeof(_omega.m));
    Matrix6d H;
    Vector6d b;
    std::memcpy(H.m, r.H, sizeof(H.m));
    std::memcpy(b.v, r.b, sizeof(b.v));
    _linearizer->setResults(H, b, _error, _inliers);
    _linearizer->setT(_T.inverse());
  }
  // the correspondences of the last device alignment into the finder
  void _fetchCorrespondences(CorrespondenceVector &cv) const {
    int n = 0;
    check(nicp_last_correspondences(Context::get(), nullptr, 0, &n));
    std::vector<int> pairs(2 * (size_t)std::max(n, 1));
    check(nicp_last_correspondences(Context::get(), pairs.data(), n, &n));
    cv.resize(n);
    for (int i = 0; i < n; ++i) cv[i] = Correspondence(pairs[2 * i], pairs[2 * i + 1]);
  }

  PointProjector *_projector = nullptr;
  Linearizer *_linearizer = nullptr;
  CorrespondenceFinder *_correspondenceFinder = nullptr;
  const Cloud *_referenceCloud = nullptr, *_currentCloud = nullptr;
  // defaults of nicp/nicp/aligner.cpp:21-41
  int _outerIterations = 10, _innerIterations = 1, _minInliers = 100;
  float _lambda = 1e3f, _translationalMinEigenRatio = 50.0f, _rotationalMinEigenRatio = 50.0f;
  bool _debug = false, _solutionValid = false;
  Isometry3f _initialGuess, _T, _referenceSensorOffset, _currentSensorOffset;
  float _error = 0.0f, _totalTime = 0.0f;
  int _inliers = 0;
  Matrix6d _omega;
  nicp_align_result _result{};
};

// Linearizer::update: one device linearisation of the finder's correspondences at _T
inline void Linearizer::update() {
  if (!_aligner) throw std::runtime_error("nicp_amd: Linearizer has no aligner");
  const CorrespondenceVector &cv = _aligner->correspondenceFinder()->correspondences();
  std::vector<int> pairs(2 * cv.size());
  for (size_t i = 0; i < cv.size(); ++i) {
    pairs[2 * i] = cv[i].referenceIndex;
    pairs[2 * i + 1] = cv[i].currentIndex;
  }
  nicp_align_params ap = _aligner->params();
  double e = 0.0;
  check(nicp_linearize(Context::get(), _aligner->referenceCloud()->handle(), _aligner->currentCloud()->handle(),
                       pairs.empty() ? nullptr : pairs.data(), (int)cv.size(), _T.matrix(), &ap, _H.m, _b.v, &e,
                       &_inliers));
  _error = (float)e;
}

}  // namespace nicp
