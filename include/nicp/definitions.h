// Value types of the nicp_amd C++ front end: the subset of the reference's
// Eigen/OpenCV vocabulary that its registration API exposes
// (nicp/nicp/definitions.h:37-74, nicp/nicp/homogeneousvector4f.h:73-88,
// nicp/nicp/bm_defs.h:10-27), as plain structs so that callers need neither
// Eigen nor OpenCV. Matrices are column-major like Eigen's defaults.
#pragma once

#include <stdint.h>

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../nicp_amd.h"

namespace nicp {

// Eigen::Isometry3f: a 4x4 column-major matrix whose bottom row is (0, 0, 0, 1)
struct Isometry3f {
  float m[16];
  Isometry3f() { setIdentity(); }
  static Isometry3f Identity() { return Isometry3f(); }
  void setIdentity() {
    std::memset(m, 0, sizeof(m));
    m[0] = m[5] = m[10] = m[15] = 1.0f;
  }
  float &operator()(int r, int c) { return m[c * 4 + r]; }
  float operator()(int r, int c) const { return m[c * 4 + r]; }
  float *matrix() { return m; }
  const float *matrix() const { return m; }
  void setBottomRow() { m[3] = m[7] = m[11] = 0.0f; m[15] = 1.0f; }
  // [R^T | -R^T t]
  Isometry3f inverse() const {
    Isometry3f o;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) o(i, j) = (*this)(j, i);
    for (int i = 0; i < 3; ++i)
      o(i, 3) = -(o(i, 0) * (*this)(0, 3) + (o(i, 1) * (*this)(1, 3) + o(i, 2) * (*this)(2, 3)));
    return o;
  }
  Isometry3f operator*(const Isometry3f &b) const {
    Isometry3f o;
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j)
        o(i, j) = (*this)(i, 0) * b(0, j) + ((*this)(i, 1) * b(1, j) + (*this)(i, 2) * b(2, j));
      o(i, 3) = ((*this)(i, 0) * b(0, 3) + ((*this)(i, 1) * b(1, 3) + (*this)(i, 2) * b(2, 3))) + (*this)(i, 3);
    }
    return o;
  }
  bool isIdentity() const {
    Isometry3f I;
    return std::memcmp(m, I.m, sizeof(m)) == 0;
  }
};

struct Matrix3f {
  float m[9];
  Matrix3f() { std::memset(m, 0, sizeof(m)); }
  float &operator()(int r, int c) { return m[c * 3 + r]; }
  float operator()(int r, int c) const { return m[c * 3 + r]; }
};

// Matrix6f / Vector6f of the Linearizer's H and b (kept in double: the device
// reduces them in FP64)
struct Matrix6d {
  double m[36];
  Matrix6d() { std::memset(m, 0, sizeof(m)); }
  double &operator()(int r, int c) { return m[c * 6 + r]; }
  double operator()(int r, int c) const { return m[c * 6 + r]; }
};
struct Vector6d {
  double v[6];
  Vector6d() { std::memset(v, 0, sizeof(v)); }
  double &operator()(int i) { return v[i]; }
  double operator()(int i) const { return v[i]; }
};

// cv::Mat_<T> in row-major rows x cols
template <typename T>
struct Image {
  int rows = 0, cols = 0;
  std::vector<T> data;
  Image() {}
  Image(int r, int c) { create(r, c); }
  void create(int r, int c) {
    rows = r;
    cols = c;
    data.assign((size_t)r * c, T());
  }
  T &operator()(int r, int c) { return data[(size_t)r * cols + c]; }
  const T &operator()(int r, int c) const { return data[(size_t)r * cols + c]; }
};
typedef Image<uint16_t> RawDepthImage;  // 16UC1
typedef Image<float> DepthImage;        // 32FC1, metres
typedef Image<int> IntImage;
typedef IntImage IndexImage;

struct Correspondence {
  int referenceIndex = -1, currentIndex = -1;
  Correspondence() {}
  Correspondence(int r, int c) : referenceIndex(r), currentIndex(c) {}
};
typedef std::vector<Correspondence> CorrespondenceVector;

inline void check(int rc) {
  if (rc != NICP_OK) throw std::runtime_error(std::string("nicp_amd: ") + nicp_last_error());
}

// One device context (HIP stream + scratch) per thread, shared by the objects the
// thread creates: the reference's objects are not thread-safe either
// (one pipeline per thread).
class Context {
 public:
  static nicp_ctx *get(int device = 0) {
    static thread_local Context c(device);
    return c.ctx_;
  }
  ~Context() {
    if (ctx_) nicp_ctx_destroy(ctx_);
  }

 private:
  explicit Context(int device) { check(nicp_ctx_create(device, &ctx_)); }
  nicp_ctx *ctx_ = nullptr;
};

}  // namespace nicp
