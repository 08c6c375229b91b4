// Helpers shared by the command-line apps: 16-bit PGM depth images, "key value"
// configuration files, the quaternion of a transform, and the parameter setup of
// setInputParameters (nicp/nicp/nicp_simple_aligner.cpp:213-268, nicp/nicp/nicp_aligner.cpp:282-340).
#pragma once

#include <cmath>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>

#include "nicp/alignerprojective.h"
#include "nicp/depthimageconverterintegralimage.h"

namespace nicp_app {

using namespace nicp;

inline bool readPGM16(const std::string &path, RawDepthImage &img) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::string magic;
  f >> magic;
  if (magic != "P5") return false;
  int w = 0, h = 0, maxv = 0;
  auto skip = [&]() {
    f >> std::ws;
    while (f.peek() == '#') {
      std::string line;
      std::getline(f, line);
      f >> std::ws;
    }
  };
  skip(); f >> w; skip(); f >> h; skip(); f >> maxv;
  f.get();
  if (w <= 0 || h <= 0 || maxv <= 255) return false;
  img.create(h, w);
  std::vector<unsigned char> buf((size_t)w * h * 2);
  f.read((char *)buf.data(), buf.size());
  if (!f) return false;
  for (size_t i = 0; i < (size_t)w * h; ++i) img.data[i] = (uint16_t)((buf[2 * i] << 8) | buf[2 * i + 1]);
  return true;
}

inline bool readConfig(const char *path, std::map<std::string, float> &params) {
  std::ifstream is(path);
  if (!is) return false;
  std::string line;
  while (std::getline(is, line)) {
    std::istringstream iss(line);
    std::string key;
    float v;
    if (iss >> key >> v) params[key] = v;
  }
  return true;
}

// Quaternion of a rotation matrix (Shepperd), normalised, as written by the reference
inline void toQuaternion(const Isometry3f &T, double q[4]) {
  double m[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) m[r][c] = T(r, c);
  double t = m[0][0] + m[1][1] + m[2][2];
  if (t > 0) {
    double s = std::sqrt(t + 1.0);
    q[3] = 0.5 * s;
    s = 0.5 / s;
    q[0] = (m[2][1] - m[1][2]) * s;
    q[1] = (m[0][2] - m[2][0]) * s;
    q[2] = (m[1][0] - m[0][1]) * s;
  } else {
    int i = 0;
    if (m[1][1] > m[0][0]) i = 1;
    if (m[2][2] > m[i][i]) i = 2;
    int j = (i + 1) % 3, k = (j + 1) % 3;
    double s = std::sqrt(m[i][i] - m[j][j] - m[k][k] + 1.0);
    q[i] = 0.5 * s;
    s = 0.5 / s;
    q[3] = (m[k][j] - m[j][k]) * s;
    q[j] = (m[j][i] + m[i][j]) * s;
    q[k] = (m[k][i] + m[i][k]) * s;
  }
  double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  for (int i = 0; i < 4; ++i) q[i] /= n;
}

// t2v of a transform (nicp/nicp/bm_se3.h:37-51): translation and the quaternion's vector
// part with w >= 0
inline void toVector6(const Isometry3f &T, float v[6]) {
  double q[4];
  toQuaternion(T, q);
  const double s = q[3] < 0 ? -1.0 : 1.0;
  for (int i = 0; i < 3; ++i) {
    v[i] = T(i, 3);
    v[3 + i] = (float)(s * q[i]);
  }
}

struct Objects {
  PinholePointProjector projector;
  StatsCalculatorIntegralImage statsCalculator;
  PointInformationMatrixCalculator pointInfo;
  NormalInformationMatrixCalculator normalInfo;
  CorrespondenceFinderProjective finder;
  Linearizer linearizer;
  AlignerProjective aligner;
};

// setInputParameters: the configuration keys the reference's apps read
inline void setInputParameters(Objects &o, std::map<std::string, float> &p) {
  auto get = [&](const char *k, float d) { auto it = p.find(k); return it == p.end() ? d : it->second; };
  Matrix3f K;
  K(0, 0) = get("fx", 525.0f); K(1, 1) = get("fy", 525.0f);
  K(0, 2) = get("cx", 319.5f); K(1, 2) = get("cy", 239.5f); K(2, 2) = 1.0f;
  if (p.count("minDistance")) o.projector.setMinDistance(p["minDistance"]);
  if (p.count("maxDistance")) o.projector.setMaxDistance(p["maxDistance"]);
  o.projector.setCameraMatrix(K);
  if (p.count("minImageRadius")) o.statsCalculator.setMinImageRadius((int)p["minImageRadius"]);
  if (p.count("maxImageRadius")) o.statsCalculator.setMaxImageRadius((int)p["maxImageRadius"]);
  if (p.count("minPoints")) o.statsCalculator.setMinPoints((int)p["minPoints"]);
  if (p.count("curvatureThreshold")) o.statsCalculator.setCurvatureThreshold(p["curvatureThreshold"]);
  if (p.count("worldRadius")) o.statsCalculator.setWorldRadius(p["worldRadius"]);
  if (p.count("informationMatrixCurvatureThreshold")) {
    o.pointInfo.setCurvatureThreshold(p["informationMatrixCurvatureThreshold"]);
    o.normalInfo.setCurvatureThreshold(p["informationMatrixCurvatureThreshold"]);
  }
  if (p.count("inlierDistanceThreshold")) o.finder.setInlierDistanceThreshold(p["inlierDistanceThreshold"]);
  if (p.count("inlierNormalAngularThreshold")) o.finder.setInlierNormalAngularThreshold(p["inlierNormalAngularThreshold"]);
  if (p.count("inlierCurvatureRatioThreshold")) o.finder.setInlierCurvatureRatioThreshold(p["inlierCurvatureRatioThreshold"]);
  if (p.count("flatCurvatureThreshold")) o.finder.setFlatCurvatureThreshold(p["flatCurvatureThreshold"]);
  if (p.count("inlierMaxChi2")) o.linearizer.setInlierMaxChi2(p["inlierMaxChi2"]);
  if (p.count("robustKernel")) o.linearizer.setRobustKernel(p["robustKernel"] != 0.0f);
  if (p.count("outerIterations")) o.aligner.setOuterIterations((int)p["outerIterations"]);
  if (p.count("innerIterations")) o.aligner.setInnerIterations((int)p["innerIterations"]);
  if (p.count("minInliers")) o.aligner.setMinInliers((int)p["minInliers"]);
  if (p.count("translationalMinEigenRatio")) o.aligner.setTranslationalMinEigenRatio(p["translationalMinEigenRatio"]);
  if (p.count("rotationalMinEigenRatio")) o.aligner.setRotationalMinEigenRatio(p["rotationalMinEigenRatio"]);
  o.aligner.setProjector(&o.projector);
  o.aligner.setCorrespondenceFinder(&o.finder);
  o.aligner.setLinearizer(&o.linearizer);
}

// the initial pose of both apps (nicp_simple_aligner.cpp:57-70, nicp_aligner.cpp:55-67):
// translation tx ty tz and quaternion qx qy qz qw (w defaults to 1)
inline Isometry3f initialPose(std::map<std::string, float> &p) {
  auto get = [&](const char *k, float d) { auto it = p.find(k); return it == p.end() ? d : it->second; };
  const double qx = get("qx", 0.0f), qy = get("qy", 0.0f), qz = get("qz", 0.0f), qw = get("qw", 1.0f);
  const double R[3][3] = {{1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)},
                          {2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)},
                          {2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)}};
  Isometry3f T;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) T(r, c) = (float)R[r][c];
  T(0, 3) = get("tx", 0.0f);
  T(1, 3) = get("ty", 0.0f);
  T(2, 3) = get("tz", 0.0f);
  T.setBottomRow();
  return T;
}

inline std::string odometryLine(const std::string &timestamp, const Isometry3f &T) {
  double q[4];
  toQuaternion(T, q);
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%s %.9g %.9g %.9g %.9g %.9g %.9g %.9g", timestamp.c_str(), T(0, 3), T(1, 3),
                T(2, 3), q[0], q[1], q[2], q[3]);
  return buf;
}

}  // namespace nicp_app
