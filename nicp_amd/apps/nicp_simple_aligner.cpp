// nicp_simple_aligner: the command-line odometry of the reference
// (nicp/nicp/nicp_simple_aligner.cpp:28-268) on the nicp_amd C++ front end.
//
//   nicp_simple_aligner configuration.conf depth_list.txt odometry.txt
//
// configuration: "key value" lines as in nicp/nicp_conf/nicp_aligner_1_1.conf
// (unknown keys and "//" lines are ignored); depth list: "timestamp file.pgm"
// lines naming 16-bit binary PGM depth images; output: one line per frame,
// "timestamp tx ty tz qx qy qz qw" of the accumulated global transform.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>

#include "nicp/alignerprojective.h"
#include "nicp/depthimageconverterintegralimage.h"

using namespace nicp;

static bool readPGM16(const std::string &path, RawDepthImage &img) {
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

static bool readConfig(const char *path, std::map<std::string, float> &params) {
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
static void toQuaternion(const Isometry3f &T, double q[4]) {
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

int main(int argc, char **argv) {
  if (argc < 4) {
    std::cout << "USAGE: nicp_simple_aligner configuration.conf depth_list.txt odometry.txt" << std::endl;
    return 0;
  }
  std::map<std::string, float> p;
  if (!readConfig(argv[1], p)) {
    std::cerr << "Error while reading input parameters" << std::endl;
    return 1;
  }
  auto get = [&](const char *k, float d) { auto it = p.find(k); return it == p.end() ? d : it->second; };
  const float depthScale = get("depthScale", 0.001f);
  const int imageScale = (int)get("imageScale", 1.0f);

  PinholePointProjector projector;
  StatsCalculatorIntegralImage statsCalculator;
  PointInformationMatrixCalculator pointInfo;
  NormalInformationMatrixCalculator normalInfo;
  CorrespondenceFinderProjective finder;
  Linearizer linearizer;
  AlignerProjective aligner;
  // setInputParameters (nicp/nicp/nicp_simple_aligner.cpp:213-268)
  Matrix3f K;
  K(0, 0) = get("fx", 525.0f); K(1, 1) = get("fy", 525.0f);
  K(0, 2) = get("cx", 319.5f); K(1, 2) = get("cy", 239.5f); K(2, 2) = 1.0f;
  if (p.count("minDistance")) projector.setMinDistance(p["minDistance"]);
  if (p.count("maxDistance")) projector.setMaxDistance(p["maxDistance"]);
  projector.setCameraMatrix(K);
  if (p.count("minImageRadius")) statsCalculator.setMinImageRadius((int)p["minImageRadius"]);
  if (p.count("maxImageRadius")) statsCalculator.setMaxImageRadius((int)p["maxImageRadius"]);
  if (p.count("minPoints")) statsCalculator.setMinPoints((int)p["minPoints"]);
  if (p.count("curvatureThreshold")) statsCalculator.setCurvatureThreshold(p["curvatureThreshold"]);
  if (p.count("worldRadius")) statsCalculator.setWorldRadius(p["worldRadius"]);
  if (p.count("informationMatrixCurvatureThreshold")) {
    pointInfo.setCurvatureThreshold(p["informationMatrixCurvatureThreshold"]);
    normalInfo.setCurvatureThreshold(p["informationMatrixCurvatureThreshold"]);
  }
  if (p.count("inlierDistanceThreshold")) finder.setInlierDistanceThreshold(p["inlierDistanceThreshold"]);
  if (p.count("inlierNormalAngularThreshold")) finder.setInlierNormalAngularThreshold(p["inlierNormalAngularThreshold"]);
  if (p.count("inlierCurvatureRatioThreshold")) finder.setInlierCurvatureRatioThreshold(p["inlierCurvatureRatioThreshold"]);
  if (p.count("flatCurvatureThreshold")) finder.setFlatCurvatureThreshold(p["flatCurvatureThreshold"]);
  if (p.count("inlierMaxChi2")) linearizer.setInlierMaxChi2(p["inlierMaxChi2"]);
  if (p.count("robustKernel")) linearizer.setRobustKernel(p["robustKernel"] != 0.0f);
  if (p.count("outerIterations")) aligner.setOuterIterations((int)p["outerIterations"]);
  if (p.count("innerIterations")) aligner.setInnerIterations((int)p["innerIterations"]);
  if (p.count("minInliers")) aligner.setMinInliers((int)p["minInliers"]);
  if (p.count("translationalMinEigenRatio")) aligner.setTranslationalMinEigenRatio(p["translationalMinEigenRatio"]);
  if (p.count("rotationalMinEigenRatio")) aligner.setRotationalMinEigenRatio(p["rotationalMinEigenRatio"]);
  aligner.setProjector(&projector);
  aligner.setCorrespondenceFinder(&finder);
  aligner.setLinearizer(&linearizer);
  DepthImageConverterIntegralImage converter(&projector, &statsCalculator, &pointInfo, &normalInfo);

  std::ifstream is(argv[2]);
  if (!is) {
    std::cerr << "Impossible to open depth image list file: " << argv[2] << std::endl;
    return 1;
  }
  std::ofstream os(argv[3]);
  if (!os) {
    std::cerr << "Impossible to open visual odometry file: " << argv[3] << std::endl;
    return 1;
  }
  std::unique_ptr<Cloud> previous, cloud;
  Isometry3f globalT;
  std::string line;
  while (std::getline(is, line)) {
    std::istringstream iss(line);
    std::string timestamp, file;
    if (!(iss >> timestamp >> file) || timestamp[0] == '#') continue;
    RawDepthImage raw;
    if (!readPGM16(file, raw)) {
      std::cerr << "cannot read 16-bit PGM " << file << std::endl;
      return 1;
    }
    // nicp_simple_aligner.cpp:138-146: convert, pyramid-scale, scale the projector once
    DepthImage depth, scaled;
    DepthImage_convert_16UC1_to_32FC1(depth, raw, depthScale);
    DepthImage_scale(scaled, depth, imageScale);
    if (!previous) {
      projector.setImageSize(raw.rows, raw.cols);
      projector.scale(1.0f / imageScale);
    }
    cloud.reset(new Cloud());
    converter.compute(*cloud, scaled);
    if (previous) {
      aligner.setReferenceCloud(previous.get());
      aligner.setCurrentCloud(cloud.get());
      aligner.setInitialGuess(Isometry3f::Identity());
      aligner.align();
      globalT = globalT * aligner.T();
      globalT.setBottomRow();
    }
    double q[4];
    toQuaternion(globalT, q);
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s %.9g %.9g %.9g %.9g %.9g %.9g %.9g", timestamp.c_str(), globalT(0, 3),
                  globalT(1, 3), globalT(2, 3), q[0], q[1], q[2], q[3]);
    os << buf << std::endl;
    std::cout << "frame " << file << " points " << cloud->size() << " inliers " << aligner.inliers() << std::endl;
    previous = std::move(cloud);
  }
  return 0;
}
