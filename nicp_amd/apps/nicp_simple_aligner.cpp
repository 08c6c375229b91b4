// nicp_simple_aligner: the command-line odometry of the reference
// (nicp/nicp/nicp_simple_aligner.cpp:28-268) on the nicp_amd C++ front end.
//
//   nicp_simple_aligner configuration.conf depth_list.txt odometry.txt
//
// configuration: "key value" lines as in nicp/nicp_conf/nicp_aligner_1_1.conf
// (unknown keys and "//" lines are ignored); depth list: "timestamp file.pgm"
// lines naming 16-bit binary PGM depth images; output: one line per frame,
// "timestamp tx ty tz qx qy qz qw" of the accumulated global transform, and each
// frame's cloud saved next to its depth image as <depth file>.nicp with that transform.
#include <cstdio>
#include <memory>

#include "app_common.h"

using namespace nicp;
using namespace nicp_app;

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

  Objects o;
  setInputParameters(o, p);
  PinholePointProjector &projector = o.projector;
  AlignerProjective &aligner = o.aligner;
  DepthImageConverterIntegralImage converter(&projector, &o.statsCalculator, &o.pointInfo, &o.normalInfo);
  converter.setKeepStats(true);  // the saved clouds carry their Stats, as the reference's do

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
  // nicp_simple_aligner.cpp:125-128: the identity sensor offset, the initial pose
  const Isometry3f sensorOffset = Isometry3f::Identity();
  Isometry3f globalT = initialPose(p);
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
    converter.compute(*cloud, scaled, sensorOffset);
    if (previous) {
      aligner.setReferenceCloud(previous.get());
      aligner.setCurrentCloud(cloud.get());
      aligner.setInitialGuess(Isometry3f::Identity());
      aligner.setSensorOffset(sensorOffset);
      aligner.align();
      globalT = globalT * aligner.T();
      globalT.setBottomRow();
    }
    // nicp_simple_aligner.cpp:175: Cloud::save(depth file + ".nicp", globalT, step 1, binary)
    float v[6];
    toVector6(globalT, v);
    if (!cloud->save((file + ".nicp").c_str(), v, 1, true)) {
      std::cerr << "cannot write " << file << ".nicp" << std::endl;
      return 1;
    }
    os << odometryLine(timestamp, globalT) << std::endl;
    std::cout << "frame " << file << " points " << cloud->size() << " inliers " << aligner.inliers() << std::endl;
    previous = std::move(cloud);
  }
  return 0;
}
