// nicp_aligner: the reference's incremental odometry with a merged local scene
// (nicp/nicp/nicp_aligner.cpp:28-232) on the nicp_amd C++ front end.
//
//   nicp_aligner configuration.conf depth_list.txt odometry.txt
//
// Every frame is registered against the local scene rendered from the last pose (the
// scene is projected into a depth image and converted again), then added to the scene
// and merged into it (Merger). Every chunkStep frames the scene is saved as
// scene-NNN.nicp and restarted. Each frame's cloud is saved next to its depth image as
// <depth file>.nicp, in scene coordinates with the global transform in its header, as the
// reference does. Configuration keys: those of nicp_simple_aligner plus chunkStep,
// depthThreshold, normalThreshold, distanceThreshold and the initial pose tx ty tz qx qy qz qw.
#include <cstdio>
#include <memory>

#include "app_common.h"
#include "nicp/merger.h"

using namespace nicp;
using namespace nicp_app;

int main(int argc, char **argv) {
  if (argc < 4) {
    std::cout << "USAGE: nicp_aligner configuration.conf depth_list.txt odometry.txt" << std::endl;
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
  const int chunkStep = (int)get("chunkStep", 10.0f);
  const Isometry3f initialT = initialPose(p);

  Objects o;
  setInputParameters(o, p);
  Merger merger;
  if (p.count("depthThreshold")) merger.setMaxPointDepth(p["depthThreshold"]);
  if (p.count("normalThreshold")) merger.setNormalThreshold(p["normalThreshold"]);
  if (p.count("distanceThreshold")) merger.setDistanceThreshold(p["distanceThreshold"]);
  DepthImageConverterIntegralImage converter(&o.projector, &o.statsCalculator, &o.pointInfo, &o.normalInfo);
  converter.setKeepStats(true);  // the saved clouds carry their Stats, as the reference's do
  merger.setDepthImageConverter(&converter);
  PinholePointProjector &projector = o.projector;
  AlignerProjective &aligner = o.aligner;

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
  std::unique_ptr<Cloud> scene(new Cloud()), subscene(new Cloud()), cloud;
  Isometry3f globalT = initialT, sceneT = initialT;
  globalT.setBottomRow();
  sceneT.setBottomRow();
  bool first = true;
  int counter = 0, rows = 0, cols = 0;
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
    DepthImage depth, scaled;
    DepthImage_convert_16UC1_to_32FC1(depth, raw, depthScale);
    DepthImage_scale(scaled, depth, imageScale);
    if (first) {
      projector.setImageSize(raw.rows, raw.cols);
      projector.scale(1.0f / imageScale);
      rows = scaled.rows;
      cols = scaled.cols;
      merger.setImageSize(rows, cols);
    }
    cloud.reset(new Cloud());
    converter.compute(*cloud, scaled);
    std::cout << "Current image " << file << std::endl;

    if (!first) {
      // the scene seen from the last pose becomes the reference cloud
      IntImage index;
      DepthImage sceneDepth;
      projector.setTransform(sceneT);
      projector.project(index, sceneDepth, *scene);
      converter.compute(*subscene, sceneDepth);
      projector.setTransform(Isometry3f::Identity());
      aligner.setReferenceCloud(subscene.get());
      aligner.setCurrentCloud(cloud.get());
      aligner.setInitialGuess(Isometry3f::Identity());
      aligner.align();
      globalT = globalT * aligner.T();
      globalT.setBottomRow();
      sceneT = sceneT * aligner.T();
      sceneT.setBottomRow();
    }

    if (!first && counter++ % chunkStep == 0) {
      char name[64];
      std::snprintf(name, sizeof(name), "scene-%03d.nicp", counter);
      float v[6];
      toVector6(sceneT.inverse() * globalT, v);
      scene->save(name, v, 1, true);
      sceneT = Isometry3f::Identity();
      scene.reset(new Cloud());
    }

    scene->add(*cloud, sceneT);
    merger.merge(scene.get(), sceneT);
    std::cout << "scene points " << scene->size() << std::endl;

    os << odometryLine(timestamp, globalT) << std::endl;
    float v[6];
    toVector6(globalT, v);
    cloud->save((file + ".nicp").c_str(), v, 1, true);
    first = false;
  }
  char name[64];
  std::snprintf(name, sizeof(name), "scene-%03d.nicp", counter);
  float v[6];
  toVector6(sceneT.inverse() * globalT, v);
  scene->save(name, v, 1, true);
  return 0;
}
