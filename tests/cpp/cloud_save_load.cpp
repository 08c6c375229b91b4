// Converts a 16-bit PGM depth image with the C++ front end (Stats kept), saves the cloud
// in both .nicp formats, loads both back and prints what the test checks.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>

#include "nicp/depthimageconverterintegralimage.h"

static bool read_pgm16(const char *path, nicp::RawDepthImage &img) {
  std::ifstream is(path, std::ios::binary);
  std::string magic;
  int w, h, maxv;
  is >> magic >> w >> h >> maxv;
  is.get();
  if (magic != "P5" || maxv != 65535) return false;
  img.create(h, w);
  for (int i = 0; i < w * h; ++i) {
    unsigned char b[2];
    is.read((char *)b, 2);
    img.data[i] = (uint16_t)(b[0] << 8 | b[1]);
  }
  return (bool)is;
}

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  nicp::RawDepthImage raw;
  if (!read_pgm16(argv[1], raw)) return 3;
  nicp::PinholePointProjector projector;
  nicp::Matrix3f K;
  K(2, 2) = 1.0f;
  K(0, 0) = 525.0f; K(1, 1) = 525.0f; K(0, 2) = 319.5f; K(1, 2) = 239.5f;
  projector.setCameraMatrix(K);
  projector.setMinDistance(0.5f);
  projector.setMaxDistance(4.5f);
  nicp::StatsCalculatorIntegralImage stats;
  stats.setCurvatureThreshold(0.2f);
  nicp::PointInformationMatrixCalculator pim;
  nicp::NormalInformationMatrixCalculator nim;
  nicp::DepthImageConverterIntegralImage conv(&projector, &stats, &pim, &nim);
  conv.setKeepStats(true);
  nicp::Cloud cloud;
  conv.compute(cloud, raw, 0.001f);
  const float T[6] = {0.5f, -0.25f, 0.125f, 0.1f, 0.0f, -0.05f};
  const std::string dir = argv[2];
  if (!cloud.save((dir + "/b.nicp").c_str(), T, 1, true)) return 4;
  if (!cloud.save((dir + "/t.nicp").c_str(), T, 1, false)) return 5;
  const auto p0 = cloud.points(), n0 = cloud.normals();
  const auto c0 = cloud.curvatures();
  for (const char *name : {"/b.nicp", "/t.nicp"}) {
    nicp::Cloud back;
    float Tb[6];
    if (!back.load(Tb, (dir + name).c_str())) return 6;
    const auto p1 = back.points(), n1 = back.normals();
    const auto c1 = back.curvatures();
    double dp = 0.0, dn = 0.0, dc = 0.0, dt = 0.0;
    for (size_t i = 0; i < p0.size() && i < p1.size(); ++i)
      for (int k = 0; k < 4; ++k) {
        dp = std::fmax(dp, std::fabs(p0[i][k] - p1[i][k]));
        dn = std::fmax(dn, std::fabs(n0[i][k] - n1[i][k]));
      }
    for (size_t i = 0; i < c0.size() && i < c1.size(); ++i) dc = std::fmax(dc, std::fabs(c0[i] - c1[i]));
    for (int k = 0; k < 6; ++k) dt = std::fmax(dt, std::fabs(T[k] - Tb[k]));
    printf("%s n0 %zu n1 %zu dp %.9g dn %.9g dc %.9g dt %.9g\n", name + 1, p0.size(), p1.size(), dp, dn, dc, dt);
  }
  return 0;
}
