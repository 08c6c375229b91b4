// Registers two panoramic 16-bit PGM depth images with the C++ front end and the
// SphericalPointProjector (nicp/nicp/sphericalpointprojector.h) and prints the transform
// (column-major, 9 significant digits) and the inlier count.
#include <cstdio>
#include <fstream>
#include <string>

#include "nicp/alignerprojective.h"
#include "nicp/depthimageconverterintegralimage.h"
#include "nicp/sphericalpointprojector.h"

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
  nicp::RawDepthImage d0, d1;
  if (!read_pgm16(argv[1], d0) || !read_pgm16(argv[2], d1)) return 3;
  nicp::SphericalPointProjector projector;
  projector.setMinDistance(0.5f);
  projector.setMaxDistance(10.0f);
  projector.setHorizontalCenter(d0.cols / 2.0f);
  projector.setVerticalCenter(d0.rows / 2.0f);
  nicp::StatsCalculatorIntegralImage stats;
  stats.setWorldRadius(0.25f);
  stats.setMinImageRadius(10);
  stats.setMaxImageRadius(100);
  stats.setMinPoints(30);
  stats.setCurvatureThreshold(0.075f);
  nicp::PointInformationMatrixCalculator pim;
  nicp::NormalInformationMatrixCalculator nim;
  nicp::DepthImageConverterIntegralImage conv(&projector, &stats, &pim, &nim);
  nicp::Cloud ref, cur;
  conv.compute(ref, d0, 0.001f);
  conv.compute(cur, d1, 0.001f);
  nicp::CorrespondenceFinderProjective finder;  // the values of nicp_aligner_1_1.conf
  finder.setInlierNormalAngularThreshold(0.95f);
  finder.setFlatCurvatureThreshold(0.02f);
  nicp::Linearizer lin;
  nicp::AlignerProjective aligner;
  aligner.setProjector(&projector);
  aligner.setCorrespondenceFinder(&finder);
  aligner.setLinearizer(&lin);
  aligner.setOuterIterations(20);
  aligner.setReferenceCloud(&ref);
  aligner.setCurrentCloud(&cur);
  aligner.setInitialGuess(nicp::Isometry3f::Identity());
  aligner.align();
  printf("T");
  for (int k = 0; k < 16; ++k) printf(" %.9g", aligner.T().matrix()[k]);
  printf("\ninliers %d points %zu %zu\n", aligner.inliers(), ref.size(), cur.size());
  return 0;
}
