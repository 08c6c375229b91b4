// Registers two 640x480 16-bit PGM depth images with the C++ front end, both converted and
// aligned with one sensor offset (DepthImageConverterIntegralImage::compute(cloud, depth,
// offset), Aligner::setSensorOffset), and prints the transform (column-major, 9 significant
// digits) and the inlier count. Arguments: the two images, then the offset's 16 column-major
// floats.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>

#include "nicp/alignerprojective.h"
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
  if (argc < 19) return 2;
  nicp::RawDepthImage d0, d1;
  if (!read_pgm16(argv[1], d0) || !read_pgm16(argv[2], d1)) return 3;
  nicp::Isometry3f offset;
  for (int k = 0; k < 16; ++k) offset.matrix()[k] = std::strtof(argv[3 + k], nullptr);
  nicp::PinholePointProjector projector;  // fx = fy = 525, (319.5, 239.5), 0.5 .. 4.5 m as the conf files
  nicp::Matrix3f K;
  K(0, 0) = 525.0f; K(1, 1) = 525.0f; K(0, 2) = 319.5f; K(1, 2) = 239.5f; K(2, 2) = 1.0f;
  projector.setCameraMatrix(K);
  projector.setMinDistance(0.5f);
  projector.setMaxDistance(4.5f);
  nicp::StatsCalculatorIntegralImage stats;
  stats.setCurvatureThreshold(0.2f);
  nicp::PointInformationMatrixCalculator pim;
  nicp::NormalInformationMatrixCalculator nim;
  nicp::DepthImageConverterIntegralImage conv(&projector, &stats, &pim, &nim);
  nicp::Cloud ref, cur;
  conv.compute(ref, d0, 0.001f, offset);
  conv.compute(cur, d1, 0.001f, offset);
  nicp::CorrespondenceFinderProjective finder;
  finder.setInlierNormalAngularThreshold(0.95f);
  finder.setFlatCurvatureThreshold(0.02f);
  nicp::Linearizer lin;
  nicp::AlignerProjective aligner;
  aligner.setProjector(&projector);
  aligner.setCorrespondenceFinder(&finder);
  aligner.setLinearizer(&lin);
  aligner.setReferenceCloud(&ref);
  aligner.setCurrentCloud(&cur);
  aligner.setInitialGuess(nicp::Isometry3f::Identity());
  aligner.setSensorOffset(offset);
  aligner.align();
  printf("T");
  for (int k = 0; k < 16; ++k) printf(" %.9g", aligner.T().matrix()[k]);
  printf("\ninliers %d points %zu %zu\n", aligner.inliers(), ref.size(), cur.size());
  return 0;
}
