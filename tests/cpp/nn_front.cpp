// Registers two unstructured clouds with the C++ front end's AlignerNN and
// CorrespondenceFinderNN (nicp/nicp/alignernn.h, correspondencefindernn.h). Each input file
// holds an int32 count n and then float32 points (n x 4), normals (n x 4), curvature (n),
// point and normal information matrices (n x 16 each). Prints the correspondences found at
// the truth-halfway transform given in argv[3..18] (column-major), then the aligned transform
// and inlier count.
#include <cstdio>
#include <cstdlib>
#include <fstream>

#include "nicp/alignernn.h"

static bool read_cloud(const char *path, nicp::Cloud &c) {
  std::ifstream is(path, std::ios::binary);
  int n = 0;
  is.read((char *)&n, sizeof(n));
  if (!is || n < 0) return false;
  std::vector<std::array<float, 4>> p(n), nr(n);
  std::vector<float> curv(n);
  std::vector<std::array<float, 16>> op(n), on(n);
  is.read((char *)p.data(), (std::streamsize)n * 16);
  is.read((char *)nr.data(), (std::streamsize)n * 16);
  is.read((char *)curv.data(), (std::streamsize)n * 4);
  is.read((char *)op.data(), (std::streamsize)n * 64);
  is.read((char *)on.data(), (std::streamsize)n * 64);
  if (!is) return false;
  c.setFromHost(p, nr, curv, op, on);
  return true;
}

int main(int argc, char **argv) {
  if (argc < 19) return 2;
  nicp::Cloud ref, cur;
  if (!read_cloud(argv[1], ref) || !read_cloud(argv[2], cur)) return 3;
  nicp::Isometry3f T;
  for (int k = 0; k < 16; ++k) T.matrix()[k] = (float)atof(argv[3 + k]);
  nicp::CorrespondenceFinderNN finder;
  finder.setInlierDistanceThreshold(0.5f);
  finder.init(ref, cur);
  finder.compute(ref, cur, T);
  printf("pairs %d", finder.numCorrespondences());
  unsigned long long h = 1469598103934665603ull;
  for (const nicp::Correspondence &c : finder.correspondences()) {
    h = (h ^ (unsigned)c.referenceIndex) * 1099511628211ull;
    h = (h ^ (unsigned)c.currentIndex) * 1099511628211ull;
  }
  printf(" %llu\n", h);
  finder.setInlierDistanceThreshold(1.0f);
  nicp::Linearizer lin;
  nicp::AlignerNN aligner;
  aligner.setCorrespondenceFinder(&finder);
  aligner.setLinearizer(&lin);
  aligner.setOuterIterations(20);
  aligner.setReferenceCloud(&ref);
  aligner.setCurrentCloud(&cur);
  aligner.setInitialGuess(nicp::Isometry3f::Identity());
  aligner.align();
  printf("T");
  for (int k = 0; k < 16; ++k) printf(" %.9g", aligner.T().matrix()[k]);
  printf("\ninliers %d correspondences %d\n", aligner.inliers(), finder.numCorrespondences());
  return 0;
}
