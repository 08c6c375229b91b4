// Writes a small cloud in both .nicp formats with nicp::CloudFile, reads the files back and
// prints what was read, one value per line, for tests/test_cloudfile.py to check.
#include <cstdio>

#include "nicp/cloudfile.h"

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  nicp::CloudFile f;
  f.resize(5);
  for (int i = 0; i < 5; ++i) {
    for (int k = 0; k < 3; ++k) {
      f.points[i][k] = 0.25f * i + 0.125f * k + 1.0f;
      f.normals[i][k] = (k == 2) ? -1.0f : 0.0f;
      f.eigenvalues[i][k] = 0.001f * (k + 1) * (i + 1);
    }
    for (int k = 0; k < 16; ++k) f.stats[i][k] = 0.5f * k - i;
    f.n[i] = 100 + i;
    f.curvature_computed[i] = (i % 2) ? 1 : 0;
    f.curvature[i] = 0.01f * i;
  }
  const float t[6] = {0.1f, -0.2f, 0.3f, 0.01f, -0.02f, 0.03f};
  for (int k = 0; k < 6; ++k) f.t2v[k] = t[k];
  const std::string b = dir + "/c.bin.nicp", x = dir + "/c.txt.nicp", z = dir + "/c.nostats.nicp",
                    s = dir + "/c.step.nicp";
  if (!f.write(b.c_str(), true) || !f.write(x.c_str(), false) || !f.write(z.c_str(), true, 1, false) ||
      !f.write(s.c_str(), true, 2))
    return 3;
  for (const std::string &name : {b, x, z}) {
    nicp::CloudFile g;
    if (!g.read(name.c_str())) return 4;
    printf("file %s %zu\n", name.c_str(), g.size());
    for (int k = 0; k < 6; ++k) printf("t %.9g\n", g.t2v[k]);
    for (size_t i = 0; i < g.size(); ++i) {
      for (int k = 0; k < 4; ++k) printf("p %.9g\n", g.points[i][k]);
      for (int k = 0; k < 4; ++k) printf("n %.9g\n", g.normals[i][k]);
      for (int k = 0; k < 16; ++k) printf("s %.9g\n", g.stats[i][k]);
      printf("c %.9g\n", g.curvature_of(i));
    }
  }
  return 0;
}
