// The .nicp cloud file of the reference (nicp/nicp/cloud.cpp:13-150), on the host:
// "NICPCLOUD <n> <binary>" line, a line with the 6-vector t2v(T), then one record per
// point. Text records are "POINTWITHSTATS px py pz nx ny nz" followed by the 16 entries
// of the Stats matrix row by row. Binary records are the raw Point, Normal and Stats
// structs of the reference's x86-64 build, 16 + 16 + 96 bytes: Point/Normal are float4;
// Stats is the column-major Matrix4f [eigenvectors | mean] at 0, int n at 64, the three
// eigenvalues at 68, the curvature-computed flag at 80 and the curvature at 84, zero
// padded to 96. A cloud saved without stats carries the reference's default Stats with a
// zero matrix: n 0, eigenvalues 0, flag false, curvature 1 (nicp/nicp/stats.h:21-27,
// cloud.cpp:134-139). Loading reads points, normals and Stats, as the reference does;
// information matrices are not part of the file.
#pragma once

#include <array>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace nicp {

struct CloudFile {
  std::vector<std::array<float, 4>> points, normals;
  std::vector<std::array<float, 16>> stats;  // column-major 4x4
  std::vector<int> n;
  std::vector<std::array<float, 3>> eigenvalues;
  std::vector<unsigned char> curvature_computed;
  std::vector<float> curvature;
  float t2v[6] = {0, 0, 0, 0, 0, 0};

  size_t size() const { return points.size(); }

  void resize(size_t s) {
    points.assign(s, {0.0f, 0.0f, 0.0f, 1.0f});
    normals.assign(s, {0.0f, 0.0f, 0.0f, 0.0f});
    std::array<float, 16> I{};
    for (int k = 0; k < 4; ++k) I[k * 5] = 1.0f;
    stats.assign(s, I);
    n.assign(s, 0);
    eigenvalues.assign(s, {0.0f, 0.0f, 0.0f});
    curvature_computed.assign(s, 0);
    curvature.assign(s, 1.0f);
  }

  // Cloud::save (cloud.cpp:99-150); has_stats = false writes the reference's zero Stats
  bool write(const char *filename, bool binary = true, int step = 1, bool has_stats = true) const {
    std::ofstream os(filename, std::ios::binary);
    if (!os || step < 1) return false;
    const size_t cnt = size();
    os << "NICPCLOUD " << cnt / step << " " << (binary ? 1 : 0) << "\n";
    for (int k = 0; k < 6; ++k) os << t2v[k] << " ";
    os << "\n";
    for (size_t i = 0; i < cnt; i += step) {
      std::array<float, 16> S{};
      if (has_stats) S = stats[i];
      if (!binary) {
        os << "POINTWITHSTATS ";
        for (int k = 0; k < 3; ++k) os << points[i][k] << " ";
        for (int k = 0; k < 3; ++k) os << normals[i][k] << " ";
        for (int r = 0; r < 4; ++r)
          for (int c = 0; c < 4; ++c) os << S[c * 4 + r] << " ";
        os << "\n";
      } else {
        unsigned char rec[128] = {0};
        std::memcpy(rec, points[i].data(), 16);
        std::memcpy(rec + 16, normals[i].data(), 16);
        std::memcpy(rec + 32, S.data(), 64);
        const int nn = has_stats ? n[i] : 0;
        const float one = 1.0f;
        std::memcpy(rec + 96, &nn, 4);
        if (has_stats) std::memcpy(rec + 100, eigenvalues[i].data(), 12);
        rec[112] = has_stats ? curvature_computed[i] : 0;
        std::memcpy(rec + 116, has_stats ? &curvature[i] : &one, 4);
        os.write((const char *)rec, sizeof(rec));
      }
    }
    return (bool)os;
  }

  // Stats::curvature (stats.h:93-98): the stored value once computed, else
  // lambda0 / (lambda0 + lambda1 + lambda2 + 1e-9) in double
  float curvature_of(size_t i) const {
    if (curvature_computed[i]) return curvature[i];
    const std::array<float, 3> &e = eigenvalues[i];
    return (float)(e[0] / ((double)(e[0] + e[1] + e[2]) + 1e-9));
  }

  // Cloud::load (cloud.cpp:13-97): true when every record was read
  bool read(const char *filename) {
    std::ifstream is(filename, std::ios::binary);
    if (!is) return false;
    std::string line;
    if (!std::getline(is, line)) return false;
    std::istringstream hs(line);
    std::string tag;
    size_t cnt = 0;
    int binary = 0;
    hs >> tag >> cnt >> binary;
    if (tag != "NICPCLOUD" || !hs) return false;
    if (!std::getline(is, line)) return false;
    std::istringstream ts(line);
    for (int k = 0; k < 6; ++k) ts >> t2v[k];
    resize(cnt);
    size_t k = 0;
    while (k < cnt && is.good()) {
      if (!binary) {
        if (!std::getline(is, line)) break;
        std::istringstream ls(line);
        std::string s;
        ls >> s;
        if (s != "POINTWITHSTATS") continue;
        for (int i = 0; i < 3 && ls; ++i) ls >> points[k][i];
        for (int i = 0; i < 3 && ls; ++i) ls >> normals[k][i];
        for (int r = 0; r < 4 && ls; ++r)
          for (int c = 0; c < 4 && ls; ++c) ls >> stats[k][c * 4 + r];
      } else {
        unsigned char rec[128];
        if (!is.read((char *)rec, sizeof(rec))) break;
        std::memcpy(points[k].data(), rec, 16);
        std::memcpy(normals[k].data(), rec + 16, 16);
        std::memcpy(stats[k].data(), rec + 32, 64);
        std::memcpy(&n[k], rec + 96, 4);
        std::memcpy(eigenvalues[k].data(), rec + 100, 12);
        curvature_computed[k] = rec[112];
        std::memcpy(&curvature[k], rec + 116, 4);
      }
      ++k;
    }
    return k == cnt;
  }
};

}  // namespace nicp
