// nicp::Cloud (nicp/nicp/cloud.h:20-193): a device-resident cloud. The
// per-point vectors of the reference (points, normals, stats, information
// matrices) are downloaded on demand.
#pragma once

#include <array>
#include <cstdio>
#include <fstream>
#include <vector>

#include "definitions.h"

namespace nicp {

class Cloud {
 public:
  Cloud() {}
  ~Cloud() { clear(); }
  Cloud(const Cloud &) = delete;
  Cloud &operator=(const Cloud &) = delete;

  void clear() {
    if (h_) nicp_cloud_destroy(h_);
    h_ = nullptr;
  }
  // takes ownership of a C handle
  void reset(nicp_cloud *h) {
    clear();
    h_ = h;
  }
  nicp_cloud *handle() const { return h_; }
  size_t size() const {
    int n = 0;
    if (h_) check(nicp_cloud_size(h_, &n));
    return (size_t)n;
  }

  // Cloud from host vectors (the reference fills _points/_normals/... directly)
  void setFromHost(const std::vector<std::array<float, 4>> &points, const std::vector<std::array<float, 4>> &normals,
                   const std::vector<float> &curvature, const std::vector<std::array<float, 16>> &pointOmegas,
                   const std::vector<std::array<float, 16>> &normalOmegas) {
    nicp_cloud *h = nullptr;
    const int n = (int)points.size();
    check(nicp_cloud_from_host(Context::get(), n, n ? points[0].data() : nullptr,
                               normals.size() == points.size() && n ? normals[0].data() : nullptr,
                               curvature.size() == points.size() && n ? curvature.data() : nullptr,
                               pointOmegas.size() == points.size() && n ? pointOmegas[0].data() : nullptr,
                               normalOmegas.size() == points.size() && n ? normalOmegas[0].data() : nullptr, &h));
    reset(h);
  }

  std::vector<std::array<float, 4>> points() const { return get(0); }
  std::vector<std::array<float, 4>> normals() const { return get(1); }
  std::vector<float> curvatures() const {
    std::vector<float> c(size());
    if (!c.empty()) check(nicp_cloud_download(h_, nullptr, nullptr, c.data(), nullptr, nullptr));
    return c;
  }
  std::vector<std::array<float, 16>> pointInformationMatrix() const { return getOmega(0); }

  // Cloud::save (nicp/nicp/cloud.cpp:99-150): "NICPCLOUD n binary", the t2v of T, then per
  // point Point, Normal and Stats. Binary records are 16 + 16 + 96 bytes: the Stats layout of
  // the reference's x86-64 build (Matrix4f eigenvectors|mean at 0, int n at 64, eigenvalues at
  // 68, curvature-computed flag at 80, curvature at 84, zero padding to 96). Stats are only
  // available for clouds converted with keep_stats; otherwise they are written as zeros, as
  // the reference does when its stats vector is empty. T is given as its 6-vector (t, qx qy qz).
  bool save(const char *filename, const float t2v6[6], int step = 1, bool binary = true) const {
    std::ofstream os(filename, std::ios::binary);
    if (!os) return false;
    const size_t n = size();
    std::vector<std::array<float, 4>> pts = points(), nrm = normals();
    std::vector<float> ev(9 * n), el(3 * n), mn(4 * n);
    std::vector<int> np(n);
    const bool have_stats = n && nicp_cloud_download_stats(h_, ev.data(), el.data(), mn.data(), np.data()) == NICP_OK;
    const std::vector<float> curv = curvatures();
    os << "NICPCLOUD " << n / step << " " << (binary ? 1 : 0) << "\n";
    os << t2v6[0] << " " << t2v6[1] << " " << t2v6[2] << " " << t2v6[3] << " " << t2v6[4] << " " << t2v6[5] << " \n";
    for (size_t i = 0; i < n; i += step) {
      float S[16] = {0};
      if (have_stats) {
        for (int c = 0; c < 3; ++c)
          for (int r = 0; r < 3; ++r) S[c * 4 + r] = ev[9 * i + c * 3 + r];
        for (int r = 0; r < 4; ++r) S[12 + r] = mn[4 * i + r];
      }
      if (!binary) {
        os << "POINTWITHSTATS ";
        for (int k = 0; k < 3; ++k) os << pts[i][k] << " ";
        for (int k = 0; k < 3; ++k) os << nrm[i][k] << " ";
        for (int r = 0; r < 4; ++r)
          for (int c = 0; c < 4; ++c) os << S[c * 4 + r] << " ";
        os << "\n";
      } else {
        unsigned char rec[128] = {0};
        std::memcpy(rec, pts[i].data(), 16);
        std::memcpy(rec + 16, nrm[i].data(), 16);
        std::memcpy(rec + 32, S, 64);
        if (have_stats) {
          std::memcpy(rec + 96, &np[i], 4);
          std::memcpy(rec + 100, &el[3 * i], 12);
          rec[112] = 1;
          std::memcpy(rec + 116, &curv[i], 4);
        }
        os.write((const char *)rec, sizeof(rec));
      }
    }
    return (bool)os;
  }
  std::vector<std::array<float, 16>> normalInformationMatrix() const { return getOmega(1); }

 private:
  std::vector<std::array<float, 4>> get(int which) const {
    std::vector<std::array<float, 4>> v(size());
    if (!v.empty())
      check(nicp_cloud_download(h_, which == 0 ? v[0].data() : nullptr, which == 1 ? v[0].data() : nullptr, nullptr,
                                nullptr, nullptr));
    return v;
  }
  std::vector<std::array<float, 16>> getOmega(int which) const {
    std::vector<std::array<float, 16>> v(size());
    if (!v.empty())
      check(nicp_cloud_download(h_, nullptr, nullptr, nullptr, which == 0 ? v[0].data() : nullptr,
                                which == 1 ? v[0].data() : nullptr));
    return v;
  }
  nicp_cloud *h_ = nullptr;
};

}  // namespace nicp
