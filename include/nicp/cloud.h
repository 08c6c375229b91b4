// nicp::Cloud (nicp/nicp/cloud.h:20-193): a device-resident cloud. The
// per-point vectors of the reference (points, normals, stats, information
// matrices) are downloaded on demand.
#pragma once

#include <array>
#include <cstdio>
#include <fstream>
#include <vector>

#include "cloudfile.h"
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

  // Cloud::save (nicp/nicp/cloud.cpp:99-150) in the file format of CloudFile; T is given as
  // its 6-vector (t, qx qy qz). Stats are written when the cloud was converted with
  // keep_stats, otherwise the reference's zero Stats.
  bool save(const char *filename, const float t2v6[6], int step = 1, bool binary = true) const {
    CloudFile f;
    const size_t n = size();
    f.resize(n);
    for (int k = 0; k < 6; ++k) f.t2v[k] = t2v6[k];
    f.points = points();
    f.normals = normals();
    std::vector<float> ev(9 * n), el(3 * n), mn(4 * n);
    std::vector<int> np(n);
    const bool have_stats =
        n && nicp_cloud_download_stats(h_, ev.data(), el.data(), mn.data(), np.data()) == NICP_OK;
    if (have_stats) {
      const std::vector<float> curv = curvatures();
      for (size_t i = 0; i < n; ++i) {
        std::array<float, 16> &S = f.stats[i];
        S.fill(0.0f);
        for (int c = 0; c < 3; ++c)
          for (int r = 0; r < 3; ++r) S[c * 4 + r] = ev[9 * i + c * 3 + r];
        for (int r = 0; r < 4; ++r) S[12 + r] = mn[4 * i + r];
        f.n[i] = np[i];
        for (int k = 0; k < 3; ++k) f.eigenvalues[i][k] = el[3 * i + k];
        f.curvature_computed[i] = 1;
        f.curvature[i] = curv[i];
      }
    }
    return f.write(filename, binary, step, have_stats);
  }

  // Cloud::load (nicp/nicp/cloud.cpp:13-97): points, normals and the curvature of the Stats
  // records into a device cloud (no information matrices: the file holds none); t2v6
  // receives the stored transform. Returns false on a malformed or short file.
  bool load(float t2v6[6], const char *filename) {
    CloudFile f;
    if (!f.read(filename)) return false;
    std::vector<float> curv(f.size());
    for (size_t i = 0; i < f.size(); ++i) curv[i] = f.curvature_of(i);
    setFromHost(f.points, f.normals, curv, {}, {});
    for (int k = 0; k < 6; ++k) t2v6[k] = f.t2v[k];
    return true;
  }

  std::vector<std::array<float, 16>> normalInformationMatrix() const { return getOmega(1); }

  // Cloud::gaussians(): the Gaussian3f records with their lazy-evaluation state
  std::vector<nicp_gaussian> gaussians() const {
    std::vector<nicp_gaussian> g(size());
    if (!g.empty()) check(nicp_cloud_download_gaussians(h_, g.data()));
    return g;
  }
  // Cloud::transformInPlace (nicp/nicp/cloud.cpp:193-205)
  void transformInPlace(const Isometry3f &T) { check(nicp_cloud_transform(h_, T.matrix())); }
  // Cloud::add (nicp/nicp/cloud.cpp:163-190): transforms `other` in place by T and appends it.
  // An empty Cloud takes over a copy of `other`'s points.
  void add(Cloud &other, const Isometry3f &T = Isometry3f::Identity()) {
    if (!h_) {
      nicp_cloud *h = nullptr;
      check(nicp_cloud_from_host(Context::get(), 0, nullptr, nullptr, nullptr, nullptr, nullptr, &h));
      h_ = h;
    }
    check(nicp_cloud_add(h_, other.handle(), T.matrix()));
  }

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
