// nicp::Cloud (nicp/nicp/cloud.h:20-193): a device-resident cloud. The
// per-point vectors of the reference (points, normals, stats, information
// matrices) are downloaded on demand.
#pragma once

#include <array>
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
