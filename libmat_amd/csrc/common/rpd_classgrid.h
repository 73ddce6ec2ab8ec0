// Weight-class grids of a sphere set (host side), shared by the host and the
// device power-neighbour builders.  Spheres are split into 8 classes by
// weight w = r^2 ([wmax/4^(c+1), wmax/4^c), the last class catches the
// rest); each class gets a uniform grid over its centres' bounding box with
// about 3 members per cell, at most 256 cells along its extent and never
// below 1/1024 of the domain box.  Candidate searches visit the classes one
// by one with each class's own weight bound.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <vector>

struct RpdClassGrid {
  double wmax = 0;
  double org[3] = {0, 0, 0};
  double cs = 1;
  int32_t dim[3] = {1, 1, 1};
  std::vector<int64_t> start;  // per grid cell, into ids
  std::vector<int32_t> ids;    // members in cell order, ascending within
};

constexpr int kRpdClasses = 8;

inline void rpd_build_class_grids(int32_t n, const double* xyzw,
                                  const double* box,
                                  std::vector<RpdClassGrid>& grids) {
  double wmax = 0;
  for (int32_t i = 0; i < n; ++i) wmax = std::max(wmax, xyzw[4 * i + 3]);
  std::vector<std::vector<int32_t>> members(kRpdClasses);
  for (int32_t i = 0; i < n; ++i) {
    double w = xyzw[4 * i + 3];
    int c = 0;
    double lim = wmax * 0.25;
    while (c < kRpdClasses - 1 && !(w >= lim)) {
      ++c;
      lim *= 0.25;
    }
    if (wmax <= 0) c = kRpdClasses - 1;
    members[c].push_back(i);
  }
  grids.assign(kRpdClasses, RpdClassGrid());
  const double box_ext =
      std::max(std::max(box[3] - box[0], box[4] - box[1]), box[5] - box[2]);
  for (int c = 0; c < kRpdClasses; ++c) {
    auto& g = grids[c];
    auto& m = members[c];
    if (m.empty()) continue;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int32_t i : m) {
      g.wmax = std::max(g.wmax, xyzw[4 * i + 3]);
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], xyzw[4 * i + a]);
        hi[a] = std::max(hi[a], xyzw[4 * i + a]);
      }
    }
    double ext = std::max(std::max(hi[0] - lo[0], hi[1] - lo[1]),
                          std::max(hi[2] - lo[2], 1e-12));
    double vol = std::max((hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]),
                          ext * ext * ext * 1e-6);
    g.cs = std::max(std::max(std::cbrt(vol * 3.0 / m.size()), ext / 256.0),
                    box_ext / 1024.0);
    int64_t cells = 1;
    for (int a = 0; a < 3; ++a) {
      g.org[a] = lo[a];
      g.dim[a] = std::max(1, (int32_t)std::ceil((hi[a] - lo[a]) / g.cs) + 1);
      cells *= g.dim[a];
    }
    g.start.assign(cells + 1, 0);
    auto cell_of = [&](int32_t i) {
      int64_t q[3];
      for (int a = 0; a < 3; ++a) {
        q[a] = (int64_t)std::floor((xyzw[4 * i + a] - g.org[a]) / g.cs);
        q[a] = std::min<int64_t>(std::max<int64_t>(q[a], 0), g.dim[a] - 1);
      }
      return q[0] + g.dim[0] * (q[1] + (int64_t)g.dim[1] * q[2]);
    };
    for (int32_t i : m) g.start[cell_of(i) + 1]++;
    for (int64_t q = 0; q < cells; ++q) g.start[q + 1] += g.start[q];
    g.ids.resize(m.size());
    std::vector<int64_t> fill(g.start.begin(), g.start.end() - 1);
    for (int32_t i : m) g.ids[fill[cell_of(i)]++] = i;
  }
}
