// Host-side input preparation for the restricted power diagram: synthetic
// meshes and spheres, sparse tet adjacency, power-diagram neighbour lists.
// See include/rpd_prep.h for the contract of each entry point.
#include "rpd_prep.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../common/rpd_classgrid.h"
#include "../common/rpd_predicates.h"

namespace {

// ------------------------------------------------------------------ RNG
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
inline uint64_t rng_u64(uint64_t seed, uint64_t stream, uint64_t index) {
  uint64_t h = mix64(seed);
  h = mix64(h ^ (stream * 0xD1B54A32D192ED03ULL));
  return mix64(h ^ index);
}
inline double rng_uniform(uint64_t seed, uint64_t stream, uint64_t index) {
  return (double)(rng_u64(seed, stream, index) >> 11) * 0x1.0p-53;
}
// standard normal by Box-Muller from draws (2*index, 2*index+1)
inline double rng_normal(uint64_t seed, uint64_t stream, uint64_t index) {
  double u1 = 1.0 - rng_uniform(seed, stream, 2 * index);  // (0, 1]
  double u2 = rng_uniform(seed, stream, 2 * index + 1);
  return std::sqrt(-2.0 * std::log(u1)) *
         std::cos(6.283185307179586476925 * u2);
}

enum : uint64_t {
  kStreamJitter = 11,
  kStreamCentre = 21,
  kStreamRadius = 31,
  kStreamGmmMean = 41,
  kStreamGmmSigma = 42,
  kStreamGmmPick = 43,
  kStreamGmmDraw = 44,
};

inline double orient(const double* a, const double* b, const double* c,
                     const double* d) {
  double ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
  double vx = c[0] - a[0], vy = c[1] - a[1], vz = c[2] - a[2];
  double wx = d[0] - a[0], wy = d[1] - a[1], wz = d[2] - a[2];
  return ux * (vy * wz - vz * wy) - uy * (vx * wz - vz * wx) +
         uz * (vx * wy - vy * wx);
}

}  // namespace

extern "C" double rpd_rng_uniform(uint64_t seed, uint64_t stream,
                                  uint64_t index) {
  return rng_uniform(seed, stream, index);
}

// ------------------------------------------------------------ grid mesh
extern "C" int rpd_synth_grid_mesh(int32_t n, double jitter_rel,
                                   uint64_t seed, double* verts,
                                   int32_t* tets) {
  if (n <= 0 || !verts || !tets) return -1;
  const int64_t nv1 = n + 1;
  const int64_t n_vert = nv1 * nv1 * nv1;
  const double h = 1.0 / n;
  auto vid = [&](int64_t i, int64_t j, int64_t k) {
    return (int32_t)(i + nv1 * (j + nv1 * k));
  };
  std::vector<int32_t> attempt(n_vert, 0);
  auto place = [&](int64_t v) {
    int64_t i = v % nv1, j = (v / nv1) % nv1, k = v / (nv1 * nv1);
    double p[3] = {i * h, j * h, k * h};
    bool interior = i > 0 && i < n && j > 0 && j < n && k > 0 && k < n;
    if (interior && jitter_rel > 0) {
      for (int a = 0; a < 3; ++a) {
        uint64_t idx = ((uint64_t)v * 3 + a) + (uint64_t)attempt[v] * (1ULL << 40);
        p[a] += jitter_rel * h * rng_normal(seed, kStreamJitter, idx);
      }
    }
    for (int a = 0; a < 3; ++a) verts[3 * v + a] = p[a];
  };
  for (int64_t v = 0; v < n_vert; ++v) place(v);

  // Kuhn split: paths 000 -> e_a -> e_a+e_b -> 111 over the 6 axis orders.
  static const int perm[6][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1},
                                 {0, 2, 1}, {1, 0, 2}, {2, 1, 0}};
  int64_t t = 0;
  for (int64_t k = 0; k < n; ++k)
    for (int64_t j = 0; j < n; ++j)
      for (int64_t i = 0; i < n; ++i)
        for (int p = 0; p < 6; ++p) {
          int64_t c[3] = {i, j, k};
          int32_t v0 = vid(c[0], c[1], c[2]);
          c[perm[p][0]] += 1;
          int32_t v1 = vid(c[0], c[1], c[2]);
          c[perm[p][1]] += 1;
          int32_t v2 = vid(c[0], c[1], c[2]);
          int32_t v3 = vid(i + 1, j + 1, k + 1);
          if (p < 3) {  // even permutation: positive as listed
            tets[4 * t + 0] = v0;
            tets[4 * t + 1] = v1;
            tets[4 * t + 2] = v2;
            tets[4 * t + 3] = v3;
          } else {  // odd permutation: swap v1, v2
            tets[4 * t + 0] = v0;
            tets[4 * t + 1] = v2;
            tets[4 * t + 2] = v1;
            tets[4 * t + 3] = v3;
          }
          ++t;
        }
  const int64_t n_tet = t;
  if (jitter_rel <= 0) return 0;
  // redraw jitter of vertices of tets that lost orientation (or became
  // slivers below 1% of the undistorted orientation determinant h^3)
  const double min_det = 0.01 * h * h * h;
  for (int round = 0; round < 200; ++round) {
    std::vector<int64_t> bad;
    for (int64_t q = 0; q < n_tet; ++q) {
      const int32_t* tv = tets + 4 * q;
      if (orient(verts + 3 * tv[0], verts + 3 * tv[1], verts + 3 * tv[2],
                 verts + 3 * tv[3]) <= min_det)
        for (int a = 0; a < 4; ++a) bad.push_back(tv[a]);
    }
    if (bad.empty()) return 0;
    std::sort(bad.begin(), bad.end());
    bad.erase(std::unique(bad.begin(), bad.end()), bad.end());
    for (int64_t v : bad) {
      attempt[v] += 1;
      place(v);
    }
  }
  return -3;
}

// -------------------------------------------------------------- spheres
extern "C" int rpd_synth_spheres(int32_t n, int32_t rule, double rmin,
                                 double rmax, uint64_t seed, double* xyzr) {
  if (n < 0 || !xyzr) return -1;
  auto bdist = [](double x, double y, double z) {
    double d = std::min(std::min(x, 1.0 - x), std::min(y, 1.0 - y));
    return std::min(d, std::min(z, 1.0 - z));
  };
  if (rule == 0 || rule == 1) {
    for (int64_t i = 0; i < n; ++i) {
      double x = rng_uniform(seed, kStreamCentre, 3 * i + 0);
      double y = rng_uniform(seed, kStreamCentre, 3 * i + 1);
      double z = rng_uniform(seed, kStreamCentre, 3 * i + 2);
      double u = rng_uniform(seed, kStreamRadius, i);
      double r = rule == 0 ? rmin + (rmax - rmin) * u
                           : 0.9 * bdist(x, y, z) * (0.5 + 0.5 * u);
      xyzr[4 * i + 0] = x;
      xyzr[4 * i + 1] = y;
      xyzr[4 * i + 2] = z;
      xyzr[4 * i + 3] = r;
    }
    return 0;
  }
  if (rule == 2) {
    const int K = 32;
    double mu[K][3], sg[K];
    for (int k = 0; k < K; ++k) {
      for (int a = 0; a < 3; ++a)
        mu[k][a] = 0.1 + 0.8 * rng_uniform(seed, kStreamGmmMean, 3 * k + a);
      sg[k] = 0.02 + 0.08 * rng_uniform(seed, kStreamGmmSigma, k);
    }
    for (int64_t i = 0; i < n; ++i) {
      int k = (int)(rng_uniform(seed, kStreamGmmPick, i) * K);
      if (k >= K) k = K - 1;
      double p[3];
      for (uint64_t att = 0;; ++att) {
        for (int a = 0; a < 3; ++a)
          p[a] = mu[k][a] +
                 sg[k] * rng_normal(seed, kStreamGmmDraw,
                                    (uint64_t)i * 3 + a + att * (1ULL << 40));
        if (p[0] > 0 && p[0] < 1 && p[1] > 0 && p[1] < 1 && p[2] > 0 &&
            p[2] < 1)
          break;
        if (att > 1000) return -3;
      }
      double r = std::exp(-3.5 + 0.6 * rng_normal(seed, kStreamRadius, i));
      r = std::min(r, bdist(p[0], p[1], p[2]));
      xyzr[4 * i + 0] = p[0];
      xyzr[4 * i + 1] = p[1];
      xyzr[4 * i + 2] = p[2];
      xyzr[4 * i + 3] = r;
    }
    return 0;
  }
  return -1;
}

// ------------------------------------------------------- tet adjacency
extern "C" int rpd_tet_adjacency(int32_t n_vert, int32_t n_tet,
                                 const int32_t* tets, int32_t* v_adj,
                                 int32_t* e_adj, int32_t* f_adj,
                                 int32_t* f_id, int32_t* n_sf_facets) {
  if (n_vert <= 0 || n_tet <= 0 || !tets) return -1;
  for (int64_t i = 0; i < 4LL * n_tet; ++i)
    if (tets[i] < 0 || tets[i] >= n_vert) return -1;
  std::fill(v_adj, v_adj + n_vert, 0);
  for (int64_t i = 0; i < 4LL * n_tet; ++i) v_adj[tets[i]]++;

  // edges: sort (key, slot) and count runs
  static const int el[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
  {
    std::vector<std::pair<uint64_t, int64_t>> ek(6LL * n_tet);
    for (int64_t t = 0; t < n_tet; ++t)
      for (int k = 0; k < 6; ++k) {
        uint64_t a = (uint32_t)tets[4 * t + el[k][0]];
        uint64_t b = (uint32_t)tets[4 * t + el[k][1]];
        if (a > b) std::swap(a, b);
        ek[6 * t + k] = {(a << 32) | b, 6 * t + k};
      }
    std::sort(ek.begin(), ek.end());
    for (size_t s = 0; s < ek.size();) {
      size_t e = s;
      while (e < ek.size() && ek[e].first == ek[s].first) ++e;
      for (size_t q = s; q < e; ++q) e_adj[ek[q].second] = (int32_t)(e - s);
      s = e;
    }
  }
  // faces: face i opposite vertex i, vertices tet_faces_lvid[i]
  static const int fl[4][3] = {{2, 1, 3}, {0, 2, 3}, {1, 0, 3}, {0, 1, 2}};
  struct FK {
    int32_t a, b, c;
    int64_t slot;
  };
  std::vector<FK> fk(4LL * n_tet);
  for (int64_t t = 0; t < n_tet; ++t)
    for (int i = 0; i < 4; ++i) {
      int32_t v[3] = {tets[4 * t + fl[i][0]], tets[4 * t + fl[i][1]],
                      tets[4 * t + fl[i][2]]};
      std::sort(v, v + 3);
      fk[4 * t + i] = {v[0], v[1], v[2], 4 * t + i};
    }
  std::sort(fk.begin(), fk.end(), [](const FK& p, const FK& q) {
    if (p.a != q.a) return p.a < q.a;
    if (p.b != q.b) return p.b < q.b;
    if (p.c != q.c) return p.c < q.c;
    return p.slot < q.slot;
  });
  std::vector<int64_t> partner(4LL * n_tet, -1);
  for (size_t s = 0; s < fk.size();) {
    size_t e = s;
    while (e < fk.size() && fk[e].a == fk[s].a && fk[e].b == fk[s].b &&
           fk[e].c == fk[s].c)
      ++e;
    if (e - s > 2) return -2;
    if (e - s == 2) {
      partner[fk[s].slot] = fk[s + 1].slot;
      partner[fk[s + 1].slot] = fk[s].slot;
    }
    s = e;
  }
  int32_t n_sf = 0;
  for (int64_t q = 0; q < 4LL * n_tet; ++q) n_sf += partner[q] < 0;
  int32_t next_sf = 0, next_in = n_sf;
  for (int64_t q = 0; q < 4LL * n_tet; ++q) {
    if (partner[q] < 0) {
      f_adj[q] = 1;
      f_id[q] = next_sf++;
    } else {
      f_adj[q] = 2;
      if (partner[q] < q)
        f_id[q] = f_id[partner[q]];
      else
        f_id[q] = next_in++;
    }
  }
  if (n_sf_facets) *n_sf_facets = n_sf;
  return 0;
}

// ------------------------------------------------- power-cell neighbours
namespace {

struct PolyVertex {
  int32_t p[3];  // plane indices
  rpd_d4 X, Q;
};

// Convex polytope clipped by halfspaces {x : e(x) >= 0}, vertices as plane
// triples with cached homogeneous coordinates; exact sign decisions.
struct Poly {
  std::vector<rpd_d4> planes;
  std::vector<int32_t> src;  // sphere id, or -1..-6 for box faces
  std::vector<PolyVertex> verts;
  int32_t failures = 0;

  int side(const PolyVertex& v, const rpd_d4& e) const {
    // sign of e(position) = sign(e.X) * sign(X.w)
    int s = rpd_filtered_sign4(e, v.X, v.Q);
    if (s == 0)
      s = rpd_exact_det4_sign(planes[v.p[0]], planes[v.p[1]], planes[v.p[2]],
                              e);
    int w = rpd_filtered_sign_w(v.X, v.Q);
    if (w == 0)
      w = rpd_exact_w_sign(planes[v.p[0]], planes[v.p[1]], planes[v.p[2]]);
    return s * w;
  }

  void make_vertex(int32_t a, int32_t b, int32_t c, PolyVertex& v) const {
    v.p[0] = a;
    v.p[1] = b;
    v.p[2] = c;
    rpd_vertex_homog(planes[a], planes[b], planes[c], v.X, v.Q);
  }

  void init_box(const double* box) {
    planes.clear();
    src.clear();
    verts.clear();
    // x >= x0, x <= x1, y >= y0, y <= y1, z >= z0, z <= z1
    planes.push_back({1, 0, 0, -box[0]});
    planes.push_back({-1, 0, 0, box[3]});
    planes.push_back({0, 1, 0, -box[1]});
    planes.push_back({0, -1, 0, box[4]});
    planes.push_back({0, 0, 1, -box[2]});
    planes.push_back({0, 0, -1, box[5]});
    for (int k = 0; k < 6; ++k) src.push_back(-1 - k);
    for (int a = 0; a < 2; ++a)
      for (int b = 2; b < 4; ++b)
        for (int c = 4; c < 6; ++c) {
          PolyVertex v;
          make_vertex(a, b, c, v);
          verts.push_back(v);
        }
  }

  // returns true if the plane cut the polytope
  bool clip(const rpd_d4& e, int32_t source) {
    const int nv = (int)verts.size();
    std::vector<char> out(nv, 0);
    int n_out = 0;
    for (int i = 0; i < nv; ++i) {
      out[i] = side(verts[i], e) < 0;
      n_out += out[i];
    }
    if (n_out == 0) return false;
    if (n_out == nv) {
      verts.clear();
      return true;
    }
    const int32_t np = (int32_t)planes.size();
    planes.push_back(e);
    src.push_back(source);
    std::vector<PolyVertex> nvts;
    nvts.reserve(nv + 2 * n_out);
    for (int i = 0; i < nv; ++i)
      if (!out[i]) nvts.push_back(verts[i]);
    for (int i = 0; i < nv; ++i) {
      if (!out[i]) continue;
      const PolyVertex& v = verts[i];
      for (int k = 0; k < 3; ++k) {
        int32_t a = v.p[k], b = v.p[(k + 1) % 3];
        int other = -1;
        for (int j = 0; j < nv; ++j) {
          if (j == i) continue;
          const int32_t* q = verts[j].p;
          bool ha = q[0] == a || q[1] == a || q[2] == a;
          bool hb = q[0] == b || q[1] == b || q[2] == b;
          if (ha && hb) {
            other = j;
            break;
          }
        }
        if (other < 0) {
          ++failures;
          continue;
        }
        if (out[other]) continue;
        PolyVertex nvx;
        make_vertex(a, b, np, nvx);
        nvts.push_back(nvx);
      }
    }
    verts.swap(nvts);
    return true;
  }

  double radius2(double cx, double cy, double cz) const {
    double r2 = 0;
    for (const auto& v : verts) {
      double x = v.X.x / v.X.w - cx, y = v.X.y / v.X.w - cy,
             z = v.X.z / v.X.w - cz;
      r2 = std::max(r2, x * x + y * y + z * z);
    }
    return r2;
  }
};

}  // namespace

struct rpd_nbr {
  std::vector<int32_t> valid;
  std::vector<int64_t> offsets;
  std::vector<int32_t> ids;
  int32_t max_degree = 0;
};

namespace {
// CSR of the symmetric closure over valid spheres: j is a neighbour of i
// iff either list names the other and both are valid; ascending ids
void symmetrize(int32_t n, const std::vector<std::vector<int32_t>>& nb,
                rpd_nbr* res) {
  std::vector<std::vector<int32_t>> sym(n);
  for (int32_t i = 0; i < n; ++i) {
    if (!res->valid[i]) continue;
    for (int32_t j : nb[i]) {
      if (!res->valid[j]) continue;
      sym[i].push_back(j);
      sym[j].push_back(i);
    }
  }
  res->offsets.assign(n + 1, 0);
  res->max_degree = 0;
  for (int32_t i = 0; i < n; ++i) {
    auto& L = sym[i];
    std::sort(L.begin(), L.end());
    L.erase(std::unique(L.begin(), L.end()), L.end());
    res->offsets[i + 1] = res->offsets[i] + (int64_t)L.size();
    res->max_degree = std::max(res->max_degree, (int32_t)L.size());
  }
  res->ids.resize(res->offsets[n]);
  for (int32_t i = 0; i < n; ++i)
    std::copy(sym[i].begin(), sym[i].end(),
              res->ids.begin() + res->offsets[i]);
}
}  // namespace

extern "C" int rpd_power_neighbors(int32_t n, const double* xyzw,
                                   const double* box, int32_t n_threads,
                                   rpd_nbr** out) {
  if (n < 0 || !out || (n > 0 && !xyzw) || !box) return -1;
  auto S = [&](int32_t i) {
    return rpd_d4{xyzw[4 * i], xyzw[4 * i + 1], xyzw[4 * i + 2],
                  xyzw[4 * i + 3]};
  };
  std::vector<RpdClassGrid> grids;
  rpd_build_class_grids(n, xyzw, box, grids);
  const int kClasses = kRpdClasses;

  auto* res = new rpd_nbr();
  res->valid.assign(n, 0);
  std::vector<std::vector<int32_t>> nb(n);
  std::atomic<int32_t> next{0};
  std::atomic<int64_t> fails{0};
  auto worker = [&]() {
    Poly poly;
    for (;;) {
      int32_t i0 = next.fetch_add(64);
      if (i0 >= n) break;
      for (int32_t i = i0; i < std::min(n, i0 + 64); ++i) {
        rpd_d4 si = S(i);
        poly.init_box(box);
        poly.failures = 0;
        double R2 = poly.radius2(si.x, si.y, si.z);
        bool empty = false;
        for (int c = 0; c < kClasses && !empty; ++c) {
          const auto& g = grids[c];
          if (g.ids.empty()) continue;
          int64_t q0[3];
          for (int a = 0; a < 3; ++a)
            q0[a] = (int64_t)std::floor(((&si.x)[a] - g.org[a]) / g.cs);
          // rings (Chebyshev shells of cells around q0) from the first one
          // that meets the grid to the last
          int64_t rmax_ring = 0, rmin_ring = 0;
          for (int a = 0; a < 3; ++a) {
            rmax_ring = std::max<int64_t>(rmax_ring,
                                 std::max<int64_t>(std::llabs(q0[a]),
                                          std::llabs(q0[a] - (g.dim[a] - 1))));
            rmin_ring = std::max<int64_t>(
                rmin_ring, std::max<int64_t>(std::max<int64_t>(-q0[a], 0),
                                             q0[a] - (g.dim[a] - 1)));
          }
          for (int64_t r = rmin_ring; r <= rmax_ring && !empty; ++r) {
            double R = std::sqrt(R2) * (1 + 1e-9) + 1e-12;
            double bound = R + std::sqrt(std::max(0.0, R * R - si.w + g.wmax));
            if (r >= 1 && (r - 1) * g.cs > bound * (1 + 1e-9) + 1e-12) break;
            for (int64_t dz = -r; dz <= r && !empty; ++dz) {
              int64_t z = q0[2] + dz;
              if (z < 0 || z >= g.dim[2]) continue;
              for (int64_t dy = -r; dy <= r && !empty; ++dy) {
                int64_t y = q0[1] + dy;
                if (y < 0 || y >= g.dim[1]) continue;
                bool face = std::llabs(dz) == r || std::llabs(dy) == r;
                for (int64_t dx = -r; dx <= r; dx += (face ? 1 : 2 * r)) {
                  int64_t x = q0[0] + dx;
                  if (x >= 0 && x < g.dim[0]) {
                    int64_t cell = x + g.dim[0] * (y + (int64_t)g.dim[1] * z);
                    for (int64_t s = g.start[cell]; s < g.start[cell + 1];
                         ++s) {
                      int32_t j = g.ids[s];
                      if (j == i) continue;
                      rpd_d4 sj = S(j);
                      double dx_ = sj.x - si.x, dy_ = sj.y - si.y,
                             dz_ = sj.z - si.z;
                      double d = std::sqrt(dx_ * dx_ + dy_ * dy_ + dz_ * dz_);
                      double Rc = std::sqrt(R2) * (1 + 1e-9) + 1e-12;
                      double bj =
                          Rc + std::sqrt(std::max(0.0, Rc * Rc - si.w + sj.w));
                      if (d > bj * (1 + 1e-9) + 1e-12) continue;
                      if (poly.clip(rpd_bisector(si, sj), j)) {
                        if (poly.verts.empty()) {
                          empty = true;
                          break;
                        }
                        R2 = poly.radius2(si.x, si.y, si.z);
                      }
                    }
                  }
                  if (r == 0 || (!face && dx == r)) break;
                  if (empty) break;
                }
              }
            }
          }
        }
        fails += poly.failures;
        if (empty || poly.verts.empty()) continue;
        res->valid[i] = 1;
        auto& L = nb[i];
        for (const auto& v : poly.verts)
          for (int k = 0; k < 3; ++k) {
            int32_t sidx = poly.src[v.p[k]];
            if (sidx >= 0) L.push_back(sidx);
          }
        std::sort(L.begin(), L.end());
        L.erase(std::unique(L.begin(), L.end()), L.end());
      }
    }
  };
  int nt = n_threads > 0 ? n_threads : (int)std::thread::hardware_concurrency();
  nt = std::max(1, std::min(nt, 64));
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t) pool.emplace_back(worker);
  for (auto& th : pool) th.join();
  if (fails.load() != 0) {
    delete res;
    return -4;
  }
  symmetrize(n, nb, res);
  *out = res;
  return 0;
}

extern "C" int rpd_nbr_from_lists(int32_t n, const int32_t* valid,
                                  const int32_t* counts, const int32_t* lists,
                                  int32_t stride, rpd_nbr** out) {
  if (n < 0 || !out || (n > 0 && (!valid || !counts || !lists)) ||
      stride < 0)
    return -1;
  auto* res = new rpd_nbr();
  res->valid.assign(valid, valid + n);
  std::vector<std::vector<int32_t>> nb(n);
  for (int32_t i = 0; i < n; ++i) {
    if (counts[i] < 0 || counts[i] > stride) {
      delete res;
      return -1;
    }
    nb[i].assign(lists + (int64_t)i * stride,
                 lists + (int64_t)i * stride + counts[i]);
    for (int32_t j : nb[i])
      if (j < 0 || j >= n) {
        delete res;
        return -1;
      }
  }
  symmetrize(n, nb, res);
  *out = res;
  return 0;
}

extern "C" int64_t rpd_nbr_total(const rpd_nbr* h) {
  return h ? (int64_t)h->ids.size() : -1;
}
extern "C" int32_t rpd_nbr_max_degree(const rpd_nbr* h) {
  return h ? h->max_degree : -1;
}
extern "C" int rpd_nbr_get(const rpd_nbr* h, int32_t* valid, int64_t* offsets,
                           int32_t* ids) {
  if (!h) return -1;
  if (valid) std::copy(h->valid.begin(), h->valid.end(), valid);
  if (offsets) std::copy(h->offsets.begin(), h->offsets.end(), offsets);
  if (ids) std::copy(h->ids.begin(), h->ids.end(), ids);
  return 0;
}
extern "C" void rpd_nbr_free(rpd_nbr* h) { delete h; }
