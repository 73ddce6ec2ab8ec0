// HIP kernels of the restricted power diagram (RPD) for CDNA4 (gfx950).
//
// Pipeline of one rpd_run (see rpd_engine.hip for the host side):
//   k_site_bbox    one wavefront per flagged sphere: clips the domain box by
//                  the neighbour bisectors pushed outwards by the largest
//                  tet radius; the bbox of that polytope contains the
//                  centroid of every tet the sphere can relate to
//   k_bin_*        spheres registered into a uniform grid of tet centroids
//   k_tet_count /  one thread per tet: the reference's tet-sphere relation
//   k_tet_fill     (voronoi.cu:154-193) over the spheres of its bin, emitted
//                  in ascending site id, plus the mask of neighbour planes
//                  that may cut the tet
//   k_clip         G cells per wavefront, 64/G lanes per cell, cell state in
//                  LDS: the reference clip loop (convex_cell.cu:680-774)
//                  with lane-parallel conflict tests (ballot), serial cavity
//                  boundary walk, lane-parallel edge/vertex creation and
//                  FP64 geometry; tiers: small capacity, full capacity,
//                  full capacity with exact arithmetic
//   k_compact_*    deterministic compaction into the output stream
#include <hip/hip_runtime.h>

#include <type_traits>

#include "rpd_device.h"
#include "rpd_capi.h"

namespace rpd {

constexpr uint32_t kEnd = 255;

__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint64_t ballot(bool p) {
  return __ballot(p);
}

__device__ __forceinline__ rpd_d4 ld4(const double4* p) {
  double4 v = *p;
  return rpd_d4{v.x, v.y, v.z, v.w};
}

// Counts of a captured pipeline (hipGraph): a launch is sized by its
// capacity and reads the count the previous kernels left on the device;
// without a device count the host's count is exact.
__device__ __forceinline__ int64_t dev_count(int64_t cap, const int64_t* d,
                                             int64_t sub = 0) {
  return d ? min(cap, *d - sub) : cap;
}
__device__ __forceinline__ int64_t dev_count32(int64_t cap, const int32_t* d,
                                               int64_t sub = 0) {
  return d ? min(cap, (int64_t)*d - sub) : cap;
}

__global__ __launch_bounds__(256) void k_fill32(int32_t* p, int64_t n,
                                                int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// A captured run's re-run lists (cnt[1] small-tier overflows, cnt[2] exact
// re-runs, cnt[6] middle-tier overflows): after a clip launch's tiers the
// counts are checked against the list positions the host reserved
// (*overflow otherwise); with next_large >= 0 the counters then start at
// the positions reserved for the next launch.
__global__ void k_check_rerun_counts(int32_t* cnt, int32_t lim_large,
                                     int32_t lim_exact, int32_t lim_full,
                                     int32_t* overflow, int32_t next_large,
                                     int32_t next_exact, int32_t next_full) {
  if (threadIdx.x != 0) return;
  if (cnt[1] > lim_large || cnt[2] > lim_exact || cnt[6] > lim_full)
    *overflow = 1;
  if (next_large >= 0) {
    cnt[1] = next_large;
    cnt[2] = next_exact;
    cnt[6] = next_full;
  }
}

// --------------------------------------------------------- predicates
// det4x4(p1.x, p2.x, p3.x, e.x, p1.y, ..., p1.w, p2.w, p3.w, e.w) in the
// reference's evaluation order (convex_cell.cu:458-459,
// common_cuda.h:195-213), with a semi-static filter: every monomial passes
// through at most 9 roundings, |err| <= gamma_9 * 24 * prod_r max_c |a_rc|,
// 24 * gamma_9 = 2.398e-14; 2.5e-14 covers the rounding of the bound.
// Products may fuse with the following sum: a fused monomial passes
// through fewer roundings, so the bound holds, and a decided sign is the
// exact one either way (undecided signs go to the exact predicate).
// The caller passes the bound eps, computed from row maxima that dominate
// the four columns': the clip kernel keeps, per cell, the largest |a|,
// |b|, |c|, |d| over all planes it has added, so each row's maximum over a
// vertex's three planes and the new plane e is at most max(cell maximum,
// |e|) -- one bound per clip step, no smaller than the per-vertex one.
// Returns +1 conflict, -1 no conflict, 0 undecided.
__device__ __forceinline__ int conflict_filtered_eps(const rpd_d4& p1,
                                                     const rpd_d4& p2,
                                                     const rpd_d4& p3,
                                                     const rpd_d4& e,
                                                     double eps) {
#pragma clang fp contract(fast)
  const double a11 = p1.x, a12 = p2.x, a13 = p3.x, a14 = e.x;
  const double a21 = p1.y, a22 = p2.y, a23 = p3.y, a24 = e.y;
  const double a31 = p1.z, a32 = p2.z, a33 = p3.z, a34 = e.z;
  const double a41 = p1.w, a42 = p2.w, a43 = p3.w, a44 = e.w;
  double m12 = a21 * a12 - a11 * a22;
  double m13 = a31 * a12 - a11 * a32;
  double m14 = a41 * a12 - a11 * a42;
  double m23 = a31 * a22 - a21 * a32;
  double m24 = a41 * a22 - a21 * a42;
  double m34 = a41 * a32 - a31 * a42;
  double m123 = m23 * a13 - m13 * a23 + m12 * a33;
  double m124 = m24 * a13 - m14 * a23 + m12 * a43;
  double m134 = m34 * a13 - m14 * a33 + m13 * a43;
  double m234 = m34 * a23 - m24 * a33 + m23 * a43;
  double det = m234 * a14 - m134 * a24 + m124 * a34 - m123 * a44;
  if (det > eps) return 1;
  if (det < -eps) return -1;
  return 0;
}

__device__ __noinline__ int conflict_exact(rpd_d4 p1, rpd_d4 p2, rpd_d4 p3,
                                           rpd_d4 e) {
  return rpd_exact_det4_sign(p1, p2, p3, e);
}

// sign of det3(normals) == 0 ? (is_vertex_perturb, convex_cell.cu:274-316)
// returns 1 degenerate, 0 fine, -1 undecided
__device__ __forceinline__ int perturb_filtered(const rpd_d4& p1,
                                                const rpd_d4& p2,
                                                const rpd_d4& p3) {
  double d, q;
  rpd_det3p(p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, p3.x, p3.y, p3.z, d, q);
  if (!(q < 1e280)) return -1;
  double eps = 6.0e-16 * q + 1e-290;
  if (d > eps || d < -eps) return 0;
  if (q == 0.0) return 1;  // all products zero: exactly degenerate
  return -1;
}

__device__ __noinline__ int perturb_exact(rpd_d4 p1, rpd_d4 p2, rpd_d4 p3) {
  return rpd_exact_w_sign(p1, p2, p3) == 0 ? 1 : 0;
}

// --------------------------------------------------------- tet planes
// face k opposite vertex k with vertices tet_faces_lvid[k]
// (convex_cell.h:24-25): {2,1,3}, {0,2,3}, {1,0,3}, {0,1,2}
__device__ __forceinline__ rpd_d4 tri(const double4& a, const double4& b,
                                      const double4& c) {
  return rpd_tri_plane(a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z);
}
__device__ __forceinline__ void tet_planes(const double4 V[4], rpd_d4 pl[4]) {
  pl[0] = tri(V[2], V[1], V[3]);
  pl[1] = tri(V[0], V[2], V[3]);
  pl[2] = tri(V[1], V[0], V[3]);
  pl[3] = tri(V[0], V[1], V[2]);
}
// initial dual triangles (convex_cell.cu:183-186)
__constant__ uint8_t c_init_ver[4][3] = {{1, 3, 2}, {0, 2, 3}, {0, 3, 1},
                                         {0, 1, 2}};
// initial edges for local vertex pairs (0,1),(0,2),(0,3),(1,2),(1,3),(2,3)
// (convex_cell.cu:202-207)
__constant__ uint8_t c_init_edge[6][2] = {{2, 3}, {1, 3}, {1, 2},
                                          {0, 3}, {0, 2}, {0, 1}};

__device__ __forceinline__ uint32_t meta_byte(const TetMeta& m, int k) {
  // k in [0,16): x:0-3, y:4-7, z:8-11, w:12-15
  uint32_t w = k < 4 ? m.x : (k < 8 ? m.y : (k < 12 ? m.z : m.w));
  return (w >> (8 * (k & 3))) & 255u;
}

// ================================================= candidate gathering
struct BinGrid {
  double org[3];
  double inv_bs;
  int32_t dim[3];
  int32_t pad;
};

// One instance of a batch: its domain box, the largest distance of a tet
// corner from its tet's centroid (rho, with margin), and its candidate bin
// grid, whose bins are numbered from bin_base in the batch-wide bin space.
struct InstInfo {
  BinGrid g;
  int64_t bin_base;
  double dom[6];
  double rho;
  BinGrid sg;        // grid of sphere centres (learned proposals)
  int64_t sg_base;
  int32_t site_off;  // first site of the instance, number of sites
  int32_t n_site;
};

__device__ __forceinline__ int bin_coord(double x, double org, double inv,
                                         int dim) {
  double f = floor((x - org) * inv);
  int c = f < 0 ? 0 : (f >= dim ? dim - 1 : (int)f);
  return c;
}

// One wavefront per sphere.  Polytope = domain box clipped by the
// neighbour bisectors, each pushed out by rho (max tet circum-radius about
// the centroid plus a margin).  Output bbox (lo, hi); empty if lo > hi.
// The bisectors are applied nearest first (ascending signed distance of
// the sphere centre to the bisector), so the polytope shrinks early and
// later planes rarely cut; that order is also written to knn_near, where
// the relation test uses it to meet a rejecting neighbour early (the test
// itself does not depend on the order).
constexpr int kBoxV = 128;      // vertices (overflow: whole-domain box)
constexpr int kBoxP = 128;      // 6 box planes + 122 neighbours
constexpr int kNearMax = 128;   // longest list ranked (else kept as is)

// inline neighbour word: the site id, with bit 31 set when the site is not
// flagged (it bounds cells but has none of its own)
constexpr uint32_t kNearUnflagged = 0x80000000u;
__device__ __forceinline__ int32_t near_word(const DevSites& S, int m) {
  return (int32_t)((uint32_t)m | (S.flags[m] == 1u ? 0u : kNearUnflagged));
}
__device__ __forceinline__ int near_id(int32_t w) { return w & 0x7fffffff; }

__global__ __launch_bounds__(64) void k_site_bbox(DevSites S,
                                                  const InstInfo* inst,
                                                  const int32_t* site_inst,
                                                  double* bbox,  // 6/site
                                                  int32_t* knn_near,
                                                  double4* knn_site_near,
                                                  int32_t* n_fallback) {
  __shared__ double4 s_pl[kBoxP];
  __shared__ double4 s_X[2][kBoxV];
  __shared__ uint32_t s_t[2][kBoxV];
  __shared__ uint8_t s_out[kBoxV];
  __shared__ double s_key[kNearMax];
  __shared__ int32_t s_near[kNearMax];
  __shared__ int32_t s_new;
  const int i = blockIdx.x;
  const int lane = threadIdx.x;
  if (i >= S.n_site) return;
  const InstInfo& I = inst[site_inst[i]];
  const double* dom = I.dom;
  const double rho = I.rho;
  double* out = bbox + 6 * (int64_t)i;
  if (S.flags[i] != 1u) {
    if (lane == 0) {
      out[0] = out[1] = out[2] = 1.0;
      out[3] = out[4] = out[5] = -1.0;
    }
    return;
  }
  const int64_t k0 = S.knn_off[i];
  const int deg = (int)(S.knn_off[i + 1] - k0);
  bool fail = deg > kBoxP - 6;
  const rpd_d4 si = ld4(S.site + i);
  const bool ranked = deg <= kNearMax;
  if (ranked) {
    for (int k = lane; k < deg; k += 64) {
      const rpd_d4 e = rpd_bisector(si, ld4(S.site + S.knn_ids[k0 + k]));
      const double nn = e.x * e.x + e.y * e.y + e.z * e.z;
      double key = (e.x * si.x + e.y * si.y + e.z * si.z + e.w) / sqrt(nn);
      if (!(key == key)) key = -INFINITY;  // keep a strict total order
      s_key[k] = key;
    }
    wsync();
    // rank = position in ascending (key, index) order: a permutation
    for (int k = lane; k < deg; k += 64) {
      const double kk = s_key[k];
      int r = 0;
      for (int u = 0; u < deg; ++u) {
        const double ku = s_key[u];
        r += (ku < kk) || (ku == kk && u < k);
      }
      s_near[r] = S.knn_ids[k0 + k];
    }
    wsync();
    for (int k = lane; k < deg; k += 64) {
      const int m = s_near[k];
      knn_near[k0 + k] = near_word(S, m);
      knn_site_near[k0 + k] = S.site[m];
    }
  } else {
    for (int k = lane; k < deg; k += 64) {
      const int m = S.knn_ids[k0 + k];
      knn_near[k0 + k] = near_word(S, m);
      knn_site_near[k0 + k] = S.site[m];
    }
  }
  // box planes: x >= x0, x <= x1, y >= y0, y <= y1, z >= z0, z <= z1
  if (lane < 6) {
    int a = lane >> 1;
    double4 p = make_double4(0, 0, 0, 0);
    double* pp = &p.x;
    if ((lane & 1) == 0) {
      pp[a] = 1.0;
      p.w = -dom[a];
    } else {
      pp[a] = -1.0;
      p.w = dom[3 + a];
    }
    s_pl[lane] = p;
  }
  wsync();
  int nv = 8, cur = 0;
  if (lane < 8) {
    int a = lane & 1, b = 2 + ((lane >> 1) & 1), c = 4 + ((lane >> 2) & 1);
    rpd_d4 X, Q;
    rpd_vertex_homog(ld4(&s_pl[a]), ld4(&s_pl[b]), ld4(&s_pl[c]), X, Q);
    s_X[0][lane] = make_double4(X.x, X.y, X.z, X.w);
    s_t[0][lane] = a | (b << 10) | (c << 20);
  }
  wsync();
  bool empty = false;
  for (int k = 0; k < deg && !fail && !empty; ++k) {
    const int m = ranked ? s_near[k] : S.knn_ids[k0 + k];
    rpd_d4 e = rpd_bisector(si, ld4(S.site + m));
    double nrm = sqrt(e.x * e.x + e.y * e.y + e.z * e.z);
    e.w += nrm * rho;
    const int p = 6 + k;
    if (lane == 0) s_pl[p] = make_double4(e.x, e.y, e.z, e.w);
    int n_out = 0;
    for (int v = lane; v < nv; v += 64) {
      double4 X = s_X[cur][v];
      double s = e.x * X.x + e.y * X.y + e.z * X.z + e.w * X.w;
      bool o = (s < 0 && X.w > 0) || (s > 0 && X.w < 0);
      s_out[v] = o;
    }
    wsync();
    for (int v0 = 0; v0 < nv; v0 += 64) {
      int v = v0 + lane;
      n_out += __popcll(ballot(v < nv && s_out[v]));
    }
    if (n_out == 0) continue;
    if (n_out == nv) {
      empty = true;
      break;
    }
    if (lane == 0) s_new = 0;
    wsync();
    const int nxt = cur ^ 1;
    const int n_keep = nv - n_out;
    // compact kept vertices in order
    int base = 0;
    for (int v0 = 0; v0 < nv; v0 += 64) {
      int v = v0 + lane;
      bool keep = v < nv && !s_out[v];
      uint64_t bm = ballot(keep);
      if (keep) {
        int pos = base + __popcll(bm & ((1ull << lane) - 1));
        s_X[nxt][pos] = s_X[cur][v];
        s_t[nxt][pos] = s_t[cur][v];
      }
      base += __popcll(bm);
    }
    // new vertices on edges between removed and kept vertices
    bool lfail = false;
    for (int v = lane; v < nv; v += 64) {
      if (!s_out[v]) continue;
      uint32_t tv = s_t[cur][v];
      int pv[3] = {(int)(tv & 1023), (int)((tv >> 10) & 1023),
                   (int)((tv >> 20) & 1023)};
      for (int q = 0; q < 3; ++q) {
        int a = pv[q], b = pv[(q + 1) % 3];
        int other = -1;
        for (int u = 0; u < nv; ++u) {
          if (u == v) continue;
          uint32_t tu = s_t[cur][u];
          int u0 = tu & 1023, u1 = (tu >> 10) & 1023, u2 = (tu >> 20) & 1023;
          bool ha = u0 == a || u1 == a || u2 == a;
          bool hb = u0 == b || u1 == b || u2 == b;
          if (ha && hb) {
            other = u;
            break;
          }
        }
        if (other < 0) {
          lfail = true;
          continue;
        }
        if (s_out[other]) continue;
        int slot = atomicAdd(&s_new, 1);
        if (n_keep + slot >= kBoxV) {
          lfail = true;
          continue;
        }
        rpd_d4 X, Q;
        rpd_vertex_homog(ld4(&s_pl[a]), ld4(&s_pl[b]), ld4(&s_pl[p]), X, Q);
        s_X[nxt][n_keep + slot] = make_double4(X.x, X.y, X.z, X.w);
        s_t[nxt][n_keep + slot] = a | (b << 10) | (p << 20);
      }
    }
    fail = fail || (ballot(lfail) != 0);
    wsync();
    nv = n_keep + min(s_new, kBoxV - n_keep);
    cur = nxt;
    wsync();
  }
  if (fail) {
    if (lane == 0) {
      for (int a = 0; a < 6; ++a) out[a] = dom[a];
      atomicAdd(n_fallback, 1);
    }
    return;
  }
  if (empty) {
    if (lane == 0) {
      out[0] = out[1] = out[2] = 1.0;
      out[3] = out[4] = out[5] = -1.0;
    }
    return;
  }
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int v = lane; v < nv; v += 64) {
    double4 X = s_X[cur][v];
    double x = X.x / X.w, y = X.y / X.w, z = X.z / X.w;
    lo[0] = fmin(lo[0], x);
    lo[1] = fmin(lo[1], y);
    lo[2] = fmin(lo[2], z);
    hi[0] = fmax(hi[0], x);
    hi[1] = fmax(hi[1], y);
    hi[2] = fmax(hi[2], z);
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int a = 0; a < 3; ++a) {
      lo[a] = fmin(lo[a], __shfl_xor(lo[a], off));
      hi[a] = fmax(hi[a], __shfl_xor(hi[a], off));
    }
  if (lane == 0) {
    double ext = fmax(fmax(dom[3] - dom[0], dom[4] - dom[1]), dom[5] - dom[2]);
    double eps = 1e-7 * ext;  // covers the rounding of the vertices
    for (int a = 0; a < 3; ++a) {
      out[a] = lo[a] - eps;
      out[3 + a] = hi[a] + eps;
    }
  }
}

// bin range of a sphere's candidate box; false if empty
__device__ __forceinline__ bool bin_range(const double* b, const BinGrid& g,
                                          int r[6]) {
  if (b[0] > b[3] || b[1] > b[4] || b[2] > b[5]) return false;
  for (int a = 0; a < 3; ++a) {
    r[a] = bin_coord(b[a], g.org[a], g.inv_bs, g.dim[a]);
    r[3 + a] = bin_coord(b[3 + a], g.org[a], g.inv_bs, g.dim[a]);
  }
  return true;
}

// number of bins each sphere registers into
__global__ __launch_bounds__(256) void k_bin_nreg(int32_t n_site,
                                                  const double* bbox,
                                                  const InstInfo* inst,
                                                  const int32_t* site_inst,
                                                  int64_t* nreg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n_site) return;
  if (i == n_site) {
    nreg[i] = 0;
    return;
  }
  int r[6];
  nreg[i] = bin_range(bbox + 6 * (int64_t)i, inst[site_inst[i]].g, r)
                ? (int64_t)(r[3] - r[0] + 1) * (r[4] - r[1] + 1) *
                      (r[5] - r[2] + 1)
                : 0;
}

// one thread per registration: key = bin << 32 | sphere
__global__ __launch_bounds__(256) void k_bin_expand(int32_t n_site,
                                                    int64_t n_reg,
                                                    const double* bbox,
                                                    const InstInfo* inst,
                                                    const int32_t* site_inst,
                                                    const int64_t* reg_off,
                                                    uint64_t* keys,
                                                    const int64_t* n_reg_dev,
                                                    uint64_t pad) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_reg) return;
  if (n_reg_dev && q >= dev_count(n_reg, n_reg_dev)) {
    keys[q] = pad;  // sorts after every registration
    return;
  }
  // sphere i with reg_off[i] <= q < reg_off[i+1]
  int lo = 0, hi = n_site;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (reg_off[mid] <= q)
      lo = mid;
    else
      hi = mid;
  }
  const int i = lo;
  const InstInfo& I = inst[site_inst[i]];
  const BinGrid& g = I.g;
  int r[6];
  bin_range(bbox + 6 * (int64_t)i, g, r);
  const int64_t k = q - reg_off[i];
  const int64_t nx = r[3] - r[0] + 1, ny = r[4] - r[1] + 1;
  const int64_t x = r[0] + k % nx, y = r[1] + (k / nx) % ny,
                z = r[2] + k / (nx * ny);
  const uint64_t bin = (uint64_t)I.bin_base + x +
                       (uint64_t)g.dim[0] * (y + (uint64_t)g.dim[1] * z);
  keys[q] = (bin << 32) | (uint32_t)i;
}

// bin_start[b] = first registration of bin >= b in the sorted keys
__global__ __launch_bounds__(256) void k_bin_starts(int64_t n_bins,
                                                    int64_t n_reg_cap,
                                                    const uint64_t* keys,
                                                    int32_t* bin_start,
                                                    int32_t* bin_ids,
                                                    const int64_t* n_reg_dev) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n_reg = dev_count(n_reg_cap, n_reg_dev);
  if (b < n_reg) bin_ids[b] = (int32_t)(keys[b] & 0xffffffffu);
  if (b > n_bins) return;
  int64_t lo = 0, hi = n_reg;  // first index with key >> 32 >= b
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)(keys[mid] >> 32) < b)
      lo = mid + 1;
    else
      hi = mid;
  }
  bin_start[b] = (int32_t)lo;
}

struct TetLoad {
  double4 V[4];
  int4 tv;
};

__device__ __forceinline__ void load_tet(const DevMesh& M, int t,
                                         TetLoad& L) {
  L.tv = M.tet_v[t];
  L.V[0] = M.vert[L.tv.x];
  L.V[1] = M.vert[L.tv.y];
  L.V[2] = M.vert[L.tv.z];
  L.V[3] = M.vert[L.tv.w];
}

// reference relation test (voronoi.cu:154-193) in binary64: sphere i
// relates to the tet iff for every neighbour m some corner v has
// pd(v, m) > pd(v, i).  The neighbours are visited nearest bisector first
// (the outcome does not depend on the order), their spheres read inline,
// one ahead.  *iters counts the neighbour tests made.
__device__ __forceinline__ bool relation(const DevSites& S, const TetLoad& L,
                                         int i, uint32_t* iters) {
  if (S.flags[i] != 1u) return false;
  const rpd_d4 si = ld4(S.site + i);
  double pdi[4];
  for (int k = 0; k < 4; ++k)
    pdi[k] = rpd_power_dist(L.V[k].x, L.V[k].y, L.V[k].z, si);
  const int64_t k0 = S.knn_off[i], k1 = S.knn_off[i + 1];
  if (k0 == k1) return true;
  double4 nxt = S.knn_site_near[k0];
  for (int64_t q = k0; q < k1; ++q) {
    const rpd_d4 sm{nxt.x, nxt.y, nxt.z, nxt.w};
    if (q + 1 < k1) nxt = S.knn_site_near[q + 1];
    ++*iters;
    bool ok = false;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      ok = ok || rpd_power_dist(L.V[k].x, L.V[k].y, L.V[k].z, sm) > pdi[k];
    if (!ok) return false;
  }
  return true;
}

__device__ __forceinline__ void tet_centroid(const TetLoad& L, double c[3]) {
  c[0] = (((L.V[0].x + L.V[1].x) + L.V[2].x) + L.V[3].x) * 0.25;
  c[1] = (((L.V[0].y + L.V[1].y) + L.V[2].y) + L.V[3].y) * 0.25;
  c[2] = (((L.V[0].z + L.V[1].z) + L.V[2].z) + L.V[3].z) * 0.25;
}

__device__ __forceinline__ int64_t tet_bin(const double c[3],
                                           const BinGrid& g) {
  int x = bin_coord(c[0], g.org[0], g.inv_bs, g.dim[0]);
  int y = bin_coord(c[1], g.org[1], g.inv_bs, g.dim[1]);
  int z = bin_coord(c[2], g.org[2], g.inv_bs, g.dim[2]);
  return x + (int64_t)g.dim[0] * (y + (int64_t)g.dim[1] * z);
}

__global__ __launch_bounds__(256) void k_tet_count(DevMesh M, DevSites S,
                                                   const InstInfo* inst,
                                                   const int32_t* tet_inst,
                                                   const int32_t* bin_start,
                                                   const int32_t* bin_ids,
                                                   int64_t* tet_count,
                                                   uint64_t* tet_pass,
                                                   unsigned long long* work,
                                                   const uint8_t* tet_sel) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M.n_tet) return;
  if (tet_sel && !tet_sel[t]) {  // outside the tet subset: no candidates
    tet_count[t] = 0;
    return;
  }
  TetLoad L;
  load_tet(M, t, L);
  double c[3];
  tet_centroid(L, c);
  const InstInfo& I = inst[tet_inst[t]];
  const int64_t b = I.bin_base + tet_bin(c, I.g);
  int64_t n = 0;
  // bit r of (pass0, pass1): the r-th site of the bin list relates
  uint64_t pass0 = 0, pass1 = 0;
  uint32_t iters = 0;
  const int q0 = bin_start[b], q1 = bin_start[b + 1];
  for (int q = q0; q < q1; ++q) {
    const int i = bin_ids[q];
    // (the candidate-box test of the centroid is left to the binning: it
    // passes 3 of 4 entries and costs more than it saves)
    const bool r = relation(S, L, i, &iters);
    n += r;
    const int d = q - q0;
    if (r && d < 64) pass0 |= 1ull << d;
    if (r && d >= 64 && d < 128) pass1 |= 1ull << (d - 64);
  }
  tet_count[t] = n;
  tet_pass[t] = pass0;
  tet_pass[(int64_t)M.n_tet + t] = pass1;
  if (work) {  // entries, neighbour tests (count_work)
    unsigned long long* w = work + 4 * (blockIdx.x & 63);
    atomicAdd(w + 0, (unsigned long long)(q1 - q0));

    atomicAdd(w + 1, (unsigned long long)iters);
  }
}

__global__ __launch_bounds__(256) void k_tet_fill(
    DevMesh M, DevSites S, const InstInfo* inst, const int32_t* tet_inst,
    const int32_t* bin_start, const int32_t* bin_ids, const int64_t* tet_off,
    const uint64_t* tet_pass, Cand* cand, int32_t* rel_ids,
    const uint8_t* tet_sel, int64_t cap, int32_t* overflow) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M.n_tet) return;
  if (tet_sel && !tet_sel[t]) return;
  // over a captured pipeline's capacity: the entries below it are still
  // written (every slot a later launch reads holds a pair of this run)
  if (tet_off[t + 1] > cap && (tet_off[t] < cap || t == 0)) overflow[0] = 1;
  if (tet_off[t] >= cap) return;
  TetLoad L;
  load_tet(M, t, L);
  double c[3];
  tet_centroid(L, c);
  const InstInfo& I = inst[tet_inst[t]];
  const int64_t b = I.bin_base + tet_bin(c, I.g);
  int64_t o = tet_off[t];
  const uint64_t pass0 = tet_pass[t];
  const uint64_t pass1 = tet_pass[(int64_t)M.n_tet + t];
  const int q0 = bin_start[b];
  for (int q = q0; q < bin_start[b + 1]; ++q) {
    const int i = bin_ids[q];
    // the first 128 outcomes come from the counting pass
    uint32_t iters = 0;
    const int d = q - q0;
    const bool r = d < 64    ? ((pass0 >> d) & 1)
                   : d < 128 ? ((pass1 >> (d - 64)) & 1)
                             : relation(S, L, i, &iters);
    if (!r) continue;
    if (o >= cap) break;
    Cand c;
    c.tet = t;
    c.seed = i;
    c.m0 = 0;
    c.m1 = 0;
    cand[o] = c;
    rel_ids[o] = i;
    ++o;
  }
}

// the bisector table of the neighbour lists: one wavefront per site
__global__ __launch_bounds__(64) void k_knn_bis(DevSites S, double4* bis,
                                                float4* bisf) {
  const int i = blockIdx.x;
  const int64_t k0 = S.knn_off[i], k1 = S.knn_off[i + 1];
  const rpd_d4 si = ld4(S.site + i);
  for (int64_t q = k0 + threadIdx.x; q < k1; q += 64) {
    const int m = S.knn_ids[q];
    double4 e = make_double4(0, 0, 0, 0);
    if (m >= 0) {
      const rpd_d4 b = rpd_bisector(si, ld4(S.site + m));
      e = make_double4(b.x, b.y, b.z, b.w);
    }
    bis[q] = e;
    bisf[q] = make_float4(rpd_mask_coef(e.x), rpd_mask_coef(e.y),
                          rpd_mask_coef(e.z), rpd_mask_coef(e.w));
  }
}

// The tet's four corners as the clip kernel's initial vertices: homogeneous
// coordinates X and cofactor magnitudes Q of the dual triangles (1,3,2),
// (0,2,3), (0,3,1), (0,1,2) of its binary64 face planes; sQ = the sum of
// a corner's four magnitudes.
struct CornerFrames {
  rpd_d4 X[4];
  double sQ[4];
  // dot(e, X~) is certainly negative (no conflict) at all four corners, so
  // the plane e cannot cut any state of a cell of this tet (every vertex of
  // the cell lies in the convex hull of the corners); decided from the
  // binary32 table entry f of e (rpd_mask_negative), one 16-byte load per
  // plane.  The few planes it leaves undecided that a tighter test would
  // skip are only clipped without effect.
  __device__ __forceinline__ bool inactive(const float4& f) const {
    return rpd_mask_negative(f.x, f.y, f.z, f.w, X, sQ);
  }
};
__device__ __forceinline__ void corner_frames(const TetLoad& L,
                                              CornerFrames& F) {
  rpd_d4 pl[4], Q[4];
  tet_planes(L.V, pl);
  rpd_vertex_homog(pl[1], pl[3], pl[2], F.X[0], Q[0]);
  rpd_vertex_homog(pl[0], pl[2], pl[3], F.X[1], Q[1]);
  rpd_vertex_homog(pl[0], pl[3], pl[1], F.X[2], Q[2]);
  rpd_vertex_homog(pl[0], pl[1], pl[2], F.X[3], Q[3]);
#pragma unroll
  for (int k = 0; k < 4; ++k) F.sQ[k] = ((Q[k].x + Q[k].y) + Q[k].z) + Q[k].w;
}
// a cell's number of clip steps: mask bits, plus the list positions past
// 128 (always clipped), saturated to a byte
__device__ __forceinline__ uint8_t clip_steps(uint64_t m0, uint64_t m1,
                                              int deg) {
  const int n = __popcll(m0) + __popcll(m1) + (deg > 128 ? deg - 128 : 0);
  return (uint8_t)(n < 255 ? n : 255);
}

// One thread per candidate cell: bit k of the mask is set unless the k-th
// neighbour's bisector is certainly positive at all four corners of the
// tet cell (then it cannot cut any state of the cell: every cell vertex
// lies in the convex hull of the corners).
// With trips != nullptr it also writes each cell's number of clip steps
// (mask bits below the degree, plus the positions past 128) and its index
// c0 + c for the step-count ordering of the clip launch.
__global__ __launch_bounds__(256) void k_cand_mask(DevMesh M, DevSites S,
                                                   int64_t n_cand,
                                                   Cand* cand,
                                                   uint8_t* trips,
                                                   int32_t* trip_ids,
                                                   int64_t c0,
                                                   const int64_t* n_dev) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cand) return;
  if (n_dev && c >= dev_count(n_cand, n_dev, c0)) {
    if (trips) {  // past the count: sorts after every cell
      trips[c] = 0;
      trip_ids[c] = (int32_t)(c0 + c);
    }
    return;
  }
  const int t = cand[c].tet, i = cand[c].seed;
  TetLoad L;
  load_tet(M, t, L);
  CornerFrames F;
  corner_frames(L, F);
  const int64_t k0 = S.knn_off[i];
  const int deg = (int)(S.knn_off[i + 1] - k0);
  uint64_t m0 = 0, m1 = 0;
  const int lim = deg < 128 ? deg : 128;
  // the next plane's load is issued before this plane's test (the test's
  // corner branches keep the compiler from hoisting it)
  float4 nxt = lim > 0 ? S.knn_bisf[k0] : make_float4(0, 0, 0, 0);
  for (int k = 0; k < lim; ++k) {
    const float4 f = nxt;
    if (k + 1 < lim) nxt = S.knn_bisf[k0 + k + 1];
    if (!F.inactive(f)) {
      if (k < 64)
        m0 |= 1ull << k;
      else
        m1 |= 1ull << (k - 64);
    }
  }
  cand[c].m0 = m0;
  cand[c].m1 = m1;
  if (trips) {
    trips[c] = clip_steps(m0, m1, deg);
    trip_ids[c] = (int32_t)(c0 + c);
  }
}

// ============================================================ clipping
// binary search in a sorted u64 array
__device__ __forceinline__ bool has_key(const uint64_t* keys, int64_t n,
                                        uint64_t k) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < k)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < n && keys[lo] == k;
}

// Completeness certificate of the learned path, evaluated on each cell
// when its status is final (tet_off == nullptr: off).  The cells of a tet's
// sphere set partition the tet iff the set is closed under facet
// neighbours: every active bisector facet (seed, j) of a cell must have j
// in the set -- the tet's CSR members, or a pair (tet, j) of the sorted
// ext keys added since.  A missing pair goes to miss (past cap: the tet is
// flagged); a cell that ends in a failure status flags its tet; a tet with
// a cell is marked in has_cell.
struct CertArgs {
  const int64_t* tet_off;
  const int32_t* mem_ids;
  const uint64_t* ext;
  int64_t n_ext;
  uint64_t* miss;
  int32_t cap;
  int32_t* n_miss;
  uint8_t* flag;
  uint8_t* has_cell;
  // (a captured run: the candidates of the launch, on the device; the miss
  // list then holds min(cap, max(4096, 2 * *cap_n)) entries, as the eager
  // path sizes it)
  const int64_t* cap_n;
};

struct ClipArgs {
  DevMesh M;
  DevSites S;
  const Cand* cand;
  const int32_t* items;   // cell indices to process (nullptr: item_base +
                          // item)
  int64_t item_base;
  int64_t n_items;
  SlotView out;           // indexed by slot_base + item, or by the cell
                          // index when slot_by_cell
  int64_t slot_base;
  int32_t slot_by_cell;   // items in step-count order: slot = cell index,
                          // and no XCD regrouping of the blocks
  int32_t use_mask;
  int32_t force_exact;
  int32_t geometry;
  int32_t* rerun_large;   // list + counter (small tier)
  int32_t* n_rerun_large;
  int32_t* rerun_exact;
  int32_t* n_rerun_exact;
  // a captured pipeline: n_items is the launch's capacity and the count is
  // on the device (n_items_d64, or n_items_d32 minus n_sub); the re-run
  // lists hold cap_large / cap_exact entries, more set overflow[0]
  const int64_t* n_items_d64;
  const int32_t* n_items_d32;
  int64_t n_sub;
  int32_t cap_large, cap_exact;
  int32_t* overflow;
  unsigned long long* work;  // 64 x 4 spread counters, or nullptr
  CertArgs cert;
};

// 128-bit vertex masks as (lo, hi) words
__device__ __forceinline__ bool mbit(uint64_t lo, uint64_t hi, int v) {
  return v < 64 ? (lo >> v) & 1 : (hi >> (v - 64)) & 1;
}
__device__ __forceinline__ uint64_t low_bits(int n) {  // bits [0, n), n <= 64
  return n >= 64 ? ~0ull : (1ull << n) - 1;
}
// bits [a, b), 0 <= a <= b <= 128
__device__ __forceinline__ void mrange(int a, int b, uint64_t& lo,
                                       uint64_t& hi) {
  lo = low_bits(min(b, 64)) & ~low_bits(min(a, 64));
  hi = low_bits(max(b, 64) - 64) & ~low_bits(max(a, 64) - 64);
}
// position of the k-th lowest (k >= 1) set bit
__device__ __forceinline__ int mkth_low(uint64_t lo, uint64_t hi, int k) {
  const int nl = __popcll(lo);
  uint64_t w = lo;
  int base = 0;
  if (k > nl) {
    k -= nl;
    w = hi;
    base = 64;
  }
  for (int i = 1; i < k; ++i) w &= w - 1;
  return base + __ffsll((long long)w) - 1;
}
// position of the k-th highest (k >= 1) set bit
__device__ __forceinline__ int mkth_high(uint64_t lo, uint64_t hi, int k) {
  const int nh = __popcll(hi);
  uint64_t w = hi;
  int base = 64;
  if (k > nh) {
    k -= nh;
    w = lo;
    base = 0;
  }
  for (int i = 1; i < k; ++i) w &= ~(1ull << (63 - __clzll((long long)w)));
  return base + 63 - __clzll((long long)w);
}

template <int L>
__device__ __forceinline__ uint64_t group_bits(uint64_t b, int g) {
  if constexpr (L == 64) {
    return b;
  } else {
    return (b >> (g * L)) & ((1ull << L) - 1);
  }
}

// ---- sums and ORs over the L lanes of a cell's group, butterfly with
// offsets L/2, ..., 1.  Groups of 4 lanes are quads: their exchanges are
// DPP quad permutations (one VALU move per 32 bits) instead of LDS-routed
// lane shuffles.
template <int C>
__device__ __forceinline__ uint32_t quad_perm(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, C, 0xF, 0xF, false);
}
template <int C>
__device__ __forceinline__ double quad_perm(double v) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const uint64_t lo = quad_perm<C>((uint32_t)u);
  const uint64_t hi = quad_perm<C>((uint32_t)(u >> 32));
  return __longlong_as_double((long long)((hi << 32) | lo));
}
constexpr int kXor2 = 0x4E;  // quad lanes (2, 3, 0, 1)
constexpr int kXor1 = 0xB1;  // quad lanes (1, 0, 3, 2)
template <int L>
__device__ __forceinline__ double group_sum(double x) {
  if constexpr (L == 4) {
    x += quad_perm<kXor2>(x);
    x += quad_perm<kXor1>(x);
  } else {
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) x += __shfl_xor(x, off);
  }
  return x;
}
template <int L>
__device__ __forceinline__ uint32_t group_or(uint32_t x) {
  if constexpr (L == 4) {
    x |= quad_perm<kXor2>(x);
    x |= quad_perm<kXor1>(x);
  } else {
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) x |= __shfl_xor(x, off);
  }
  return x;
}

// ---- FP64 geometry helpers (tolerance-checked, not bit-compared)
// 1/x: hardware reciprocal estimate refined by two Newton steps (relative
// error ~1 ulp for the normal, finite x used here)
__device__ __forceinline__ double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
// foot F = C - s n_p of C on plane p, given 1/|n_p|^2; returns s =
// (n_p.C + d_p) / |n_p|^2
__device__ __forceinline__ double foot_plane(const rpd_d4& p, double inv_nn,
                                             const double C[3], double F[3]) {
#pragma clang fp contract(fast)
  const double s = (p.x * C[0] + p.y * C[1] + p.z * C[2] + p.w) * inv_nn;
  F[0] = C[0] - s * p.x;
  F[1] = C[1] - s * p.y;
  F[2] = C[2] - s * p.z;
  return s;
}
// foot of C on the line through A with direction n_p x n_q
__device__ __forceinline__ void foot_line(const rpd_d4& p, const rpd_d4& q,
                                          const double A[3], const double C[3],
                                          double E[3]) {
#pragma clang fp contract(fast)
  const double d0 = p.y * q.z - p.z * q.y;
  const double d1 = p.z * q.x - p.x * q.z;
  const double d2 = p.x * q.y - p.y * q.x;
  const double dd = d0 * d0 + d1 * d1 + d2 * d2;
  const double s =
      ((C[0] - A[0]) * d0 + (C[1] - A[1]) * d1 + (C[2] - A[2]) * d2) *
      rcp64(dd);
  E[0] = A[0] + s * d0;
  E[1] = A[1] + s * d1;
  E[2] = A[2] + s * d2;
}
// twice the signed area of triangle (P, Q, A) along the (unnormalised)
// normal n
__device__ __forceinline__ double tri_area2(const double P[3],
                                            const double Q[3],
                                            const double A[3],
                                            const rpd_d4& n) {
#pragma clang fp contract(fast)
  const double u0 = Q[0] - P[0], u1 = Q[1] - P[1], u2 = Q[2] - P[2];
  const double v0 = A[0] - P[0], v1 = A[1] - P[1], v2 = A[2] - P[2];
  return (u1 * v2 - u2 * v1) * n.x + (u2 * v0 - u0 * v2) * n.y +
         (u0 * v1 - u1 * v0) * n.z;
}
__device__ __forceinline__ void vertex_pos(const rpd_d4& a, const rpd_d4& b,
                                           const rpd_d4& c, double A[3]) {
  rpd_d4 X, Q;
  rpd_vertex_homog(a, b, c, X, Q);
  const double r = rcp64(X.w);
  A[0] = X.x * r;
  A[1] = X.y * r;
  A[2] = X.z * r;
}

// resident wavefronts per SIMD the compiler must allow: 4 for the 16-cell
// shape with the tightest capacities (its register count otherwise lands
// just above the 4-wave limit of 128; measured 8 % faster), no constraint
// otherwise (the same hint slows the 8-cell shapes)
template <int G, class C>
constexpr int clip_min_waves() {
  return (G == 16 && C::V <= 16) ? 4 : 1;
}

// kCert: the learned path's launches, which evaluate the completeness
// certificate on every final cell (a.cert); the plain path's launches are
// compiled without it
template <int G, class C, bool kExact, bool kCert>
__global__ __launch_bounds__(64, (clip_min_waves<G, C>())) void k_clip(
    ClipArgs a) {
  constexpr int L = 64 / G;
  static_assert(L >= 4, "the tet's four planes are set up one per lane");
  constexpr int V = C::V, P = C::P, E = C::E;
  constexpr int VS = (V + L - 1) / L;
  constexpr bool kFull = (P == RPD_MAX_P);
  __shared__ double4 s_pl[G][P];
  __shared__ int32_t s_src[G][P];
  __shared__ uint8_t s_h[G][P];
  __shared__ uint8_t s_bn[G][P];
  __shared__ uint8_t s_cyc[G][P + 1];
  __shared__ uint32_t s_ver[G][V];
  // s_adj[p]: bit q (q < p) set iff the edge list holds the pair (q, p)
  using AdjT = typename std::conditional<
      (P <= 16), uint16_t,
      typename std::conditional<(P <= 32), uint32_t, uint64_t>::type>::type;
  __shared__ AdjT s_adj[G][P];
  // per cell: the largest |a|, |b|, |c|, |d| over its planes (filter bound)
  __shared__ double s_mx[G][4];

  const int lane = threadIdx.x;
  const int g = lane / L, j = lane % L;
  // XCD-aware order: consecutive items stay on one XCD (8 XCDs,
  // round-robin dispatch of workgroups)
  const int64_t nb = gridDim.x;
  int64_t blk = blockIdx.x;
  if (nb >= 8 && !a.slot_by_cell) {
    const int64_t per = nb / 8;
    const int64_t xcd = blk % 8, pos = blk / 8;
    if (pos < per) blk = xcd * per + pos;
  }
  const int64_t item = blk * G + g;
  const int64_t n_items =
      a.n_items_d64 ? dev_count(a.n_items, a.n_items_d64)
                    : dev_count32(a.n_items, a.n_items_d32, a.n_sub);
  if (item >= n_items) return;
  const int64_t cidx = a.items ? a.items[item] : a.item_base + item;
  const int64_t slot = a.slot_by_cell ? cidx : a.slot_base + item;
  // the slot's edge records are written at its start as the edges are
  // appended (the reference never removes an edge), so they need no LDS
  // copy; the vertices and plane sources follow the last one at the end
  uint32_t* const wslot = a.out.words + slot * a.out.stride_words;
  uint32_t* const wedge = wslot;
  const Cand cd = a.cand[cidx];
  const int t = cd.tet, s = cd.seed;
  const int4 tv = a.M.tet_v[t];
  const TetMeta tm = a.M.tet_meta[t];
  const int64_t k0 = a.S.knn_off[s];
  const int deg = (int)(a.S.knn_off[s + 1] - k0);
  uint64_t m0 = a.use_mask ? cd.m0 : ~0ull;
  uint64_t m1 = a.use_mask ? cd.m1 : ~0ull;

  // ---- init the cell as its tet (convex_cell.cu:116-214)
  {
    if (j < 4) {
      // face j = tet_faces_lvid[j]: {2,1,3}, {0,2,3}, {1,0,3}, {0,1,2}
      const int ia = j == 0 ? tv.z : (j == 2 ? tv.y : tv.x);
      const int ib = j == 0 ? tv.y : (j == 1 ? tv.z : (j == 2 ? tv.x : tv.y));
      const int ic = j == 3 ? tv.z : tv.w;
      const rpd_d4 p = tri(a.M.vert[ia], a.M.vert[ib], a.M.vert[ic]);
      s_pl[g][j] = make_double4(p.x, p.y, p.z, p.w);
      s_src[g][j] = -1 - j;
      s_h[g][j] = (uint8_t)meta_byte(tm, j);
      s_ver[g][j] = c_init_ver[j][0] | (c_init_ver[j][1] << 8) |
                    (c_init_ver[j][2] << 16) | (meta_byte(tm, 4 + j) << 24);
      s_adj[g][j] = (AdjT)((1u << j) - 1u);  // the six edges of the tet
    }
    for (int q = j; q < 6; q += L)
      wedge[q] = c_init_edge[q][0] | (c_init_edge[q][1] << 8) |
                 (meta_byte(tm, 8 + q) << 16);
  }
  wsync();
  if (j < 4) {  // row maxima of the four tet planes, one row per lane
    double m = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double4 pq = s_pl[g][q];
      const double v = j == 0 ? pq.x : j == 1 ? pq.y : j == 2 ? pq.z : pq.w;
      m = fmax(m, fabs(v));
    }
    s_mx[g][j] = m;
  }
  wsync();
  int nb_v = 4, nb_p = 4, nb_e = 6;
  uint32_t status = RPD_SUCCESS;
  uint32_t n_det = 0, n_bis = 0, n_exd = 0;  // executed work (count_work)
  bool on_plane = false;  // exact tier: a vertex lay exactly on a plane
  // next neighbour position after k that may cut the cell
  auto next_pos = [&](int k) {
    int nk = k + 1;
    if (nk < 64) {
      uint64_t r = m0 >> nk;
      if (r)
        nk += __ffsll((long long)r) - 1;
      else
        nk = 64;
    }
    if (nk >= 64 && nk < 128) {
      uint64_t r = m1 >> (nk - 64);
      if (r)
        nk += __ffsll((long long)r) - 1;
      else
        nk = 128;
    }
    return nk;
  };
  // the plane of the next step is loaded one step ahead (the masks fix the
  // positions in advance), so its latency overlaps the current step
  int k = -1;
  int kp = next_pos(-1);
  int mp = -1;
  rpd_d4 ep{0, 0, 0, 0};
  if (kp < deg) {
    mp = a.S.knn_ids[k0 + kp];
    ep = ld4(a.S.knn_bis + k0 + kp);
  }
  for (;;) {
    // at the reference's plane capacity every further neighbour overflows,
    // whether or not it can cut (new_plane checks the count before the
    // plane is tested, convex_cell.cu:561-565), so a masked one does too
    if (kFull && nb_p >= P && k + 1 < deg) {
      status = RPD_VERTEX_OVERFLOW;
      break;
    }
    k = kp;
    if (k >= deg) break;
    if (nb_p >= P) {
      status = kFull ? RPD_VERTEX_OVERFLOW : kStatusRerunLarge;
      break;
    }
    const int m = mp;
    const rpd_d4 e = ep;
    kp = next_pos(k);
    if (kp < deg) {
      mp = a.S.knn_ids[k0 + kp];
      ep = ld4(a.S.knn_bis + k0 + kp);
    }
    if (m == -1) break;
    const int cur = nb_p;
    n_bis += 1;
    n_det += nb_v;

    // ---- conflict tests, lane-parallel (convex_cell.cu:706-721); one
    // filter bound per step from the cell's row maxima and the new plane
    const double mx = fmax(s_mx[g][0], fabs(e.x));
    const double my = fmax(s_mx[g][1], fabs(e.y));
    const double mz = fmax(s_mx[g][2], fabs(e.z));
    const double mw = fmax(s_mx[g][3], fabs(e.w));
    const double bnd = (mx * my) * (mz * mw);
    // an unrepresentable bound leaves every sign undecided
    const double eps = bnd < 1e280 ? 2.5e-14 * bnd + 1e-290 : INFINITY;
    uint64_t conf_lo = 0, conf_hi = 0;
    bool undecided = false;
#pragma unroll
    for (int q = 0; q < VS; ++q) {
      // slots no cell of the wavefront has reached are not issued
      if (q > 0 && ballot(nb_v > q * L) == 0) break;
      const int v = q * L + j;
      int c = -1;
      if (v < nb_v) {
        const uint32_t w = s_ver[g][v];
        const rpd_d4 p1 = ld4(&s_pl[g][w & 255]);
        const rpd_d4 p2 = ld4(&s_pl[g][(w >> 8) & 255]);
        const rpd_d4 p3 = ld4(&s_pl[g][(w >> 16) & 255]);
        c = a.force_exact ? 0 : conflict_filtered_eps(p1, p2, p3, e, eps);
        if (kExact && c == 0) {
          const int sg = conflict_exact(p1, p2, p3, e);
          on_plane = on_plane || sg == 0;
          c = sg > 0 ? 1 : -1;
          ++n_exd;
        }
      }
      const uint64_t bc = group_bits<L>(ballot(c > 0), g);
      const uint64_t bu = group_bits<L>(ballot(c == 0), g);
      if (q * L < 64)
        conf_lo |= bc << (q * L);
      else
        conf_hi |= bc << (q * L - 64);
      undecided = undecided || bu != 0;
    }
    if (undecided) {
      status = kStatusRerunExact;
      break;
    }
    const int nr = __popcll(conf_lo) + __popcll(conf_hi);
    if (nr == 0) continue;  // the plane does not cut: popped (nb_p--)
    if (nr == nb_v) {
      status = RPD_NO_INTERSECTION;
      break;
    }
    if (j == 0) {
      s_pl[g][cur] = make_double4(e.x, e.y, e.z, e.w);
      s_src[g][cur] = m;
      s_h[g][cur] = 1;  // F_CELL_ADJ_DEFAULT
      s_adj[g][cur] = 0;
      s_mx[g][0] = mx;
      s_mx[g][1] = my;
      s_mx[g][2] = mz;
      s_mx[g][3] = mw;
    }
    nb_p = cur + 1;

    // ---- move conflicting vertices to the tail exactly as the serial
    // swap loop does (i scans from the front, end from the back, a
    // conflicting vertex at i is swapped with the one at end - 1), in
    // closed form: with nk = nb_v - nr kept vertices, the front pass reads
    // F = [0, m), m = nk + (vertex nk conflicts), the back pass reads
    // B = [m, nb_v) downwards.  A kept vertex of F stays; the j-th
    // conflicting vertex of F goes to nb_v - 1 (j = 1) or just below the
    // (j-1)-th kept vertex of B; the k-th kept vertex of B takes the place
    // of the k-th conflicting vertex of F; a conflicting vertex v of B goes
    // to v - 1.
    {
      const int nk = nb_v - nr;
      const int m = nk + (int)mbit(conf_lo, conf_hi, nk);
      uint64_t f_lo, f_hi, b_lo, b_hi;  // F's conflicts, B's kept vertices
      mrange(0, m, f_lo, f_hi);
      mrange(m, nb_v, b_lo, b_hi);
      f_lo &= conf_lo;
      f_hi &= conf_hi;
      b_lo &= ~conf_lo;
      b_hi &= ~conf_hi;
      uint32_t val[VS];
      int dst[VS];
#pragma unroll
      for (int q = 0; q < VS; ++q) {
        val[q] = 0;
        dst[q] = -1;
      }
#pragma unroll
      for (int q = 0; q < VS; ++q) {
        if (q > 0 && ballot(nb_v > q * L) == 0) break;
        const int v = q * L + j;
        if (v < nb_v) {
          val[q] = s_ver[g][v];
          const bool c = mbit(conf_lo, conf_hi, v);
          if (v < m) {
            if (!c) {
              dst[q] = v;
            } else {
              uint64_t r_lo, r_hi;
              mrange(0, v + 1, r_lo, r_hi);
              const int jj = __popcll(f_lo & r_lo) + __popcll(f_hi & r_hi);
              dst[q] = jj == 1 ? nb_v - 1 : mkth_high(b_lo, b_hi, jj - 1) - 1;
            }
          } else if (!c) {
            uint64_t r_lo, r_hi;
            mrange(v, nb_v, r_lo, r_hi);
            const int kk = __popcll(b_lo & r_lo) + __popcll(b_hi & r_hi);
            dst[q] = mkth_low(f_lo, f_hi, kk);
          } else {
            dst[q] = v - 1;
          }
        }
      }
      wsync();
#pragma unroll
      for (int q = 0; q < VS; ++q)
        if (dst[q] >= 0) s_ver[g][dst[q]] = val[q];
      wsync();
    }
    const int nv_old = nb_v;
    nb_v -= nr;
    (void)nv_old;

    // ---- cavity boundary (convex_cell.cu:618-678), serial per cell
    for (int p = j; p < P; p += L) s_bn[g][p] = kEnd;
    wsync();
    uint32_t first = kEnd;
    {
      int iter = 0, t = nb_v, r = nr;
      while (r > 0) {
        if (iter++ > 65535) {
          status = RPD_INCONSISTENT_BOUNDARY;
          break;
        }
        const uint32_t tw = s_ver[g][t];
        const uint32_t pa = tw & 255, pb = (tw >> 8) & 255,
                       pc = (tw >> 16) & 255;
        const uint32_t ba = s_bn[g][pa], bb = s_bn[g][pb], bc = s_bn[g][pc];
        const bool in0 = ba != kEnd, in1 = bb != kEnd, in2 = bc != kEnd;
        const bool n0 = bb == pa, n1 = bc == pb, n2 = ba == pc;
        bool simple = true;
        if (!n0 && !n1 && in1) simple = false;
        if (!n1 && !n2 && in2) simple = false;
        if (!n2 && !n0 && in0) simple = false;
        if (!n0 && !n1 && !n2) {
          if (first == kEnd) {
            wsync();
            s_bn[g][pa] = pb;
            s_bn[g][pb] = pc;
            s_bn[g][pc] = pa;
            first = pa;
            wsync();
          } else {
            simple = false;
          }
        }
        if (!simple) {
          ++t;
          if (t == nb_v + r) t = nb_v;
          continue;
        }
        wsync();
        if (!n0) s_bn[g][pa] = pb;
        if (!n1) s_bn[g][pb] = pc;
        if (!n2) s_bn[g][pc] = pa;
        wsync();
        if (n0 && n1) {
          if (first == pb) first = s_bn[g][pb];
          wsync();
          s_bn[g][pb] = kEnd;
          wsync();
        }
        if (n1 && n2) {
          if (first == pc) first = s_bn[g][pc];
          wsync();
          s_bn[g][pc] = kEnd;
          wsync();
        }
        if (n2 && n0) {
          if (first == pa) first = s_bn[g][pa];
          wsync();
          s_bn[g][pa] = kEnd;
          wsync();
        }
        const int last = nb_v + r - 1;
        const uint32_t lw = s_ver[g][last];
        wsync();
        s_ver[g][t] = lw;
        s_ver[g][last] = tw;
        wsync();
        t = nb_v;
        --r;
      }
    }
    if (status != RPD_SUCCESS) break;
    if (first == kEnd) {
      if (kFull && nb_p >= RPD_MAX_P && k + 1 < deg) {
        status = RPD_VERTEX_OVERFLOW;
        break;
      }
      continue;
    }

    // ---- boundary cycle
    int nbd = 0;
    {
      uint32_t cir = first;
      do {
        if (nbd >= P || cir >= (uint32_t)P) break;
        if (j == 0) s_cyc[g][nbd] = (uint8_t)cir;
        ++nbd;
        cir = s_bn[g][cir];
      } while (cir != first);
      if (cir != first) {
        status = RPD_INCONSISTENT_BOUNDARY;
        break;
      }
    }
    wsync();
    // ---- new edges (cur, cir) (convex_cell.cu:604-616, 752-758)
    if (nb_e + nbd > E) {
      status = kFull ? RPD_EDGE_OVERFLOW : kStatusRerunLarge;
      break;
    }
    const uint32_t hcur = s_h[g][cur];
    uint32_t adj_lo = 0, adj_hi = 0;  // the planes now joined to cur
    for (int q = j; q < nbd; q += L) {
      const uint32_t c2 = s_cyc[g][q];
      const uint32_t hz = max(hcur, (uint32_t)s_h[g][c2]);
      const uint32_t lo = min((uint32_t)cur, c2), hi = max((uint32_t)cur, c2);
      wedge[nb_e + q] = lo | (hi << 8) | (hz << 16);
      if (P <= 32 || c2 < 32)
        adj_lo |= 1u << (c2 & 31);
      else
        adj_hi |= 1u << (c2 - 32);
    }
    adj_lo = group_or<L>(adj_lo);
    if constexpr (P > 32) adj_hi = group_or<L>(adj_hi);
    if (j == 0) s_adj[g][cur] = (AdjT)(((uint64_t)adj_hi << 32) | adj_lo);
    wsync();
    // ---- new vertices (cur, cir, next(cir)) (convex_cell.cu:760-772)
    int first_bad = 1 << 30;   // first index failing overflow or perturb
    int bad_kind = 0;          // 1 overflow, 2 perturb, 3 edge lookup, 4 undecided
    for (int q0 = 0; q0 < nbd; q0 += L) {
      const int q = q0 + j;
      int kind = 0;
      if (q < nbd) {
        const uint32_t c1 = s_cyc[g][q];
        const uint32_t c2 = s_cyc[g][(q + 1) % nbd];
        if (nb_v + q + 1 >= V) {
          kind = 1;
        } else {
          // adjacency = max over the three edges (first match in list)
          // the new edges (cur, c1), (cur, c2) carry max(h_cur, h_c)
          const uint32_t z1 = max(hcur, (uint32_t)s_h[g][c1]);
          const uint32_t z2 = max(hcur, (uint32_t)s_h[g][c2]);
          const uint32_t lo = min(c1, c2), hi = max(c1, c2);
          // the edge (lo, hi), without a scan of the list: a pair of tet
          // faces is one of the six initial records (list positions 0-5,
          // pairs (2,3), (1,3), (1,2), (0,3), (0,2), (0,1)); a pair with a
          // bisector was appended once, when its later plane was clipped,
          // with z = max of the two planes' adjacency bytes
          const bool found = (s_adj[g][hi] >> lo) & 1;
          uint32_t z3 = 0;
          if (found) {
            if (hi < 4)
              z3 = meta_byte(tm, 8 + 5 - (lo * (7 - lo) / 2 + (hi - lo - 1)));
            else
              z3 = max((uint32_t)s_h[g][lo], (uint32_t)s_h[g][hi]);
          }
          if (!found) {
            kind = 3;
          } else {
            const rpd_d4 pc = ld4(&s_pl[g][cur]);
            const rpd_d4 p1 = ld4(&s_pl[g][c1]);
            const rpd_d4 p2 = ld4(&s_pl[g][c2]);
            int pf = a.force_exact ? -1 : perturb_filtered(pc, p1, p2);
            if (kExact && pf < 0) pf = perturb_exact(pc, p1, p2);
            if (pf == 1)
              kind = 2;
            else if (pf < 0)
              kind = 4;
            const uint32_t z = max(max(z1, z2), z3);
            s_ver[g][nb_v + q] = cur | (c1 << 8) | (c2 << 16) | (z << 24);
          }
        }
      }
      // first failing index in this pass (in order), kinds by priority of
      // position; overflow at q stops before the perturb test of q
      const uint64_t bad = group_bits<L>(ballot(kind != 0), g);
      if (bad && first_bad == (1 << 30)) {
        const int qb = q0 + __ffsll((long long)bad) - 1;
        first_bad = qb;
        // broadcast the kind of lane (qb - q0) of this group
        bad_kind = __shfl(kind, g * L + (qb - q0));
      }
    }
    wsync();
    if (first_bad != (1 << 30)) {
      if (bad_kind == 1)
        status = kFull ? RPD_TRIANGLE_OVERFLOW : kStatusRerunLarge;
      else if (bad_kind == 2)
        status = RPD_NEEDS_PERTURB;
      else if (bad_kind == 3)
        status = RPD_INCONSISTENT_BOUNDARY;
      else
        status = kStatusRerunExact;
      break;
    }
    nb_e += nbd;
    nb_v += nbd;
    if (kFull && nb_p >= RPD_MAX_P && k + 1 < deg) {
      status = RPD_VERTEX_OVERFLOW;
      break;
    }
  }

  // ---- write the slot
  const SlotView& o = a.out;
  if (status == RPD_SUCCESS && a.geometry) {
    // (tolerance-checked arithmetic: products and sums may fuse)
#pragma clang fp contract(fast)
    // FP64 geometry in one pass over the vertices.  Around each vertex A
    // lie 6 tets (C, X, Y, A) whose bases (X, Y, A) tile A's corner of its
    // three faces: X, Y run over the feet of C on A's 3 planes and on its 3
    // edge lines (the decomposition of the reference's
    // get_tet_decomposition_of_vertex, convex_cell.cu:986-1006).  Their
    // volumes give the cell volume and centroid, their bases the facet
    // areas (signed along each plane normal, so feet outside a face
    // cancel).  C is the cell's vertex 0, so rounding scales with the cell.
    // A base (X, Y, A) lies in A's plane p, so six times the signed volume
    // of (X, Y, C, A) is -area2(X, Y, A; n_p) * s_p, with s_p the foot
    // parameter of C on p (C - F_p = s_p n_p): one product per tet.
    double Cc[3];
    {
      const uint32_t w0 = s_ver[g][0];
      vertex_pos(ld4(&s_pl[g][w0 & 255]), ld4(&s_pl[g][(w0 >> 8) & 255]),
                 ld4(&s_pl[g][(w0 >> 16) & 255]), Cc);
    }
    constexpr int PA = (P + L - 1) / L;  // planes per lane
    double acc[PA];
#pragma unroll
    for (int q = 0; q < PA; ++q) acc[q] = 0.0;
    double vol6 = 0, mx = 0, my = 0, mz = 0;
    for (int q0 = 0; q0 < nb_v; q0 += L) {
      const int v = q0 + j;
      uint32_t ia = kEnd, ib = kEnd, ic = kEnd;
      double ca = 0, cb = 0, cc = 0;
      if (v < nb_v) {
        const uint32_t w = s_ver[g][v];
        ia = w & 255;
        ib = (w >> 8) & 255;
        ic = (w >> 16) & 255;
        const rpd_d4 pa = ld4(&s_pl[g][ia]);
        const rpd_d4 pb = ld4(&s_pl[g][ib]);
        const rpd_d4 pc = ld4(&s_pl[g][ic]);
        double A[3];
        vertex_pos(pa, pb, pc, A);
        // walk the six (foot, foot) pairs around A so that each foot dies
        // after its two tets and base triangles
        double Fa[3], Xp[3], Xn[3];
        const double sa = foot_plane(
            pa, rcp64(pa.x * pa.x + pa.y * pa.y + pa.z * pa.z), Cc, Fa);
#define RPD_ACC(P0, P1, PL, SP, CONTRIB)                         \
  {                                                              \
    const double ar2 = tri_area2(P0, P1, A, PL);                 \
    const double tvl = -ar2 * SP;                                \
    vol6 += tvl;                                                 \
    mx += tvl * (((P0[0] + P1[0]) + Cc[0]) + A[0]);              \
    my += tvl * (((P0[1] + P1[1]) + Cc[1]) + A[1]);              \
    mz += tvl * (((P0[2] + P1[2]) + Cc[2]) + A[2]);              \
    CONTRIB += ar2;                                              \
  }
        foot_line(pa, pb, A, Cc, Xp);                               // Eab
        RPD_ACC(Fa, Xp, pa, sa, ca)
        const double sb = foot_plane(
            pb, rcp64(pb.x * pb.x + pb.y * pb.y + pb.z * pb.z), Cc, Xn);
        RPD_ACC(Xp, Xn, pb, sb, cb)                                 // Eab, Fb
        foot_line(pb, pc, A, Cc, Xp);                               // Ebc
        RPD_ACC(Xn, Xp, pb, sb, cb)                                 // Fb, Ebc
        const double sc = foot_plane(
            pc, rcp64(pc.x * pc.x + pc.y * pc.y + pc.z * pc.z), Cc, Xn);
        RPD_ACC(Xp, Xn, pc, sc, cc)                                 // Ebc, Fc
        foot_line(pc, pa, A, Cc, Xp);                               // Eca
        RPD_ACC(Xn, Xp, pc, sc, cc)                                 // Fc, Eca
        RPD_ACC(Xp, Fa, pa, sa, ca)                                 // Eca, Fa
#undef RPD_ACC
      }
      // per-plane sums over the group's lanes (fixed butterfly order);
      // lane p % L keeps plane p
#pragma unroll
      for (int pi = 0; pi < PA; ++pi) {
#pragma unroll
        for (int pj = 0; pj < L; ++pj) {
          const uint32_t p = (uint32_t)(pi * L + pj);
          if (p < (uint32_t)P && p < (uint32_t)nb_p) {
            double sp = (ia == p ? ca : 0.0) + (ib == p ? cb : 0.0) +
                        (ic == p ? cc : 0.0);
            sp = group_sum<L>(sp);
            if (j == pj) acc[pi] += sp;
          }
        }
      }
    }
    vol6 = group_sum<L>(vol6);
    mx = group_sum<L>(mx);
    my = group_sum<L>(my);
    mz = group_sum<L>(mz);
    if (j == 0) {
      double* gm = o.geom + 4 * slot;
      const double vol = vol6 / 6.0;
      const double inv = vol6 != 0 ? 1.0 / (4.0 * vol6) : 0.0;
      gm[0] = fabs(vol);
      gm[1] = mx * inv;
      gm[2] = my * inv;
      gm[3] = mz * inv;
    }
#pragma unroll
    for (int pi = 0; pi < PA; ++pi) {
      const int p = pi * L + j;
      if (p < nb_p) {
        const rpd_d4 pp = ld4(&s_pl[g][p]);
        const double nn = pp.x * pp.x + pp.y * pp.y + pp.z * pp.z;
        o.areas[slot * o.stride_areas + p] =
            fabs(acc[pi]) * 0.5 / sqrt(nn);
      }
    }
  }
  if (status == RPD_SUCCESS) {
    // vertices and plane sources right after the edges (in place)
    for (int v = j; v < nb_v; v += L) wslot[nb_e + v] = s_ver[g][v];
    for (int p = j; p < nb_p; p += L)
      wslot[nb_e + nb_v + p] = (uint32_t)s_src[g][p];
  }
  if constexpr (kExact) {
    // a vertex exactly on a plane: the configuration is degenerate and a
    // sphere whose cell meets the tet only in a face, an edge or a point
    // (a zero-volume cell, which the plain path emits) may touch this cell
    // without leaving an active facet here; the certificate cannot see it,
    // so the tet goes to the full relation (the exact tier runs one cell
    // per wavefront, so the ballot covers exactly this cell's lanes)
    if (kCert && ballot(on_plane) != 0 && j == 0) a.cert.flag[t] = 1;
  }
  if (kCert && status != kStatusRerunLarge && status != kStatusRerunExact &&
      status != RPD_NO_INTERSECTION) {
    const CertArgs& c = a.cert;
    if (status != RPD_SUCCESS) {
      if (j == 0) c.flag[t] = 1;
    } else {
      if (j == 0) c.has_cell[t] = 1;
      // planes with a vertex (the active facets), OR over the group
      uint64_t act0 = 0, act1 = 0;
      for (int v = j; v < nb_v; v += L) {
        const uint32_t w = s_ver[g][v];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const uint32_t p = (w >> (8 * q)) & 255u;
          if (p < 64)
            act0 |= 1ull << p;
          else
            act1 |= 1ull << (p - 64);
        }
      }
#pragma unroll
      for (int off = L / 2; off > 0; off >>= 1) {
        act0 |= (uint64_t)__shfl_xor((unsigned long long)act0, off);
        act1 |= (uint64_t)__shfl_xor((unsigned long long)act1, off);
      }
      const int64_t o0 = c.tet_off[t], o1 = c.tet_off[t + 1];
      for (int p = j; p < nb_p; p += L) {
        const bool active =
            p < 64 ? ((act0 >> p) & 1) : ((act1 >> (p - 64)) & 1);
        const int32_t src = s_src[g][p];
        if (!active || src < 0) continue;
        int64_t lo = o0, hi = o1;
        while (lo < hi) {
          const int64_t mid = (lo + hi) >> 1;
          if (c.mem_ids[mid] < src)
            lo = mid + 1;
          else
            hi = mid;
        }
        if (lo < o1 && c.mem_ids[lo] == src) continue;
        const uint64_t key = ((uint64_t)(uint32_t)t << 32) | (uint32_t)src;
        if (c.n_ext && has_key(c.ext, c.n_ext, key)) continue;
        const int r = atomicAdd(c.n_miss, 1);
        const int64_t cap =
            c.cap_n ? min((int64_t)c.cap, max((int64_t)4096, 2 * *c.cap_n))
                    : (int64_t)c.cap;
        if (r < cap)
          c.miss[r] = key;
        else
          c.flag[t] = 1;
      }
    }
  }
  if (a.work) {
    if (kExact) {  // exact determinants are per lane: sum over the group
#pragma unroll
      for (int d = 1; d < L; d <<= 1) n_exd += __shfl_xor(n_exd, d, 64);
    }
    if (j == 0) {
      unsigned long long* w = a.work + 4 * (blockIdx.x & 63);
      atomicAdd(w + 0, (unsigned long long)n_det);
      atomicAdd(w + 1, (unsigned long long)n_bis);
      if (status == RPD_SUCCESS) atomicAdd(w + 2, (unsigned long long)nb_v);
      if (n_exd) atomicAdd(w + 3, (unsigned long long)n_exd);
    }
  }
  if (j == 0) {
    o.hdr[slot] = (uint32_t)nb_v | ((uint32_t)nb_p << 8) |
                  ((uint32_t)nb_e << 16) | (status << 24);
    if (status == kStatusRerunLarge) {
      int r = atomicAdd(a.n_rerun_large, 1);
      if (r < a.cap_large)
        a.rerun_large[r] = (int32_t)cidx;
      else
        a.overflow[0] = 1;
    } else if (status == kStatusRerunExact) {
      int r = atomicAdd(a.n_rerun_exact, 1);
      if (r < a.cap_exact)
        a.rerun_exact[r] = (int32_t)cidx;
      else
        a.overflow[0] = 1;
    }
  }
}

// ========================================================== compaction
// per cell: location of its final slot; tier 0 small (index = cell),
// tier 1 large / 2 exact (index = re-run position)
__global__ __launch_bounds__(256) void k_mark_rerun(const int32_t* list,
                                                    int32_t r0, int32_t r1,
                                                    int32_t tier,
                                                    int32_t* loc,
                                                    const int32_t* r1_dev) {
  const int r = r0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (r1_dev ? min(r1, *r1_dev) : r1)) return;
  loc[list[r]] = (tier << 28) | r;
}

struct Sizes3 {
  int64_t cells, words, areas;
};
struct Sizes3Sum {
  __host__ __device__ Sizes3 operator()(const Sizes3& a,
                                        const Sizes3& b) const {
    return Sizes3{a.cells + b.cells, a.words + b.words, a.areas + b.areas};
  }
};

struct TierViews {
  SlotView t[4];  // small, full capacity, exact, middle
};

// Compaction in blocks of kCB consecutive output positions r (cell
// perm[r], or r when perm is null).  A cell's sizes pack into one u64
// (cells in bits 0-15, areas 16-35, words 36-63), so one integer sum
// scans all three: a block holds at most 256 cells, 256 x 64 areas and
// 256 x 312 words.
constexpr int kCB = 256;

struct CellRef {
  int tier;
  int64_t idx;
  uint32_t hdr;
};

__device__ __forceinline__ CellRef cell_ref(int64_t r, const int32_t* perm,
                                            const int32_t* loc,
                                            const TierViews& tv) {
  const int64_t c = perm ? perm[r] : r;
  const int32_t l = loc[c];
  CellRef x;
  x.tier = l >> 28;
  x.idx = x.tier ? (l & ((1 << 28) - 1)) : c;
  x.hdr = tv.t[x.tier].hdr[x.idx];
  return x;
}

__device__ __forceinline__ uint64_t packed_size(uint32_t hd) {
  if ((hd >> 24) != RPD_SUCCESS) return 0;
  const uint64_t nv = hd & 255, np = (hd >> 8) & 255, ne = (hd >> 16) & 255;
  return 1ull | (np << 16) | ((nv + np + ne) << 36);
}

// per block: totals of its cells (bsum[b]) and the dropped cells by status
// per-status counts of the cells not emitted, spread over kDropSlots copies
// of the 9 counters (a block adds to copy blockIdx % kDropSlots) so that
// the blocks' atomics do not all meet at one address
constexpr int kDropSlots = 64;
__global__ __launch_bounds__(kCB) void k_block_sizes(int64_t n_cells,
                                                     const int32_t* perm,
                                                     const int32_t* loc,
                                                     TierViews tv,
                                                     Sizes3* bsum,
                                                     int64_t* drops,
                                                     const int64_t* n_dev) {
  typedef hipcub::BlockReduce<uint64_t, kCB> Red;
  n_cells = dev_count(n_cells, n_dev);
  __shared__ typename Red::TempStorage tmp;
  __shared__ int32_t h[9];
  if (threadIdx.x < 9) h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * kCB + threadIdx.x;
  uint64_t z = 0;
  if (r < n_cells) {
    const CellRef x = cell_ref(r, perm, loc, tv);
    const uint32_t st = x.hdr >> 24;
    z = packed_size(x.hdr);
    if (st != RPD_SUCCESS && st < 9) atomicAdd(&h[st], 1);
  }
  const uint64_t tot = Red(tmp).Sum(z);
  if (threadIdx.x == 0)
    bsum[blockIdx.x] = Sizes3{(int64_t)(tot & 0xffff),
                              (int64_t)(tot >> 36),
                              (int64_t)((tot >> 16) & 0xfffff)};
  __syncthreads();
  if (threadIdx.x < 9 && h[threadIdx.x])
    atomicAdd((unsigned long long*)&drops[9 * (blockIdx.x % kDropSlots) +
                                          threadIdx.x],
              (unsigned long long)h[threadIdx.x]);
}

// block b writes its cells from boff[b] on: cell records and geometry one
// per thread, payload words and areas by the whole block over the block's
// contiguous output range (each output element finds its cell by binary
// search in the block's offsets), so every store is coalesced.
__global__ __launch_bounds__(kCB) void k_compact(
    int64_t n_cells, const int32_t* perm, const Cand* cand,
    const int32_t* loc, TierViews tv, const Sizes3* boff, rpd_cell* out_cells,
    uint32_t* out_words, double* out_geom, double* out_areas,
    int32_t geometry, const int64_t* n_dev, Sizes3 cap, int32_t* overflow) {
  typedef hipcub::BlockScan<uint64_t, kCB> Scan;
  __shared__ typename Scan::TempStorage tmp;
  __shared__ uint32_t s_w[kCB + 1];
  __shared__ uint32_t s_a[kCB + 1];
  __shared__ const uint32_t* s_wsrc[kCB];
  __shared__ const double* s_asrc[kCB];
  // a slot holds the edges at 0 and the vertices and plane sources after
  // them; the stream holds vertices, plane sources, edges
  __shared__ uint32_t s_lay[kCB];  // nb_v | nb_p << 8 | nb_e << 16
  const int i = threadIdx.x;
  const int64_t r = (int64_t)blockIdx.x * kCB + i;
  n_cells = dev_count(n_cells, n_dev);
  {  // a captured pipeline's output capacity: the block's whole range
    const Sizes3 B0 = boff[blockIdx.x], B1 = boff[blockIdx.x + 1];
    if (B1.cells > cap.cells || B1.words > cap.words ||
        B1.areas > cap.areas) {
      if (i == 0) overflow[0] = 1;
      return;
    }
    (void)B0;
  }
  CellRef x{0, 0, 0};
  int64_t c = 0;
  uint64_t z = 0;
  if (r < n_cells) {
    c = perm ? perm[r] : r;
    x = cell_ref(r, perm, loc, tv);
    z = packed_size(x.hdr);
  }
  uint64_t off, tot;
  Scan(tmp).ExclusiveSum(z, off, tot);
  const Sizes3 B = boff[blockIdx.x];
  const uint32_t w_off = (uint32_t)(off >> 36);
  const uint32_t a_off = (uint32_t)((off >> 16) & 0xfffff);
  const int64_t c_off = B.cells + (int64_t)(off & 0xffff);
  s_w[i] = w_off;
  s_a[i] = a_off;
  if (i == 0) {
    s_w[kCB] = (uint32_t)(tot >> 36);
    s_a[kCB] = (uint32_t)((tot >> 16) & 0xfffff);
  }
  const SlotView& sv = tv.t[x.tier];
  s_wsrc[i] = sv.words + x.idx * sv.stride_words;
  s_asrc[i] = sv.areas + x.idx * sv.stride_areas;
  s_lay[i] = x.hdr & 0xffffffu;
  if (z) {
    rpd_cell rc;
    rc.tet_id = cand[c].tet;
    rc.seed = cand[c].seed;
    rc.nb_v = (uint8_t)(x.hdr & 255);
    rc.nb_p = (uint8_t)((x.hdr >> 8) & 255);
    rc.nb_e = (uint8_t)((x.hdr >> 16) & 255);
    rc.status = RPD_SUCCESS;
    rc.reserved = 0;
    rc.word_off = (uint64_t)(B.words + w_off);
    rc.area_off = (uint64_t)(B.areas + a_off);
    out_cells[c_off] = rc;
    if (geometry) {
      const double* gs = sv.geom + 4 * x.idx;
      double* gd = out_geom + 4 * c_off;
      gd[0] = gs[0];
      gd[1] = gs[1];
      gd[2] = gs[2];
      gd[3] = gs[3];
    }
  }
  __syncthreads();
  // the cell holding element e: the last k with s[k] <= e (cells without
  // output have s[k] == s[k + 1] and are skipped by "last")
  auto owner = [](const uint32_t* sa, uint32_t e) {
    int lo = 0, hi = kCB;  // sa[lo] <= e < sa[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (sa[mid] <= e)
        lo = mid;
      else
        hi = mid;
    }
    return lo;
  };
  // Each wavefront copies a contiguous quarter of the block's elements, 64
  // consecutive ones per step (coalesced stores); a lane finds the cell of
  // its first element by binary search and then walks forward, past about
  // 64 / (elements per cell) cells a step.  kU steps at a time: their
  // sources first, then their loads together, then their stores.
  constexpr int kU = 4;
  const int wv = i >> 6, ln = i & 63;
  auto range = [&](uint32_t n, uint32_t* q0, uint32_t* q1) {
    *q0 = (uint32_t)((uint64_t)n * wv / (kCB / 64));
    *q1 = (uint32_t)((uint64_t)n * (wv + 1) / (kCB / 64));
  };
  const uint32_t W = s_w[kCB];
  uint32_t* wd = out_words + B.words;
  {
    uint32_t q0, q1;
    range(W, &q0, &q1);
    int k = q0 + ln < q1 ? owner(s_w, q0 + ln) : 0;
    for (uint32_t e0 = q0 + ln; e0 < q1; e0 += 64 * kU) {
      const uint32_t* src[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const uint32_t e = e0 + 64 * u;
        src[u] = nullptr;
        if (e < q1) {
          while (s_w[k + 1] <= e) ++k;
          const uint32_t w = e - s_w[k], lay = s_lay[k];
          const uint32_t nvp = (lay & 255) + ((lay >> 8) & 255);
          src[u] = s_wsrc[k] + (w < nvp ? (lay >> 16) + w : w - nvp);
        }
      }
      uint32_t v[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) v[u] = src[u] ? *src[u] : 0u;
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (src[u]) wd[e0 + 64 * u] = v[u];
    }
  }
  if (geometry) {
    const uint32_t A = s_a[kCB];
    double* ad = out_areas + B.areas;
    uint32_t q0, q1;
    range(A, &q0, &q1);
    int k = q0 + ln < q1 ? owner(s_a, q0 + ln) : 0;
    for (uint32_t e0 = q0 + ln; e0 < q1; e0 += 64 * kU) {
      const double* src[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const uint32_t e = e0 + 64 * u;
        src[u] = nullptr;
        if (e < q1) {
          while (s_a[k + 1] <= e) ++k;
          src[u] = s_asrc[k] + (e - s_a[k]);
        }
      }
      double v[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) v[u] = src[u] ? *src[u] : 0.0;
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (src[u]) ad[e0 + 64 * u] = v[u];
    }
  }
}

// active-plane statistics of the candidate masks
__global__ __launch_bounds__(256) void k_mask_stats(int64_t n, const Cand* cand,
                                                    DevSites S,
                                                    unsigned long long* acc) {
  __shared__ unsigned long long s_a, s_b;
  if (threadIdx.x == 0) s_a = s_b = 0;
  __syncthreads();
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) {
    const Cand cd = cand[c];
    const int64_t deg = S.knn_off[cd.seed + 1] - S.knn_off[cd.seed];
    int64_t act = __popcll(cd.m0) + __popcll(cd.m1);
    if (deg > 128) act += deg - 128;
    atomicAdd(&s_a, (unsigned long long)act);
    atomicAdd(&s_b, (unsigned long long)deg);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&acc[0], s_a);
    atomicAdd(&acc[1], s_b);
  }
}

}  // namespace rpd
