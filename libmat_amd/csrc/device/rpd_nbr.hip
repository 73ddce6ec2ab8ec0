// Regular-triangulation neighbours of a sphere set on the device: the
// power cell of every sphere inside the domain box, clipped by the power
// bisectors of the spheres that can reach it, with exact signs (filtered
// dot product of the plane with the vertex's homogeneous coordinates, exact
// expansion fallback).  A sphere is valid (not hidden) iff its cell is not
// empty; its neighbours are the spheres whose bisector carries a face of
// the cell.  This is the restatement of the reference's CGAL regular
// triangulation step (triangulation.cxx:62-144, get_RT_neighbors_given_
// spheres :396-438) that the host builder rpd_power_neighbors computes; the
// per-sphere lists are symmetrised on the host (rpd_nbr_from_lists).
//
// Candidates come from weight-class grids (rpd_classgrid.h): class by
// class, shells of grid cells at Chebyshev distance r around the sphere's
// cell, until no sphere of the class can reach the cell: a sphere j cuts
// the cell of i only if d_ij < R + sqrt(max(0, R^2 - w_i + w_j)), R the
// cell's largest vertex distance from c_i.

namespace rpd {

constexpr int kNbP = 192;   // planes of one cell (6 box planes + bisectors)
constexpr int kNbV = 256;   // vertices of one cell
constexpr int kNbOut = 160;  // neighbours written per sphere

struct NbGrid {
  double org[3];
  double cs;
  double wmax;
  int32_t dim[3];
  int32_t used;
  int64_t cell_base;  // into start (cells + 1 entries)
};
struct NbGrids {
  NbGrid g[8];
};

// cell of Chebyshev shell r (k in [0, shell size)) around the origin
__device__ __forceinline__ void shell_cell(int r, int k, int& dx, int& dy,
                                           int& dz) {
  if (r == 0) {
    dx = dy = dz = 0;
    return;
  }
  const int side = 2 * r + 1, face = side * side;
  if (k < 2 * face) {  // the two z faces
    dz = k < face ? -r : r;
    const int q = k % face;
    dy = q / side - r;
    dx = q % side - r;
    return;
  }
  const int q = k - 2 * face;  // rings of the slabs between them
  const int per = 8 * r;
  dz = -r + 1 + q / per;
  const int p = q % per;
  if (p < side) {
    dy = -r;
    dx = -r + p;
  } else if (p < 2 * side) {
    dy = r;
    dx = -r + (p - side);
  } else {
    const int p2 = p - 2 * side;  // 0 .. 4r-3: the two columns
    dx = (p2 & 1) ? r : -r;
    dy = -r + 1 + (p2 >> 1);
  }
}

__device__ __forceinline__ int shell_size(int r) {
  return r == 0 ? 1 : 2 * (2 * r + 1) * (2 * r + 1) + (2 * r - 1) * 8 * r;
}

// sign of the bisector e at vertex (planes p1, p2, p3; homogeneous X, Q)
__device__ __forceinline__ int nb_side(const rpd_d4& p1, const rpd_d4& p2,
                                       const rpd_d4& p3, const rpd_d4& X,
                                       const rpd_d4& Q, const rpd_d4& e) {
  int s = rpd_filtered_sign4(e, X, Q);
  if (s == 0) s = rpd_exact_det4_sign(p1, p2, p3, e);
  int w = rpd_filtered_sign_w(X, Q);
  if (w == 0) w = rpd_exact_w_sign(p1, p2, p3);
  return s * w;
}

// One wavefront per sphere.  valid[i]: 1 valid, 0 hidden, -1 capacity or
// topology failure; counts[i] / lists[i * kNbOut ..]: neighbour ids.
__global__ __launch_bounds__(64) void k_power_nbrs(
    int32_t n, const double4* __restrict__ sph, const double* box, NbGrids G,
    const int64_t* __restrict__ start, const int32_t* __restrict__ ids,
    int32_t* valid, int32_t* counts, int32_t* lists) {
  __shared__ double4 s_pl[kNbP];
  __shared__ int32_t s_src[kNbP];
  __shared__ double4 s_X[2][kNbV];
  __shared__ double4 s_Q[2][kNbV];
  __shared__ uint32_t s_t[2][kNbV];
  __shared__ uint8_t s_out[kNbV];
  __shared__ uint8_t s_used[kNbP];
  __shared__ int16_t s_map[kNbP];
  __shared__ int32_t s_new;
  const int i = blockIdx.x;
  const int lane = threadIdx.x;
  if (i >= n) return;
  const double4 sd = sph[i];
  const rpd_d4 si{sd.x, sd.y, sd.z, sd.w};
  // the domain box: x >= x0, x <= x1, y >= y0, y <= y1, z >= z0, z <= z1
  if (lane < 6) {
    const int a = lane >> 1;
    double4 p = make_double4(0, 0, 0, 0);
    double* pp = &p.x;
    if ((lane & 1) == 0) {
      pp[a] = 1.0;
      p.w = -box[a];
    } else {
      pp[a] = -1.0;
      p.w = box[3 + a];
    }
    s_pl[lane] = p;
    s_src[lane] = -1 - lane;
  }
  wsync();
  if (lane < 8) {
    const int a = lane & 1, b = 2 + ((lane >> 1) & 1), c = 4 + ((lane >> 2) & 1);
    rpd_d4 X, Q;
    rpd_vertex_homog(ld4(&s_pl[a]), ld4(&s_pl[b]), ld4(&s_pl[c]), X, Q);
    s_X[0][lane] = make_double4(X.x, X.y, X.z, X.w);
    s_Q[0][lane] = make_double4(Q.x, Q.y, Q.z, Q.w);
    s_t[0][lane] = a | (b << 10) | (c << 20);
  }
  wsync();
  int nv = 8, np = 6, cur = 0;
  bool empty = false, fail = false;
  auto radius2 = [&]() {
    double r2 = 0;
    for (int v = lane; v < nv; v += 64) {
      const double4 X = s_X[cur][v];
      const double x = X.x / X.w - si.x, y = X.y / X.w - si.y,
                   z = X.z / X.w - si.z;
      r2 = fmax(r2, x * x + y * y + z * z);
    }
    for (int off = 32; off > 0; off >>= 1) r2 = fmax(r2, __shfl_xor(r2, off));
    return r2;
  };
  double R2 = radius2();

  // clip the cell by the bisector of sphere j (every lane calls it)
  auto clip = [&](int j) {
    const double4 s2 = sph[j];
    const rpd_d4 e = rpd_bisector(si, rpd_d4{s2.x, s2.y, s2.z, s2.w});
    int n_out = 0;
    for (int v0 = 0; v0 < nv; v0 += 64) {
      const int v = v0 + lane;
      bool o = false;
      if (v < nv) {
        const uint32_t t = s_t[cur][v];
        const double4 X = s_X[cur][v], Q = s_Q[cur][v];
        o = nb_side(ld4(&s_pl[t & 1023]), ld4(&s_pl[(t >> 10) & 1023]),
                    ld4(&s_pl[(t >> 20) & 1023]),
                    rpd_d4{X.x, X.y, X.z, X.w}, rpd_d4{Q.x, Q.y, Q.z, Q.w},
                    e) < 0;
        s_out[v] = o;
      }
      n_out += __popcll(ballot(o));
    }
    wsync();
    if (n_out == 0) return;
    if (n_out == nv) {
      empty = true;
      return;
    }
    if (np >= kNbP) {
      // planes no vertex uses any more are dropped and the rest renumbered
      // (box planes stay at 0..5)
      for (int q = lane; q < np; q += 64) s_used[q] = q < 6;
      wsync();
      for (int v = lane; v < nv; v += 64) {
        const uint32_t t = s_t[cur][v];
        s_used[t & 1023] = 1;
        s_used[(t >> 10) & 1023] = 1;
        s_used[(t >> 20) & 1023] = 1;
      }
      wsync();
      int kept = 0;
      for (int q0 = 0; q0 < np; q0 += 64) {
        const int q = q0 + lane;
        const bool u = q < np && s_used[q];
        const uint64_t bm = ballot(u);
        const int to = kept + __popcll(bm & ((1ull << lane) - 1));
        double4 pl = make_double4(0, 0, 0, 0);
        int32_t sr = 0;
        if (u) {
          pl = s_pl[q];
          sr = s_src[q];
        }
        if (q < np) s_map[q] = (int16_t)(u ? to : -1);
        wsync();
        if (u) {
          s_pl[to] = pl;
          s_src[to] = sr;
        }
        wsync();
        kept += __popcll(bm);
      }
      for (int v = lane; v < nv; v += 64) {
        const uint32_t t = s_t[cur][v];
        s_t[cur][v] = (uint32_t)s_map[t & 1023] |
                      ((uint32_t)s_map[(t >> 10) & 1023] << 10) |
                      ((uint32_t)s_map[(t >> 20) & 1023] << 20);
      }
      wsync();
      np = kept;
      if (np >= kNbP) {
        fail = true;
        return;
      }
    }
    const int p = np;
    if (lane == 0) {
      s_pl[p] = make_double4(e.x, e.y, e.z, e.w);
      s_src[p] = j;
      s_new = 0;
    }
    wsync();
    const int nxt = cur ^ 1;
    const int n_keep = nv - n_out;
    int base = 0;
    for (int v0 = 0; v0 < nv; v0 += 64) {
      const int v = v0 + lane;
      const bool keep = v < nv && !s_out[v];
      const uint64_t bm = ballot(keep);
      if (keep) {
        const int pos = base + __popcll(bm & ((1ull << lane) - 1));
        s_X[nxt][pos] = s_X[cur][v];
        s_Q[nxt][pos] = s_Q[cur][v];
        s_t[nxt][pos] = s_t[cur][v];
      }
      base += __popcll(bm);
    }
    // new vertices where an edge joins a removed and a kept vertex: each
    // removed vertex looks across its three edges
    bool lfail = false;
    for (int v = lane; v < nv; v += 64) {
      if (!s_out[v]) continue;
      const uint32_t tv = s_t[cur][v];
      const int pv[3] = {(int)(tv & 1023), (int)((tv >> 10) & 1023),
                         (int)((tv >> 20) & 1023)};
      for (int q = 0; q < 3; ++q) {
        const int a = pv[q], b = pv[(q + 1) % 3];
        int other = -1;
        for (int u = 0; u < nv; ++u) {
          if (u == v) continue;
          const uint32_t tu = s_t[cur][u];
          const int u0 = tu & 1023, u1 = (tu >> 10) & 1023,
                    u2 = (tu >> 20) & 1023;
          const bool ha = u0 == a || u1 == a || u2 == a;
          const bool hb = u0 == b || u1 == b || u2 == b;
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
        const int slot = atomicAdd(&s_new, 1);
        if (n_keep + slot >= kNbV) {
          lfail = true;
          continue;
        }
        rpd_d4 X, Q;
        rpd_vertex_homog(ld4(&s_pl[a]), ld4(&s_pl[b]), ld4(&s_pl[p]), X, Q);
        s_X[nxt][n_keep + slot] = make_double4(X.x, X.y, X.z, X.w);
        s_Q[nxt][n_keep + slot] = make_double4(Q.x, Q.y, Q.z, Q.w);
        s_t[nxt][n_keep + slot] = a | (b << 10) | (p << 20);
      }
    }
    fail = fail || (ballot(lfail) != 0);
    wsync();
    nv = n_keep + min(s_new, kNbV - n_keep);
    np = p + 1;
    cur = nxt;
    wsync();
    R2 = radius2();
  };

  for (int c = 0; c < 8 && !empty && !fail; ++c) {
    const NbGrid& g = G.g[c];
    if (!g.used) continue;
    int64_t q0[3], rmin = 0, rmax = 0;
    for (int a = 0; a < 3; ++a) {
      q0[a] = (int64_t)floor(((&si.x)[a] - g.org[a]) / g.cs);
      const int64_t lo = q0[a], hi = q0[a] - (g.dim[a] - 1);
      const int64_t d1 = lo < 0 ? -lo : lo, d2 = hi < 0 ? -hi : hi;
      if (d1 > rmax) rmax = d1;
      if (d2 > rmax) rmax = d2;
      if (-lo > rmin) rmin = -lo;  // rings before the grid: none
      if (hi > rmin) rmin = hi;
    }
    for (int64_t r = rmin; r <= rmax && !empty && !fail; ++r) {
      {
        const double R = sqrt(R2) * (1 + 1e-9) + 1e-12;
        const double bound = R + sqrt(fmax(0.0, R * R - si.w + g.wmax));
        if (r >= 1 && (double)(r - 1) * g.cs > bound * (1 + 1e-9) + 1e-12)
          break;
      }
      const int S = shell_size((int)r);
      for (int k0 = 0; k0 < S && !empty && !fail; k0 += 64) {
        // this lane's cell of the shell and its member range
        int64_t a0 = 0, a1 = 0;
        const int k = k0 + lane;
        if (k < S) {
          int dx, dy, dz;
          shell_cell((int)r, k, dx, dy, dz);
          const int64_t x = q0[0] + dx, y = q0[1] + dy, z = q0[2] + dz;
          if (x >= 0 && y >= 0 && z >= 0 && x < g.dim[0] && y < g.dim[1] &&
              z < g.dim[2]) {
            const int64_t cell =
                g.cell_base + x + (int64_t)g.dim[0] * (y + (int64_t)g.dim[1] * z);
            a0 = start[cell];
            a1 = start[cell + 1];
          }
        }
        int len = (int)(a1 - a0);
        for (int off = 32; off > 0; off >>= 1) len = max(len, __shfl_xor(len, off));
        for (int m = 0; m < len && !empty && !fail; ++m) {
          // each lane offers its m-th member if it may reach the cell
          int j = -1;
          if (a0 + m < a1) {
            j = ids[a0 + m];
            if (j == i) {
              j = -1;
            } else {
              const double4 sj = sph[j];
              const double ddx = sj.x - si.x, ddy = sj.y - si.y,
                           ddz = sj.z - si.z;
              const double d = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
              const double Rc = sqrt(R2) * (1 + 1e-9) + 1e-12;
              const double bj = Rc + sqrt(fmax(0.0, Rc * Rc - si.w + sj.w));
              if (d > bj * (1 + 1e-9) + 1e-12) j = -1;
            }
          }
          uint64_t todo = ballot(j >= 0);
          while (todo && !empty && !fail) {
            const int src = __ffsll((long long)todo) - 1;
            todo &= todo - 1;
            clip(__shfl(j, src));
          }
        }
      }
    }
  }
  if (fail || empty || nv == 0) {
    if (lane == 0) {
      valid[i] = fail ? -1 : 0;
      counts[i] = 0;
    }
    return;
  }
  // neighbours: the bisectors that carry a vertex of the cell
  for (int p = lane; p < np; p += 64) s_used[p] = 0;
  wsync();
  for (int v = lane; v < nv; v += 64) {
    const uint32_t t = s_t[cur][v];
    s_used[t & 1023] = 1;
    s_used[(t >> 10) & 1023] = 1;
    s_used[(t >> 20) & 1023] = 1;
  }
  wsync();
  int cnt = 0;
  for (int p0 = 6; p0 < np; p0 += 64) {
    const int p = p0 + lane;
    const bool u = p < np && s_used[p];
    const uint64_t bm = ballot(u);
    if (u) {
      const int pos = cnt + __popcll(bm & ((1ull << lane) - 1));
      if (pos < kNbOut) lists[(int64_t)i * kNbOut + pos] = s_src[p];
    }
    cnt += __popcll(bm);
  }
  if (lane == 0) {
    valid[i] = cnt > kNbOut ? -1 : 1;
    counts[i] = min(cnt, kNbOut);
  }
}

}  // namespace rpd
