// Kernels over the emitted cells of a run (included by rpd_engine.hip).
//
// Per-site centroids: the Lloyd update of compute_new_site
// (src/rpd3d/voronoi.cu:40-69) over the barycentre sums that the
// reference's clip kernel accumulates per seed with float atomics
// (convex_cell.cu:1009-1054).  Here the sums are FP64 and in a fixed order:
// the cells are sorted by seed (a stable radix sort keeps each seed's cells
// in emission order), and each site's cells are summed by one wavefront,
// lane l taking entries l, l + 64, ... in order, then a fixed butterfly.

namespace rpd {

// keys = seed, values = cell index
__global__ __launch_bounds__(256) void k_cell_seed_keys(int64_t n,
                                                        const rpd_cell* cells,
                                                        uint32_t* keys,
                                                        int32_t* vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = (uint32_t)cells[i].seed;
  vals[i] = (int32_t)i;
}

// start[s] = first position of seed s in the sorted keys (start[n_site] = n)
__global__ __launch_bounds__(256) void k_seed_starts(int32_t n_site, int64_t n,
                                                     const uint32_t* keys,
                                                     int64_t* start) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s > n_site) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)keys[mid] < s)
      lo = mid + 1;
    else
      hi = mid;
  }
  start[s] = lo;
}

// one wavefront per site: out[4 s ..] = (sum vol * centroid) / sum vol,
// sum vol; a site without volume keeps its centre (compute_new_site
// updates only sites with cell_vol != 0)
__global__ __launch_bounds__(256) void k_site_centroid(
    int32_t n_site, const int64_t* start, const int32_t* idx,
    const double* geom, const double4* site, double* out) {
  const int lane = threadIdx.x & 63;
  const int64_t s = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (s >= n_site) return;
  const int64_t b = start[s], e = start[s + 1];
  double v = 0, x = 0, y = 0, z = 0;
  for (int64_t i = b + lane; i < e; i += 64) {
    const double* g = geom + 4 * (int64_t)idx[i];
    const double w = g[0];
    v += w;
    x += w * g[1];
    y += w * g[2];
    z += w * g[3];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v += __shfl_xor(v, off, 64);
    x += __shfl_xor(x, off, 64);
    y += __shfl_xor(y, off, 64);
    z += __shfl_xor(z, off, 64);
  }
  if (lane == 0) {
    const double4 c = site[s];
    double* o = out + 4 * s;
    if (v != 0) {
      o[0] = x / v;
      o[1] = y / v;
      o[2] = z / v;
    } else {
      o[0] = c.x;
      o[1] = c.y;
      o[2] = c.z;
    }
    o[3] = v;
  }
}


// ============================================================ adjacency
// The per-sphere power-cell adjacency of get_all_voro_info
// (src/rpd3d_base/rpd_update.cxx:80-340) and the connected components of
// update_pc_cc_info / update_pc_facet_cc_info / update_pc_edge_cc_info
// (rpd_update.cxx:409-521), over the cells of the last run, as sorted key
// streams instead of per-sphere maps:
//   tet-face records  (seed, global tet face id) -> (cell, local face):
//                     two records with one key = two cells of a sphere
//                     across a tet face (cell_neighbors)
//   facet records     (seed, neighbour site) -> cell (facet_neigh_to_cells)
//   edge records      (seed, neighbour pair) -> cell (e_to_cells)
// Components by union-find on the device: a root is only ever hooked under
// a smaller root, so every component ends rooted at its smallest element
// whatever the order of the hooks.

constexpr int kAdjPerBlock = 4;  // cells (one wavefront each) per block

// Per cell: the active planes (planes carrying a vertex) and active edges
// (both planes active and shared by at least two vertices) of
// ConvexCellHost::reload_active (voronoi_defs.cxx:76-106); the cell Euler
// sum of cal_cell_euler (voronoi_defs.cxx:190-221) with its rounding and
// summation order (binary32 vertex terms 1/#adjacent cells, binary64 face
// and edge terms, binary64 sums, binary32 result); and the number of
// records the cell adds of each kind.
__global__ __launch_bounds__(256) void k_adj_cell(
    int64_t n, const rpd_cell* cells, const uint32_t* words, DevMesh M,
    uint64_t* act, float* euler, int64_t* n_tf, int64_t* n_fac,
    int64_t* n_edg) {
  __shared__ unsigned long long s_vm[kAdjPerBlock][64][2];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t c = (int64_t)blockIdx.x * kAdjPerBlock + w;
  if (c >= n) return;
  const rpd_cell h = cells[c];
  const int nv = h.nb_v, np = h.nb_p, ne = h.nb_e;
  const uint32_t* W = words + h.word_off;
  s_vm[w][lane][0] = 0;
  s_vm[w][lane][1] = 0;
  wsync();
  // vertex sets of the planes (nb_v <= 96, nb_p <= 64)
  for (int v = lane; v < nv; v += 64) {
    const uint32_t x = W[v];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const uint32_t p = (x >> (8 * q)) & 255u;
      if (p < 64) atomicOr(&s_vm[w][p][v >> 6], 1ull << (v & 63));
    }
  }
  wsync();
  const bool pa =
      lane < np && (s_vm[w][lane][0] | s_vm[w][lane][1]) != 0;
  const uint64_t actp = ballot(pa);
  const int32_t src = lane < np ? (int32_t)W[nv + lane] : -1;
  const uint64_t bis = ballot(lane < np && src >= 0);
  uint64_t acte[3];
  int n_be = 0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int e = q * 64 + lane;
    bool a = false, ab = false;
    if (e < ne) {
      const uint32_t ew = W[nv + np + e];
      const uint32_t x = ew & 255u, y = (ew >> 8) & 255u;
      if (x < 64 && y < 64 && ((actp >> x) & 1) && ((actp >> y) & 1)) {
        const int k = __popcll(s_vm[w][x][0] & s_vm[w][y][0]) +
                      __popcll(s_vm[w][x][1] & s_vm[w][y][1]);
        a = k >= 2;
        ab = a && ((bis >> x) & 1) && ((bis >> y) & 1);
      }
    }
    acte[q] = ballot(a);
    n_be += __popcll(ballot(ab));
  }
  if (lane == 0) {
    double sv = 0, sf = 0, se = 0;
    for (int v = 0; v < nv; ++v)
      sv += (double)(1.0f / (float)(int)(W[v] >> 24));
    const TetMeta tm = M.tet_meta[h.tet_id];
    for (int p = 0; p < np; ++p) {
      if (!((actp >> p) & 1)) continue;
      const int32_t s = (int32_t)W[nv + p];
      const double hh = s < 0 ? (double)(float)meta_byte(tm, -1 - s) : 1.0;
      sf += 1.0 / hh;
    }
    for (int e = 0; e < ne; ++e)
      if ((acte[e >> 6] >> (e & 63)) & 1)
        se += 1.0 / (double)((W[nv + np + e] >> 16) & 255u);
    euler[c] = (float)(sv - se + sf);
    act[4 * c + 0] = actp;
    act[4 * c + 1] = acte[0];
    act[4 * c + 2] = acte[1];
    act[4 * c + 3] = acte[2];
    n_tf[c] = __popcll(actp & ~bis);
    n_fac[c] = __popcll(actp & bis);
    n_edg[c] = n_be;
  }
}

// position of site j in the ascending neighbour list of site s (-1: none)
__device__ __forceinline__ int knn_pos(const DevSites& S, int s, int j) {
  int64_t lo = S.knn_off[s], hi = S.knn_off[s + 1];
  const int64_t b = lo;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (S.knn_ids[mid] < j)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < S.knn_off[s + 1] && S.knn_ids[lo] == j) ? (int)(lo - b) : -1;
}

constexpr int kEdgePosBits = 12;  // neighbour-list positions in edge keys

// The records of each cell at its offsets.  err[0] is set if a bisector's
// site is missing from the seed's list or sits beyond the edge key's
// position range.
__global__ __launch_bounds__(256) void k_adj_fill(
    int64_t n, const rpd_cell* cells, const uint32_t* words, DevMesh M,
    DevSites S, const uint64_t* act, const int64_t* tf_off,
    const int64_t* fac_off, const int64_t* edg_off, uint64_t* tf_key,
    int32_t* tf_val, uint64_t* fac_key, int32_t* fac_val, uint64_t* edg_key,
    int32_t* edg_val, int32_t* err) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t c = (int64_t)blockIdx.x * kAdjPerBlock + w;
  if (c >= n) return;
  const rpd_cell h = cells[c];
  const int nv = h.nb_v, np = h.nb_p, ne = h.nb_e;
  const uint32_t* W = words + h.word_off;
  const uint64_t actp = act[4 * c];
  const uint64_t below = (1ull << lane) - 1ull;
  const int32_t src = lane < np ? (int32_t)W[nv + lane] : -1;
  const uint64_t bis = ballot(lane < np && src >= 0);
  const bool mine = (actp >> lane) & 1;
  const uint64_t seed = (uint64_t)(uint32_t)h.seed;
  if (mine && src < 0) {  // tet face k of the cell's tet
    const int k = -1 - src;
    const int4 f4 = M.tet_fid[h.tet_id];
    const int fid = k == 0 ? f4.x : k == 1 ? f4.y : k == 2 ? f4.z : f4.w;
    const int64_t r = tf_off[c] + __popcll(actp & ~bis & below);
    tf_key[r] = (seed << 32) | (uint32_t)fid;
    tf_val[r] = (int32_t)(c * 4 + k);
  } else if (mine) {  // bisector facet toward site src
    const int64_t r = fac_off[c] + __popcll(actp & bis & below);
    fac_key[r] = (seed << 32) | (uint32_t)src;
    fac_val[r] = (int32_t)c;
  }
  // edges between two bisectors: key (seed, list position pair)
  int64_t r0 = edg_off[c];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int e = q * 64 + lane;
    bool ab = false;
    uint64_t key = 0;
    if (e < ne && ((act[4 * c + 1 + q] >> lane) & 1)) {
      const uint32_t ew = W[nv + np + e];
      const uint32_t x = ew & 255u, y = (ew >> 8) & 255u;
      if (((bis >> x) & 1) && ((bis >> y) & 1)) {
        ab = true;
        const int px = knn_pos(S, h.seed, (int32_t)W[nv + x]);
        const int py = knn_pos(S, h.seed, (int32_t)W[nv + y]);
        if (px < 0 || py < 0 || px >= (1 << kEdgePosBits) ||
            py >= (1 << kEdgePosBits)) {
          err[0] = 1;
        } else {
          const uint64_t lo = (uint64_t)min(px, py), hi = (uint64_t)max(px, py);
          key = (seed << (2 * kEdgePosBits)) | (lo << kEdgePosBits) | hi;
        }
      }
    }
    const uint64_t m = ballot(ab);
    if (ab) {
      const int64_t r = r0 + __popcll(m & below);
      edg_key[r] = key;
      edg_val[r] = (int32_t)c;
    }
    r0 += __popcll(m);
  }
}

// Per facet record (thread per record): the facet Euler sum of
// cal_halfplane_facet_euler (voronoi_defs.cxx:223-267) of its cell toward
// its neighbour site j, on the first active plane whose id pair holds j --
// a tet face's pair is (global face id, -1), so a face whose id equals j
// comes before the bisector -- with the cell Euler sum's rounding.  Face
// ids are instance-local and site ids batch-wide, so the site id is taken
// relative to its instance's first site (site_off[site_inst[seed]]) for
// that comparison, as a run of the instance alone numbers it.
__global__ __launch_bounds__(256) void k_fac_euler(
    int64_t n, const uint64_t* key, const int32_t* cell, const rpd_cell* cells,
    const uint32_t* words, DevMesh M, const uint64_t* act,
    const int32_t* site_inst, const int64_t* site_off, float* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t c = cell[i];
  const int32_t j = (int32_t)(uint32_t)key[i];
  const rpd_cell h = cells[c];
  const int32_t jl = j - (int32_t)site_off[site_inst[h.seed]];
  const int nv = h.nb_v, np = h.nb_p, ne = h.nb_e;
  const uint32_t* W = words + h.word_off;
  const uint64_t actp = act[4 * c];
  const int4 f4 = M.tet_fid[h.tet_id];
  uint32_t pid = 255;
  for (int q = 0; q < np; ++q) {
    if (!((actp >> q) & 1)) continue;
    const int32_t sq = (int32_t)W[nv + q];
    if (sq < 0) {
      const int k = -1 - sq;
      const int fid = k == 0 ? f4.x : k == 1 ? f4.y : k == 2 ? f4.z : f4.w;
      if (fid != jl) continue;
    } else if (sq != j) {
      continue;
    }
    pid = (uint32_t)q;
    break;
  }
  double sv = 0, se = 0;
  for (int v = 0; v < nv; ++v) {
    const uint32_t x = W[v];
    if ((x & 255u) == pid || ((x >> 8) & 255u) == pid ||
        ((x >> 16) & 255u) == pid)
      sv += (double)(1.0f / (float)(int)(x >> 24));
  }
  for (int e = 0; e < ne; ++e) {
    const uint32_t ew = W[nv + np + e];
    if ((ew & 255u) != pid && ((ew >> 8) & 255u) != pid) continue;
    if (!((act[4 * c + 1 + (e >> 6)] >> (e & 63)) & 1)) continue;
    se += 1.0 / (double)((ew >> 16) & 255u);
  }
  out[i] = (float)(sv - se + 1.0);
}

// tet-face records sorted by key: equal neighbours = one face, two cells
__global__ __launch_bounds__(256) void k_adj_pairs(int64_t n,
                                                   const uint64_t* key,
                                                   const int32_t* val,
                                                   int32_t* nbr,
                                                   int32_t* n_pairs,
                                                   int32_t* err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 >= n || key[i] != key[i + 1]) return;
  if (i + 2 < n && key[i + 2] == key[i]) {  // a face has two sides
    err[1] = 1;
    return;
  }
  const int32_t a = val[i], b = val[i + 1];
  nbr[a] = b >> 2;  // (cell a>>2, local face a&3) -> cell b>>2
  nbr[b] = a >> 2;
  // one atomic per wavefront: its first pairing lane adds the wave's count
  const uint64_t m = __ballot(1);
  if ((threadIdx.x & 63) == (uint32_t)(__ffsll((long long)m) - 1))
    atomicAdd(n_pairs, (int)__popcll(m));
}

__device__ __forceinline__ int32_t uf_root(int32_t* par, int32_t x) {
  int32_t p = __hip_atomic_load(par + x, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
  while (p != x) {
    x = p;
    p = __hip_atomic_load(par + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return x;
}

// hooks the larger root under the smaller; retries if that root was hooked
// meanwhile (each failed exchange means another hook made progress)
__device__ void uf_union(int32_t* par, int32_t a, int32_t b) {
  for (;;) {
    a = uf_root(par, a);
    b = uf_root(par, b);
    if (a == b) return;
    const int32_t lo = min(a, b), hi = max(a, b);
    if (atomicCAS(par + hi, hi, lo) == hi) return;
  }
}

__global__ __launch_bounds__(256) void k_uf_init(int64_t n, int32_t* par) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) par[i] = (int32_t)i;
}

__global__ __launch_bounds__(256) void k_uf_flatten(int64_t n, int32_t* par) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) par[i] = uf_root(par, (int32_t)i);
}

// cells of one sphere across a tet face are one component
__global__ __launch_bounds__(256) void k_uf_cells(int64_t n,
                                                  const int32_t* nbr,
                                                  int32_t* par) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int32_t m = nbr[4 * c + k];
    if (m > c) uf_union(par, (int32_t)c, m);
  }
}

// Records sorted by (key, cell): record i joins the record of the same key
// in each neighbour cell of its cell (a binary search in the key's run).
__global__ __launch_bounds__(256) void k_uf_records(int64_t n,
                                                    const uint64_t* key,
                                                    const int32_t* cell,
                                                    const int32_t* nbr,
                                                    int32_t* par) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t K = key[i];
  const int32_t c = cell[i];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int32_t m = nbr[4 * (int64_t)c + k];
    if (m <= c) continue;
    // first record at or after i with (K, cell >= m)
    int64_t lo = i + 1, hi = n;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (key[mid] < K || (key[mid] == K && cell[mid] < m))
        lo = mid + 1;
      else
        hi = mid;
    }
    if (lo < n && key[lo] == K && cell[lo] == m)
      uf_union(par, (int32_t)i, (int32_t)lo);
  }
}

// run starts of a sorted key stream: flag[i] = 1 where a new key begins
__global__ __launch_bounds__(256) void k_run_flags(int64_t n,
                                                   const uint64_t* key,
                                                   int64_t* flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_run_starts(int64_t n,
                                                    const int64_t* flag,
                                                    const int64_t* pos,
                                                    int64_t* start) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) start[pos[i]] = i;
}

// One thread per run of facet (kind 0) or edge (kind 1) records: the run's
// site and neighbour(s), first record, number of cells and of components;
// for facets the binary64 sum of the records' facet Euler values in cell
// order (is_to_fix_facet_euler); per record its cell and the smallest cell
// of its component.
__global__ __launch_bounds__(256) void k_adj_runs(
    int64_t n_runs, int64_t n, int kind, const int64_t* start,
    const uint64_t* key, const int32_t* cell, const int32_t* par,
    const float* rec_euler, DevSites S, int32_t* runs, double* run_euler,
    int32_t* rec_out) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_runs) return;
  const int64_t b = start[r], e = r + 1 < n_runs ? start[r + 1] : n;
  const uint64_t K = key[b];
  int32_t ncc = 0;
  double es = 0;
  for (int64_t i = b; i < e; ++i) {
    ncc += par[i] == i;
    if (rec_euler) es += (double)rec_euler[i];
    rec_out[2 * i] = cell[i];
    rec_out[2 * i + 1] = cell[par[i]];
  }
  if (kind == 0) {
    int32_t* o = runs + 5 * r;
    o[0] = (int32_t)(K >> 32);
    o[1] = (int32_t)(uint32_t)K;
    o[2] = (int32_t)b;
    o[3] = (int32_t)(e - b);
    o[4] = ncc;
    run_euler[r] = es;
  } else {
    const int s = (int)(K >> (2 * kEdgePosBits));
    const int64_t k0 = S.knn_off[s];
    const int pl = (int)((K >> kEdgePosBits) & ((1u << kEdgePosBits) - 1));
    const int ph = (int)(K & ((1u << kEdgePosBits) - 1));
    int32_t* o = runs + 6 * r;
    o[0] = s;
    o[1] = S.knn_ids[k0 + pl];
    o[2] = S.knn_ids[k0 + ph];
    o[3] = (int32_t)b;
    o[4] = (int32_t)(e - b);
    o[5] = ncc;
  }
}

// The cells' Euler values and component roots in seed order (cells sorted
// by seed, ascending within a seed), so that the per-site loop below reads
// contiguous memory
__global__ __launch_bounds__(256) void k_adj_gather(int64_t n,
                                                    const int32_t* idx,
                                                    const int32_t* par,
                                                    const float* euler,
                                                    float* eul_sorted,
                                                    uint8_t* root_sorted) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t c = idx[i];
  eul_sorted[i] = euler[c];
  root_sorted[i] = par[c] == c;
}

// One thread per site over its cells in ascending order: number of cells,
// of components, and the Euler number euler_sum - num_cells of
// is_to_fix_voro_euler (binary64 sum of the cells' binary32 values, in
// that order)
__global__ __launch_bounds__(256) void k_adj_sites(
    int32_t n_site, const int64_t* start, const float* eul_sorted,
    const uint8_t* root_sorted, int32_t* site_cells, int32_t* site_cc,
    double* site_euler) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_site) return;
  const int64_t b = start[s], e = start[s + 1];
  int32_t ncc = 0;
  double es = 0;
  for (int64_t i = b; i < e; ++i) {
    ncc += root_sorted[i];
    es += (double)eul_sorted[i];
  }
  site_cells[s] = (int32_t)(e - b);
  site_cc[s] = ncc;
  site_euler[s] = es - (double)(e - b);
}

}  // namespace rpd
