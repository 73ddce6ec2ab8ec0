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

}  // namespace rpd
