// Learned candidate proposal for the RPD (device side).
//
// Per tet, the owners (spheres of least power distance) of its centroid and
// of its 4 corners are found by greedy descent on the neighbour graph -- in
// a power diagram a sphere that does not own a point always has a
// neighbour of smaller power distance to it; each owner certainly has a
// cell in the tet.  The tet's 32 proposals are the owners and their
// neighbours (flagged spheres only) of least key = min over the 5 points of
// the power distance minus the point owner's.  A small MLP (48 -> 128 ->
// 128 -> 32, ReLU, BF16 MFMA with FP32 accumulation) scores them from 48
// features: the centroid in domain coordinates, and for the first 8
// proposals their centre offset and radius over the tet's radius and their
// key over its square.  Proposals scoring above a threshold (owners
// always) are clipped; a completeness
// certificate then checks every tet: each active bisector facet of each of
// its cells must lead to a sphere of its proposal set, no cell may have
// been dropped, and some cell must exist (the cells of a set closed under
// facet neighbours partition the tet).  Tets that fail are re-run with the
// reference relation (voronoi.cu:154-193) over every sphere of their
// instance.  Cells are clipped exactly as on the plain path, so the output
// is the plain path's.
//
// Weights are row-major [in][out] (fp32 master copy); the MFMA kernels read
// bf16 copies: transposed [out][in] for the forward pass (B operand of
// X * W) and [in][out] for the backward pass (B operand of dZ * W^T).
// v_mfma_f32_16x16x16_bf16 lane layout: A[m = l%16][k = 4(l/16) + i],
// B[k = 4(l/16) + i][n = l%16], D[m = 4(l/16) + i][n = l%16], i = 0..3.

namespace rpd {

constexpr int kIn = 48, kHid = 128, kOut = 32;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kIn * kHid;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid * kOut;
constexpr int kNParam = kOffB3 + kOut;  // 26912
constexpr int kNFeat = 3 + 8 * 5;       // 43 used of 48

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// binary32 -> bfloat16, round to nearest even (finite inputs)
__device__ __host__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u;
  __builtin_memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}

struct MlpDev {
  const uint16_t* w1t;  // [128][48]
  const uint16_t* w2t;  // [128][128]
  const uint16_t* w3t;  // [32][128]
  const uint16_t* w2;   // [128][128]
  const uint16_t* w3;   // [128][32]
  const float* b1;
  const float* b2;
  const float* b3;
};

__device__ __forceinline__ bf16x4 ld_bf4(const uint16_t* p) {
  const uint2 v = *reinterpret_cast<const uint2*>(p);
  bf16x4 r;
  r[0] = (short)(v.x & 0xffffu);
  r[1] = (short)(v.x >> 16);
  r[2] = (short)(v.y & 0xffffu);
  r[3] = (short)(v.y >> 16);
  return r;
}
__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// ------------------------------------------------------------- features
// neighbour lists in list order with their spheres inline
__global__ __launch_bounds__(256) void k_site_inline(DevSites S, int64_t n_knn,
                                                     int32_t* knn_near,
                                                     double4* knn_site_near) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_knn) return;
  const int m = S.knn_ids[q];
  knn_near[q] = m;
  knn_site_near[q] = S.site[m];
}

__device__ __forceinline__ double pdist(const double c[3], const double4& s) {
  const double dx = c[0] - s.x, dy = c[1] - s.y, dz = c[2] - s.z;
  return ((dx * dx + dy * dy) + dz * dz) - s.w;
}

// keys of the site grid: (cell << 32) | site, per instance grid sg
__global__ __launch_bounds__(256) void k_sg_keys(DevSites S,
                                                 const InstInfo* inst,
                                                 const int32_t* site_inst,
                                                 uint64_t* keys) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S.n_site) return;
  const InstInfo& I = inst[site_inst[i]];
  const double4 s = S.site[i];
  const int x = bin_coord(s.x, I.sg.org[0], I.sg.inv_bs, I.sg.dim[0]);
  const int y = bin_coord(s.y, I.sg.org[1], I.sg.inv_bs, I.sg.dim[1]);
  const int z = bin_coord(s.z, I.sg.org[2], I.sg.inv_bs, I.sg.dim[2]);
  const uint64_t b = (uint64_t)I.sg_base + x +
                     (uint64_t)I.sg.dim[0] * (y + (uint64_t)I.sg.dim[1] * z);
  keys[i] = (b << 32) | (uint32_t)i;
}

// grid start for a descent at point c: least power distance among the
// spheres of c's cell of the sphere grid (rings up to 2), else the first
// sphere of the instance
__device__ __forceinline__ int grid_start(const DevSites& S,
                                         const InstInfo& I,
                                         const int32_t* sg_start,
                                         const int32_t* sg_ids,
                                         const double c[3], double* best) {
  int cur = -1;
  *best = INFINITY;
  const int cx = bin_coord(c[0], I.sg.org[0], I.sg.inv_bs, I.sg.dim[0]);
  const int cy = bin_coord(c[1], I.sg.org[1], I.sg.inv_bs, I.sg.dim[1]);
  const int cz = bin_coord(c[2], I.sg.org[2], I.sg.inv_bs, I.sg.dim[2]);
  for (int ring = 0; ring <= 2 && cur < 0; ++ring)
    for (int dz = -ring; dz <= ring; ++dz)
      for (int dy = -ring; dy <= ring; ++dy)
        for (int dx = -ring; dx <= ring; ++dx) {
          if (max(abs(dx), max(abs(dy), abs(dz))) != ring) continue;
          const int xx = cx + dx, yy = cy + dy, zz = cz + dz;
          if (xx < 0 || yy < 0 || zz < 0 || xx >= I.sg.dim[0] ||
              yy >= I.sg.dim[1] || zz >= I.sg.dim[2])
            continue;
          const int64_t b = I.sg_base + xx +
                            (int64_t)I.sg.dim[0] *
                                (yy + (int64_t)I.sg.dim[1] * zz);
          for (int q = sg_start[b]; q < sg_start[b + 1]; ++q) {
            const int s = sg_ids[q];
            const double pd = pdist(c, S.site[s]);
            if (pd < *best || (pd == *best && s < cur)) {
              *best = pd;
              cur = s;
            }
          }
        }
  if (cur < 0 && I.n_site > 0) {
    cur = I.site_off;
    *best = pdist(c, S.site[cur]);
  }
  return cur;
}

// greedy descent to the owner of c: (pd, id) decreases strictly; in a
// power diagram a sphere that does not own c has a neighbour closer to it
__device__ __forceinline__ int descend(const DevSites& S, const double c[3],
                                       int cur, double* best) {
  for (int step = 0; step < (1 << 16); ++step) {
    int nxt = cur;
    double nb = *best;
    for (int64_t q = S.knn_off[cur]; q < S.knn_off[cur + 1]; ++q) {
      const int id = S.knn_near[q];
      const double pd = pdist(c, S.knn_site_near[q]);
      if (pd < nb || (pd == nb && id < nxt)) {
        nb = pd;
        nxt = id;
      }
    }
    if (nxt == cur) break;
    cur = nxt;
    *best = nb;
  }
  return cur;
}

// owner of every mesh vertex (-1 if its instance has no sphere)
__global__ __launch_bounds__(256) void k_vertex_owner(
    DevMesh M, DevSites S, const InstInfo* inst, const int32_t* vert_inst,
    const int32_t* sg_start, const int32_t* sg_ids, int32_t* owner) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= M.n_vert) return;
  const double4 p = M.vert[v];
  const double c[3] = {p.x, p.y, p.z};
  double best;
  const int st = grid_start(S, inst[vert_inst[v]], sg_start, sg_ids, c, &best);
  owner[v] = st < 0 ? -1 : descend(S, c, st, &best);
}

// One thread per tet.  The owners of its centroid and of its 4 corners
// belong to its sphere set; the proposals are the union of their 1-rings
// (with them), ranked by key(s) = min over the 5 points x of
// pd(x, s) - pd(x, owner of x) -- 0 exactly for the owners -- then id;
// 32 kept, flagged spheres only.  omask marks the owner slots.
__global__ __launch_bounds__(64) void k_knn_features(
    DevMesh M, DevSites S, const InstInfo* inst, const int32_t* tet_inst,
    const int32_t* vowner, int32_t* knn, uint16_t* X, uint32_t* omask) {
  __shared__ float s_key[kOut][64];
  __shared__ int32_t s_id[kOut][64];
  const int lane = threadIdx.x;
  const int t = blockIdx.x * 64 + lane;
  if (t >= M.n_tet) return;
  TetLoad L;
  load_tet(M, t, L);
  double P[5][3];
  tet_centroid(L, P[0]);
  for (int k = 0; k < 4; ++k) {
    P[1 + k][0] = L.V[k].x;
    P[1 + k][1] = L.V[k].y;
    P[1 + k][2] = L.V[k].z;
  }
  double rho2 = 0;
  for (int k = 0; k < 4; ++k) {
    const double dx = P[1 + k][0] - P[0][0], dy = P[1 + k][1] - P[0][1],
                 dz = P[1 + k][2] - P[0][2];
    rho2 = fmax(rho2, (dx * dx + dy * dy) + dz * dz);
  }
  const InstInfo& I = inst[tet_inst[t]];
  uint16_t* x = X + (int64_t)t * kIn;
  int32_t* kt = knn + (int64_t)t * kOut;
  int own[5];
  double pmin[5];
  own[1] = vowner[L.tv.x];
  own[2] = vowner[L.tv.y];
  own[3] = vowner[L.tv.z];
  own[4] = vowner[L.tv.w];
  int cnt = 0;
  uint32_t om = 0;
  if (own[1] >= 0) {
    for (int k = 1; k < 5; ++k) pmin[k] = pdist(P[k], S.site[own[k]]);
    // centroid owner: descend from the corner owner closest to it
    int st = own[1];
    double best = pdist(P[0], S.site[st]);
    for (int k = 2; k < 5; ++k) {
      const double pd = pdist(P[0], S.site[own[k]]);
      if (pd < best || (pd == best && own[k] < st)) {
        best = pd;
        st = own[k];
      }
    }
    own[0] = descend(S, P[0], st, &best);
    pmin[0] = best;
    auto key_of = [&](const double4& s) {
      double k = INFINITY;
#pragma unroll
      for (int j = 0; j < 5; ++j) k = fmin(k, pdist(P[j], s) - pmin[j]);
      return k;
    };
    auto ins = [&](float key, int id) {
      for (int q = 0; q < cnt; ++q)
        if (s_id[q][lane] == id) return;
      if (cnt == kOut &&
          !(key < s_key[kOut - 1][lane] ||
            (key == s_key[kOut - 1][lane] && id < s_id[kOut - 1][lane])))
        return;
      int j = cnt < kOut ? cnt : kOut - 1;
      while (j > 0 && (s_key[j - 1][lane] > key ||
                       (s_key[j - 1][lane] == key && s_id[j - 1][lane] > id))) {
        s_key[j][lane] = s_key[j - 1][lane];
        s_id[j][lane] = s_id[j - 1][lane];
        --j;
      }
      s_key[j][lane] = key;
      s_id[j][lane] = id;
      if (cnt < kOut) ++cnt;
    };
    for (int j = 0; j < 5; ++j) {
      const int o = own[j];
      bool seen = false;
      for (int q = 0; q < j; ++q) seen = seen || own[q] == o;
      if (seen) continue;
      if (S.flags[o] == 1u) ins(0.f, o);
      for (int64_t q = S.knn_off[o]; q < S.knn_off[o + 1]; ++q) {
        const int id = S.knn_near[q];
        if (S.flags[id] != 1u) continue;
        ins((float)key_of(S.knn_site_near[q]), id);
      }
    }
    for (int k = 0; k < cnt; ++k)
      if (s_key[k][lane] == 0.f) {
        const int id = s_id[k][lane];
        bool is_owner = false;
        for (int j = 0; j < 5; ++j) is_owner = is_owner || own[j] == id;
        if (is_owner) om |= 1u << k;
      }
  }
  omask[t] = om;
  for (int k = 0; k < kOut; ++k) kt[k] = k < cnt ? s_id[k][lane] : -1;
  const double ext =
      fmax(fmax(I.dom[3] - I.dom[0], I.dom[4] - I.dom[1]), I.dom[5] - I.dom[2]);
  const double inv_ext = ext > 0 ? 1.0 / ext : 0.0;
  const double inv_r = rho2 > 0 ? 1.0 / sqrt(rho2) : 0.0;
  float f[kIn];
#pragma unroll
  for (int q = 0; q < kIn; ++q) f[q] = 0.f;
  f[0] = (float)((P[0][0] - I.dom[0]) * inv_ext);
  f[1] = (float)((P[0][1] - I.dom[1]) * inv_ext);
  f[2] = (float)((P[0][2] - I.dom[2]) * inv_ext);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (k < cnt) {
      const double4 sp = S.site[s_id[k][lane]];
      f[3 + 5 * k + 0] = (float)((sp.x - P[0][0]) * inv_r);
      f[3 + 5 * k + 1] = (float)((sp.y - P[0][1]) * inv_r);
      f[3 + 5 * k + 2] = (float)((sp.z - P[0][2]) * inv_r);
      f[3 + 5 * k + 3] = (float)(sqrt(sp.w) * inv_r);
      f[3 + 5 * k + 4] = (float)(s_key[k][lane] * (inv_r * inv_r));
    }
  }
#pragma unroll
  for (int q = 0; q < kIn; q += 4) {
    uint2 v;
    v.x = (uint32_t)f2bf(f[q]) | ((uint32_t)f2bf(f[q + 1]) << 16);
    v.y = (uint32_t)f2bf(f[q + 2]) | ((uint32_t)f2bf(f[q + 3]) << 16);
    *reinterpret_cast<uint2*>(x + q) = v;
  }
}

// Labels from a plain run's output cells (sorted by tet, site): proposal k
// of tet t is positive iff that sphere has a cell in t; mask = valid slot.
__global__ __launch_bounds__(256) void k_labels(int32_t n_tet,
                                                const rpd_cell* cells,
                                                int64_t n_cells,
                                                const int32_t* knn, float* Y,
                                                float* Mk) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tet) return;
  int64_t lo = 0, hi = n_cells;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (cells[mid].tet_id < t)
      lo = mid + 1;
    else
      hi = mid;
  }
  int64_t e = lo;
  while (e < n_cells && cells[e].tet_id == t) ++e;
  for (int k = 0; k < kOut; ++k) {
    const int id = knn[(int64_t)t * kOut + k];
    bool y = false;
    for (int64_t q = lo; q < e && id >= 0; ++q) y = y || cells[q].seed == id;
    Y[(int64_t)t * kOut + k] = y ? 1.f : 0.f;
    Mk[(int64_t)t * kOut + k] = id >= 0 ? 1.f : 0.f;
  }
}

// ------------------------------------------------------------------ MLP
// One wave per 16-row tile.  Activations pass through LDS to turn the
// D layout of one layer into the A layout of the next.
constexpr int kActStride = kHid + 8;  // bf16 per LDS row (bank spread)

struct WaveLds {
  uint16_t h1[16][kActStride];
  uint16_t h2[16][kActStride];
  uint16_t dz[16][kOut + 8];
};

__device__ __forceinline__ void wsync_l() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// X tile rows (A operand, 3 k-blocks); rows with id < 0 are zero
__device__ __forceinline__ void load_x(const uint16_t* X, int64_t row_id,
                                       int l, bf16x4 a[3]) {
#pragma unroll
  for (int kb = 0; kb < 3; ++kb) {
    if (row_id >= 0) {
      a[kb] = ld_bf4(X + row_id * kIn + 16 * kb + 4 * (l / 16));
    } else {
      a[kb] = bf16x4{0, 0, 0, 0};
    }
  }
}

// layers 1-2 of one tile into LDS (h1, h2) and layer 3 logits z[nb][i]
__device__ __forceinline__ void mlp_forward_tile(const MlpDev& P,
                                                 const bf16x4 a1[3], int l,
                                                 WaveLds& W, f32x4 z[2]) {
  const int col = l % 16, rq = 4 * (l / 16);
#pragma unroll
  for (int nb = 0; nb < 8; ++nb) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 3; ++kb)
      acc = mfma(a1[kb], ld_bf4(P.w1t + (16 * nb + col) * kIn + 16 * kb + rq),
                 acc);
    const float b = P.b1[16 * nb + col];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      W.h1[rq + i][16 * nb + col] = f2bf(fmaxf(acc[i] + b, 0.f));
  }
  wsync_l();
  bf16x4 a2[8];
#pragma unroll
  for (int kb = 0; kb < 8; ++kb) a2[kb] = ld_bf4(&W.h1[col][16 * kb + rq]);
#pragma unroll
  for (int nb = 0; nb < 8; ++nb) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 8; ++kb)
      acc = mfma(a2[kb],
                 ld_bf4(P.w2t + (16 * nb + col) * kHid + 16 * kb + rq), acc);
    const float b = P.b2[16 * nb + col];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      W.h2[rq + i][16 * nb + col] = f2bf(fmaxf(acc[i] + b, 0.f));
  }
  wsync_l();
  bf16x4 a3[8];
#pragma unroll
  for (int kb = 0; kb < 8; ++kb) a3[kb] = ld_bf4(&W.h2[col][16 * kb + rq]);
#pragma unroll
  for (int nb = 0; nb < 2; ++nb) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 8; ++kb)
      acc = mfma(a3[kb],
                 ld_bf4(P.w3t + (16 * nb + col) * kHid + 16 * kb + rq), acc);
    const float b = P.b3[16 * nb + col];
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] += b;
    z[nb] = acc;
  }
}

// Inference: selection mask per row (bit k: logit_k > thr); optional logits
__global__ __launch_bounds__(256) void k_mlp_infer(MlpDev P,
                                                   const uint16_t* X,
                                                   int64_t n, float thr,
                                                   uint32_t* sel,
                                                   float* logits) {
  __shared__ WaveLds s_w[4];
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t n_tiles = (n + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * 4 + wv; tile < n_tiles;
       tile += (int64_t)gridDim.x * 4) {
    const int64_t row0 = tile * 16;
    const int64_t my_row = row0 + l % 16;
    bf16x4 a1[3];
    load_x(X, my_row < n ? my_row : -1, l, a1);
    f32x4 z[2];
    mlp_forward_tile(P, a1, l, s_w[wv], z);
    const int rq = 4 * (l / 16), col = l % 16;
    uint64_t bm[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) bm[i][nb] = __ballot(z[nb][i] > thr);
    if (l < 16) {
      // row l = 4 g + i: lanes 16 g .. 16 g + 15 of ballot (i, nb)
      const int g = l / 4, i = l % 4;
      uint64_t b0 = i == 0 ? bm[0][0] : (i == 1 ? bm[1][0] : (i == 2 ? bm[2][0] : bm[3][0]));
      uint64_t b1 = i == 0 ? bm[0][1] : (i == 1 ? bm[1][1] : (i == 2 ? bm[2][1] : bm[3][1]));
      const uint32_t m = (uint32_t)((b0 >> (16 * g)) & 0xffffu) |
                         ((uint32_t)((b1 >> (16 * g)) & 0xffffu) << 16);
      if (row0 + l < n) sel[row0 + l] = m;
    }
    if (logits) {
#pragma unroll
      for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int64_t r = row0 + rq + i;
          if (r < n) logits[r * kOut + 16 * nb + col] = z[nb][i];
        }
    }
    wsync_l();
  }
}

// Training, forward and backward of one minibatch (rows idx[r], -1 = pad):
// stores the bf16 activations and output gradients the weight-gradient
// kernel needs, and per-tile loss sums.  Loss: pos_w-weighted binary cross
// entropy over valid slots, times inv_count.
__global__ __launch_bounds__(256) void k_mlp_fb(
    MlpDev P, const uint16_t* X, const float* Y, const float* Mk,
    const int32_t* idx, int32_t n_rows, float pos_w, float inv_count,
    uint16_t* XB, uint16_t* H1, uint16_t* H2, uint16_t* DZ1, uint16_t* DZ2,
    uint16_t* DZ3, float* loss_tile) {
  __shared__ WaveLds s_w[4];
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = l % 16, rq = 4 * (l / 16);
  const int n_tiles = n_rows / 16;
  const int tile = blockIdx.x * 4 + wv;
  if (tile >= n_tiles) return;
  WaveLds& W = s_w[wv];
  const int64_t row0 = (int64_t)tile * 16;
  const int64_t src = idx[row0 + col];
  bf16x4 a1[3];
  load_x(X, src, l, a1);
  // the gathered input rows, for dW1
#pragma unroll
  for (int kb = 0; kb < 3; ++kb) {
    uint2 v;
    v.x = (uint32_t)(uint16_t)a1[kb][0] | ((uint32_t)(uint16_t)a1[kb][1] << 16);
    v.y = (uint32_t)(uint16_t)a1[kb][2] | ((uint32_t)(uint16_t)a1[kb][3] << 16);
    *reinterpret_cast<uint2*>(XB + (row0 + col) * kIn + 16 * kb + rq) = v;
  }
  f32x4 z[2];
  mlp_forward_tile(P, a1, l, W, z);
  // h1, h2 tiles to global (row-major [row][128])
  for (int q = l; q < 16 * kHid; q += 64) {
    const int r = q / kHid, cc = q % kHid;
    H1[(row0 + r) * kHid + cc] = W.h1[r][cc];
    H2[(row0 + r) * kHid + cc] = W.h2[r][cc];
  }
  // output gradient and loss (D layout: row rq + i, column 16 nb + col)
  float lsum = 0.f;
#pragma unroll
  for (int nb = 0; nb < 2; ++nb)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = rq + i, k = 16 * nb + col;
      const int64_t s = idx[row0 + r];
      float y = 0.f, m = 0.f;
      if (s >= 0) {
        y = Y[s * kOut + k];
        m = Mk[s * kOut + k];
      }
      const float zz = z[nb][i];
      const float w = y > 0.5f ? pos_w : 1.f;
      const float p = 1.f / (1.f + expf(-zz));
      const float g = m * w * (p - y) * inv_count;
      // softplus(z) - y z, stable
      const float sp = fmaxf(zz, 0.f) + log1pf(expf(-fabsf(zz)));
      lsum += m * w * (sp - y * zz);
      const uint16_t gb = f2bf(g);
      W.dz[r][k] = gb;
      DZ3[(row0 + r) * kOut + k] = gb;
    }
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
  if (l == 0) loss_tile[tile] = lsum * inv_count;
  wsync_l();
  // dH2 = dZ3 W3^T, dZ2 = dH2 * [h2 > 0]
  bf16x4 ad[2];
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) ad[kb] = ld_bf4(&W.dz[col][16 * kb + rq]);
#pragma unroll
  for (int nb = 0; nb < 8; ++nb) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
      acc = mfma(ad[kb], ld_bf4(P.w3 + (16 * nb + col) * kOut + 16 * kb + rq),
                 acc);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = rq + i, cc = 16 * nb + col;
      const float d = bf2f(W.h2[r][cc]) > 0.f ? acc[i] : 0.f;
      const uint16_t db = f2bf(d);
      DZ2[(row0 + r) * kHid + cc] = db;
      acc[i] = bf2f(db);
    }
    // keep dZ2 for the next product (h2 rows already used for this block)
#pragma unroll
    for (int i = 0; i < 4; ++i) W.h2[rq + i][16 * nb + col] = f2bf(acc[i]);
  }
  wsync_l();
  // dH1 = dZ2 W2^T, dZ1 = dH1 * [h1 > 0]
  bf16x4 a2[8];
#pragma unroll
  for (int kb = 0; kb < 8; ++kb) a2[kb] = ld_bf4(&W.h2[col][16 * kb + rq]);
#pragma unroll
  for (int nb = 0; nb < 8; ++nb) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 8; ++kb)
      acc = mfma(a2[kb], ld_bf4(P.w2 + (16 * nb + col) * kHid + 16 * kb + rq),
                 acc);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = rq + i, cc = 16 * nb + col;
      const float d = bf2f(W.h1[r][cc]) > 0.f ? acc[i] : 0.f;
      DZ1[(row0 + r) * kHid + cc] = f2bf(d);
    }
  }
}

// Weight gradient partials: part[chunk][in][out] = sum over the chunk's
// rows of A[row][in] * DZ[row][out] (MFMA over 16-row k-steps); one wave
// per 16x16 output block, blockIdx.y = chunk.
__global__ __launch_bounds__(64) void k_mlp_wgrad(const uint16_t* A,
                                                  const uint16_t* DZ,
                                                  int32_t in, int32_t out,
                                                  int32_t tiles_per_chunk,
                                                  int32_t n_tiles,
                                                  float* part) {
  const int l = threadIdx.x;
  const int nbo = out / 16;
  const int mb = blockIdx.x / nbo, nb = blockIdx.x % nbo;
  const int chunk = blockIdx.y;
  const int col = l % 16, rq = 4 * (l / 16);
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  const int t0 = chunk * tiles_per_chunk;
  const int t1 = min(n_tiles, t0 + tiles_per_chunk);
  for (int tile = t0; tile < t1; ++tile) {
    const int64_t r0 = (int64_t)tile * 16 + rq;
    bf16x4 a, b;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a[i] = (short)A[(r0 + i) * in + 16 * mb + col];
      b[i] = (short)DZ[(r0 + i) * out + 16 * nb + col];
    }
    acc = mfma(a, b, acc);
  }
  float* p = part + (int64_t)chunk * in * out;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    p[(int64_t)(16 * mb + rq + i) * out + 16 * nb + col] = acc[i];
}

// Gradient of every parameter (fixed summation order), optionally an Adam
// step on the fp32 master copy.
struct AdamArgs {
  float lr, beta1, beta2, eps;
  float bc1, bc2;  // 1 - beta^step
  int32_t apply;
};
__global__ __launch_bounds__(256) void k_mlp_update(
    float* master, float* m1, float* m2, const float* part1,
    const float* part2, const float* part3, int32_t n_chunks,
    const uint16_t* DZ1, const uint16_t* DZ2, const uint16_t* DZ3,
    int32_t n_rows, AdamArgs a, float* grad_out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNParam) return;
  float g = 0.f;
  if (p < kOffB1) {
    for (int c = 0; c < n_chunks; ++c) g += part1[(int64_t)c * kIn * kHid + p];
  } else if (p < kOffW2) {
    const int j = p - kOffB1;
    for (int r = 0; r < n_rows; ++r) g += bf2f(DZ1[(int64_t)r * kHid + j]);
  } else if (p < kOffB2) {
    const int q = p - kOffW2;
    for (int c = 0; c < n_chunks; ++c)
      g += part2[(int64_t)c * kHid * kHid + q];
  } else if (p < kOffW3) {
    const int j = p - kOffB2;
    for (int r = 0; r < n_rows; ++r) g += bf2f(DZ2[(int64_t)r * kHid + j]);
  } else if (p < kOffB3) {
    const int q = p - kOffW3;
    for (int c = 0; c < n_chunks; ++c)
      g += part3[(int64_t)c * kHid * kOut + q];
  } else {
    const int j = p - kOffB3;
    for (int r = 0; r < n_rows; ++r) g += bf2f(DZ3[(int64_t)r * kOut + j]);
  }
  if (grad_out) grad_out[p] = g;
  if (a.apply) {
    const float mm = a.beta1 * m1[p] + (1.f - a.beta1) * g;
    const float vv = a.beta2 * m2[p] + (1.f - a.beta2) * g * g;
    m1[p] = mm;
    m2[p] = vv;
    master[p] -= a.lr * (mm / a.bc1) / (sqrtf(vv / a.bc2) + a.eps);
  }
}

// bf16 copies of the master weights in the layouts the MFMA kernels read
__global__ __launch_bounds__(256) void k_mlp_pack(const float* master,
                                                  uint16_t* w1t, uint16_t* w2t,
                                                  uint16_t* w3t, uint16_t* w2,
                                                  uint16_t* w3, float* bias) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < kIn * kHid) {  // W1[in][out] -> w1t[out][in]
    const int i = q / kHid, o = q % kHid;
    w1t[o * kIn + i] = f2bf(master[kOffW1 + q]);
  }
  if (q < kHid * kHid) {
    const int i = q / kHid, o = q % kHid;
    const uint16_t v = f2bf(master[kOffW2 + q]);
    w2t[o * kHid + i] = v;
    w2[q] = v;
  }
  if (q < kHid * kOut) {
    const int i = q / kOut, o = q % kOut;
    const uint16_t v = f2bf(master[kOffW3 + q]);
    w3t[o * kHid + i] = v;
    w3[q] = v;
  }
  if (q < kHid) bias[q] = master[kOffB1 + q];
  if (q < kHid) bias[kHid + q] = master[kOffB2 + q];
  if (q < kOut) bias[2 * kHid + q] = master[kOffB3 + q];
}

// ------------------------------------------------- learned candidate sets
// the non-learned rule: the first k proposals by power distance
__global__ __launch_bounds__(256) void k_sel_topk(int32_t n_tet, int32_t k,
                                                  uint32_t* sel) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tet) return;
  sel[t] = k >= 32 ? ~0u : (k <= 0 ? 0u : ((1u << k) - 1u));
}

// proposals of tet t: the owner slots always, others if selected
__device__ __forceinline__ int proposals(const int32_t* knn, uint32_t sel,
                                         int t, int32_t out[kOut]) {
  int n = 0;
  for (int k = 0; k < kOut; ++k) {
    const int id = knn[(int64_t)t * kOut + k];
    if (id < 0) break;
    if ((sel >> k) & 1u) out[n++] = id;
  }
  // ascending site id (the clip order of the candidate list)
  for (int a = 1; a < n; ++a) {
    const int v = out[a];
    int b = a;
    while (b > 0 && out[b - 1] > v) {
      out[b] = out[b - 1];
      --b;
    }
    out[b] = v;
  }
  return n;
}

__global__ __launch_bounds__(64) void k_learn_count(int32_t n_tet,
                                                    const int32_t* knn,
                                                    const uint32_t* sel,
                                                    const uint32_t* omask,
                                                    int64_t* count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tet) return;
  int n = 0;
  const uint32_t s = sel[t] | omask[t];
  for (int k = 0; k < kOut; ++k) {
    if (knn[(int64_t)t * kOut + k] < 0) break;
    n += ((s >> k) & 1u) ? 1 : 0;
  }
  count[t] = n;
}

__global__ __launch_bounds__(64) void k_learn_fill(int32_t n_tet,
                                                   const int32_t* knn,
                                                   const uint32_t* sel,
                                                   const uint32_t* omask,
                                                   const int64_t* off,
                                                   Cand* cand,
                                                   int32_t* rel_ids) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tet) return;
  int32_t ids[kOut];
  const int n = proposals(knn, sel[t] | omask[t], t, ids);
  const int64_t o = off[t];
  for (int k = 0; k < n; ++k) {
    Cand c;
    c.tet = t;
    c.seed = ids[k];
    c.m0 = 0;
    c.m1 = 0;
    cand[o + k] = c;
    rel_ids[o + k] = ids[k];
  }
}

// Certificate over candidates [0, n): flags tets whose proposal set is not
// shown complete; marks for exclusion the near-zero cells of spheres the
// reference relation rejects (only such slivers can differ between the
// relation's rounding and the exact clip).
__global__ __launch_bounds__(256) void k_certify(
    int64_t n, DevMesh M, DevSites S, const Cand* cand, const int32_t* loc,
    TierViews tv, const int64_t* tet_off, const int32_t* rel_ids,
    int32_t geometry, uint8_t* flag, uint8_t* has_cell, uint8_t* excl) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int t = cand[c].tet, s = cand[c].seed;
  const int32_t l = loc[c];
  const int tier = l >> 28;
  const int64_t idx = tier ? (l & ((1 << 28) - 1)) : c;
  const SlotView& sv = tv.t[tier];
  const uint32_t hd = sv.hdr[idx];
  const uint32_t st = hd >> 24;
  excl[c] = 0;
  if (st == RPD_NO_INTERSECTION) return;
  if (st != RPD_SUCCESS) {
    flag[t] = 1;
    return;
  }
  TetLoad L;
  load_tet(M, t, L);
  bool keep = true;
  if (geometry) {
    const double4 a = L.V[0], b = L.V[1], cc = L.V[2], d = L.V[3];
    const double ux = b.x - a.x, uy = b.y - a.y, uz = b.z - a.z;
    const double vx = cc.x - a.x, vy = cc.y - a.y, vz = cc.z - a.z;
    const double wx = d.x - a.x, wy = d.y - a.y, wz = d.z - a.z;
    const double vol6 = fabs(ux * (vy * wz - vz * wy) - uy * (vx * wz - vz * wx) +
                             uz * (vx * wy - vy * wx));
    if (sv.geom[4 * idx] < 1e-9 * vol6) {
      uint32_t it = 0;
      keep = relation(S, L, s, &it);
    }
  } else {
    uint32_t it = 0;
    keep = relation(S, L, s, &it);
  }
  if (!keep) {
    excl[c] = 1;
    return;
  }
  has_cell[t] = 1;
  const int nv = hd & 255, np = (hd >> 8) & 255;
  const uint32_t* w = sv.words + idx * sv.stride_words;
  uint64_t act0 = 0, act1 = 0;
  for (int v = 0; v < nv; ++v) {
    const uint32_t x = w[v];
    for (int k = 0; k < 3; ++k) {
      const uint32_t p = (x >> (8 * k)) & 255u;
      if (p < 64)
        act0 |= 1ull << p;
      else
        act1 |= 1ull << (p - 64);
    }
  }
  const int64_t o0 = tet_off[t], o1 = tet_off[t + 1];
  for (int p = 0; p < np; ++p) {
    const bool active = p < 64 ? ((act0 >> p) & 1) : ((act1 >> (p - 64)) & 1);
    const int32_t src = (int32_t)w[nv + p];
    if (!active || src < 0) continue;
    int64_t lo = o0, hi = o1;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (rel_ids[mid] < src)
        lo = mid + 1;
      else
        hi = mid;
    }
    if (lo == o1 || rel_ids[lo] != src) {
      flag[t] = 1;
      return;
    }
  }
}

__global__ __launch_bounds__(256) void k_flag_final(int32_t n_tet,
                                                    uint8_t* flag,
                                                    const uint8_t* has_cell,
                                                    int32_t* flag32) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tet) return;
  const uint8_t f = flag[t] || !has_cell[t];
  flag[t] = f;
  flag32[t] = f;
}

__global__ __launch_bounds__(256) void k_flag_list(int32_t n_tet,
                                                   const int32_t* flag32,
                                                   const int32_t* pos,
                                                   int32_t* list) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tet) return;
  if (flag32[t]) list[pos[t]] = t;
}

// Fallback: the reference relation of each listed tet against every sphere
// of its instance; one wave per tet, sites in ascending order.
__global__ __launch_bounds__(64) void k_brute(
    int32_t n_list, const int32_t* list, DevMesh M, DevSites S,
    const InstInfo* inst, const int32_t* tet_inst, const int64_t* off,
    int64_t* count, Cand* cand, int64_t cand_base) {
  const int f = blockIdx.x;
  if (f >= n_list) return;
  const int lane = threadIdx.x;
  const int t = list[f];
  TetLoad L;
  load_tet(M, t, L);
  const InstInfo& I = inst[tet_inst[t]];
  int64_t o = off ? cand_base + off[f] : 0;
  int64_t n = 0;
  for (int s0 = 0; s0 < I.n_site; s0 += 64) {
    const int s = I.site_off + s0 + lane;
    bool r = false;
    if (s0 + lane < I.n_site) {
      uint32_t it = 0;
      r = relation(S, L, s, &it);
    }
    const uint64_t b = __ballot(r);
    if (off && r) {
      Cand c;
      c.tet = t;
      c.seed = s;
      c.m0 = 0;
      c.m1 = 0;
      cand[o + n + __popcll(b & ((1ull << lane) - 1))] = c;
    }
    n += __popcll(b);
  }
  if (count && lane == 0) count[f] = n;
}

// keys of the final candidate order; excluded candidates sort last
__global__ __launch_bounds__(256) void k_final_keys(
    int64_t n, int64_t n_first, const Cand* cand, const uint8_t* flag,
    const uint8_t* excl, uint64_t* keys, int32_t* vals, int32_t* n_incl) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const Cand cd = cand[c];
  const bool out = c < n_first && (flag[cd.tet] || excl[c]);
  keys[c] = out ? ~0ull : (((uint64_t)(uint32_t)cd.tet << 32) | (uint32_t)cd.seed);
  vals[c] = (int32_t)c;
  if (!out) atomicAdd(n_incl, 1);
}

// per-tet CSR of the final candidate sets
__global__ __launch_bounds__(256) void k_final_rel(int64_t n_incl,
                                                   const Cand* cand,
                                                   const int32_t* perm,
                                                   int64_t* tet_count,
                                                   int32_t* rel_ids) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_incl) return;
  const Cand cd = cand[perm[r]];
  rel_ids[r] = cd.seed;
  atomicAdd((unsigned long long*)&tet_count[cd.tet], 1ull);
}

}  // namespace rpd
