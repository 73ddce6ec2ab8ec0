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
// certificate, evaluated by the clip kernel on each cell whose status is
// final (CertArgs), then checks every tet: each active bisector facet of each of
// its cells must lead to a sphere of its proposal set, no cell may have
// been dropped, and some cell must exist (the cells of a set closed under
// facet neighbours partition the tet).  Only proposals passing the
// reference relation (voronoi.cu:154-193) are clipped.  A missing facet
// neighbour is added (if related) and clipped, and the new cells are
// certified in turn, until every tet's set is closed; tets with a dropped
// cell, without any cell, or not closed after a few rounds are re-run with
// the relation over every sphere of their instance.  Cells are clipped as
// on the plain path, so the output is the plain path's.
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

// four consecutive bf16 (8-byte aligned) as an MFMA operand
__device__ __forceinline__ bf16x4 ld_bf4(const uint16_t* p) {
  return *reinterpret_cast<const bf16x4*>(p);
}
// two binary32 -> two bfloat16 in one word (v_cvt_pk_bf16_f32: round to
// nearest even, as f2bf for finite inputs)
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t f2bf2(float a, float b) {
  const f32x2 v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
}
__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// ------------------------------------------------------------- features
// neighbour lists in list order with their spheres inline (flag in the
// id word, near_word)
// One wavefront per site: its neighbour list nearest bisector first (the
// signed distance of the site's centre to each bisector, then the list
// position), written with the neighbour spheres inline.  The relation test
// meets a rejecting neighbour early in that order (its outcome does not
// depend on the order); lists longer than kNearMax keep their order.
__global__ __launch_bounds__(64) void k_site_inline(DevSites S,
                                                    int32_t* knn_near,
                                                    double4* knn_site_near) {
  __shared__ double s_key[kNearMax];
  __shared__ int32_t s_near[kNearMax];
  const int i = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t k0 = S.knn_off[i];
  const int deg = (int)(S.knn_off[i + 1] - k0);
  if (deg > kNearMax) {
    for (int k = lane; k < deg; k += 64) {
      const int m = S.knn_ids[k0 + k];
      knn_near[k0 + k] = near_word(S, m);
      knn_site_near[k0 + k] = S.site[m];
    }
    return;
  }
  const rpd_d4 si = ld4(S.site + i);
  for (int k = lane; k < deg; k += 64) {
    const rpd_d4 e = rpd_bisector(si, ld4(S.site + S.knn_ids[k0 + k]));
    double key = (e.x * si.x + e.y * si.y + e.z * si.z + e.w) /
                 sqrt(e.x * e.x + e.y * e.y + e.z * e.z);
    s_key[k] = key == key ? key : -INFINITY;  // a total order
  }
  wsync();
  for (int k = lane; k < deg; k += 64) {
    const double kk = s_key[k];
    int r = 0;
    for (int u = 0; u < deg; ++u) {
      const double ku = s_key[u];
      r += ku < kk || (ku == kk && u < k);
    }
    s_near[r] = S.knn_ids[k0 + k];
  }
  wsync();
  for (int k = lane; k < deg; k += 64) {
    const int m = s_near[k];
    knn_near[k0 + k] = near_word(S, m);
    knn_site_near[k0 + k] = S.site[m];
  }
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
      const int id = near_id(S.knn_near[q]);
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

// the median of x, a and p where p <= a: max(p, min(x, a)).  32-bit words
// take the full median form, max(min(x, a), min(max(x, a), p)), which
// compiles to one v_med3_u32 (v_med3_i32 for signed words); 64-bit words
// keep the two-step form (no 64-bit median instruction).
template <typename T>
__device__ __forceinline__ T slot_median(T x, T a, T p) {
  if constexpr (sizeof(T) == 4) {
    return max(min(x, a), min(max(x, a), p));
  } else {
    const T m = x < a ? x : a;
    return p < m ? m : p;
  }
}

// Proposal list entries: (key, id) packed into one unsigned word so that
// integer order is (key, id) order.  The key keeps the top 16 bits of its
// non-negative binary32 pattern (8 exponent and 8 mantissa bits, truncated;
// owners 0, others at least 0x100, finite keys at most 0xff00); the id is
// instance-local in 16 bits (instances of at most 65,536 sites), else
// batch-wide in 32 bits of a 64-bit entry (same order: local = batch-wide
// id - the instance's first site).  ~0 is the empty slot.
template <bool kWide>
struct PropEntry {
  using T = typename std::conditional<kWide, uint64_t, uint32_t>::type;
  static constexpr int kShift = kWide ? 32 : 16;
  static constexpr T kEmpty = ~T(0);
  __device__ __forceinline__ static uint32_t key16(float key) {
    return __float_as_uint(key) >> 15;
  }
  __device__ __forceinline__ static T pack(uint32_t k16, int32_t id,
                                           int32_t site_off) {
    return kWide ? (T)(((uint64_t)k16 << 32) | (uint32_t)id)
                 : (T)((k16 << 16) | (uint32_t)(id - site_off));
  }
  __device__ __forceinline__ static int32_t id(T e, int32_t site_off) {
    return kWide ? (int32_t)(uint32_t)e : (int32_t)(e & 0xffffu) + site_off;
  }
  __device__ __forceinline__ static float key(T e) {
    return __uint_as_float((uint32_t)(e >> kShift) << 15);
  }
};

// One thread per tet.  The owners of its centroid and of its 4 corners
// belong to its sphere set; the proposals are the union of their 1-rings
// (with them), ranked by key(s) = min over the 5 points x of
// pd(x, s) - pd(x, owner of x) -- 0 for the owners -- then id; 32 kept,
// flagged spheres only.  omask marks the owner slots.  Keys only rank
// proposals (the certificate makes the output exact whatever they are), so
// they are evaluated in binary32 around the centroid c: with u = s - c and
// the corner offsets d_x = x - c, pd(x, s) - pd_min(x) =
// (|u|^2 - w) - 2 d_x.u + (|d_x|^2 - pd_min(x)), and kept to 16 bits
// (PropEntry), so that a 32-entry sorted list of 32-bit words costs one
// median instruction per slot and insertion.
template <bool kWide>
__device__ __forceinline__ void knn_features(
    const DevMesh& M, const DevSites& S, const InstInfo* inst,
    const int32_t* tet_inst, const int32_t* vowner, int32_t* knn, uint16_t* X,
    uint32_t* omask) {
  using PE = PropEntry<kWide>;
  using E = typename PE::T;
  const int t = blockIdx.x * 64 + threadIdx.x;
  if (t >= M.n_tet) return;
  TetLoad L;
  load_tet(M, t, L);
  double C[3];
  tet_centroid(L, C);
  float D[4][3];
  float rho2f = 0.f;
  for (int k = 0; k < 4; ++k) {
    D[k][0] = (float)(L.V[k].x - C[0]);
    D[k][1] = (float)(L.V[k].y - C[1]);
    D[k][2] = (float)(L.V[k].z - C[2]);
    rho2f = fmaxf(rho2f, (D[k][0] * D[k][0] + D[k][1] * D[k][1]) +
                             D[k][2] * D[k][2]);
  }
  const InstInfo& I = inst[tet_inst[t]];
  const int32_t soff = I.site_off;
  uint16_t* x = X + (int64_t)t * kIn;
  int32_t* kt = knn + (int64_t)t * kOut;
  int own[5];
  own[1] = vowner[L.tv.x];
  own[2] = vowner[L.tv.y];
  own[3] = vowner[L.tv.z];
  own[4] = vowner[L.tv.w];
  // the 32 least entries, sorted, in registers.  A listed sphere comes back
  // with the same key (same operands, same operations), so an entry equal
  // to a listed one is a repeat and is dropped (made empty).  Inserting x
  // into the sorted a[0..31]: slot q becomes the median of a[q-1], x and
  // a[q] (a[q] if x >= a[q], x if a[q-1] <= x < a[q], else a[q-1]), slot 0
  // min(x, a[0]) -- each slot from old values only, so the slots update in
  // place from the last down, fully unrolled and without branches (an
  // empty or too large x leaves the list as it is).
  E lst[kOut];
#pragma unroll
  for (int q = 0; q < kOut; ++q) lst[q] = PE::kEmpty;
  // chk: the sphere may be listed already (entries of the second and later
  // rings; the owners and the first ring hold distinct ids)
  auto ins = [&](E x, bool chk) {
    if (chk) {
      bool dup = false;
#pragma unroll
      for (int q = 0; q < kOut; ++q) dup |= lst[q] == x;
      x = dup ? PE::kEmpty : x;
    }
    if (!(x < lst[kOut - 1])) return;  // past the last slot (or empty)
#pragma unroll
    for (int q = kOut - 1; q > 0; --q) lst[q] = slot_median(x, lst[q], lst[q - 1]);
    lst[0] = x < lst[0] ? x : lst[0];
  };
  uint32_t om = 0;
  if (own[1] >= 0) {
    // centroid owner: descend from the corner owner closest to it
    double pc[5];
    for (int k = 1; k < 5; ++k) pc[k] = pdist(C, S.site[own[k]]);
    int st = own[1];
    double best = pc[1];
    for (int k = 2; k < 5; ++k)
      if (pc[k] < best || (pc[k] == best && own[k] < st)) {
        best = pc[k];
        st = own[k];
      }
    own[0] = descend(S, C, st, &best);
    // q_x = |d_x|^2 - pd_min(x), binary32
    float qx[5];
    qx[0] = (float)(-best);
    for (int k = 1; k < 5; ++k) {
      const double4 so = S.site[own[k]];
      const double P[3] = {L.V[k - 1].x, L.V[k - 1].y, L.V[k - 1].z};
      const float dd = (D[k - 1][0] * D[k - 1][0] + D[k - 1][1] * D[k - 1][1]) +
                       D[k - 1][2] * D[k - 1][2];
      qx[k] = dd - (float)pdist(P, so);
    }
    // rings: the centroid owner's, and a corner owner's only if that
    // owner is not already in it (the rings overlap otherwise); which
    // corner owners the centroid owner's ring holds (flagged or not) is
    // noted while that ring is scanned
    uint32_t in_ring = 0;  // bit k: own[k] is in the centroid owner's ring
    int n_own = 0;         // distinct flagged owners listed (key 0)
    for (int j = 0; j < 5; ++j) {
      const int o = own[j];
      bool seen = false;
      for (int q = 0; q < j; ++q) seen = seen || own[q] == o;
      if (seen) continue;
      if (S.flags[o] == 1u) {
        ins(PE::pack(0u, o, soff), false);
        ++n_own;
      }
      if (j > 0 && ((in_ring >> j) & 1u)) continue;
      const int64_t e0 = S.knn_off[o], e1 = S.knn_off[o + 1];
      // kRing ring entries per round: their loads are in flight together
      constexpr int kRing = 2;
      for (int64_t q0 = e0; q0 < e1; q0 += kRing) {
        int32_t w4[kRing];
        double4 s4[kRing];
#pragma unroll
        for (int u = 0; u < kRing; ++u) {
          const int64_t q = q0 + u < e1 ? q0 + u : e1 - 1;
          w4[u] = S.knn_near[q];
          s4[u] = S.knn_site_near[q];
        }
#pragma unroll
        for (int u = 0; u < kRing; ++u) {
          if (q0 + u >= e1) continue;  // past the end
          if (j == 0) {
            const int nid = near_id(w4[u]);
#pragma unroll
            for (int a = 1; a < 5; ++a)
              in_ring |= (uint32_t)(own[a] == nid) << a;
          }
          if (w4[u] < 0) continue;  // unflagged
          const int id = w4[u];
          bool is_owner = false;
#pragma unroll
          for (int a = 0; a < 5; ++a) is_owner = is_owner || own[a] == id;
          if (is_owner) continue;  // listed with key 0
          const float ux = (float)(s4[u].x - C[0]);
          const float uy = (float)(s4[u].y - C[1]);
          const float uz = (float)(s4[u].z - C[2]);
          const float base =
              (float)(((s4[u].x - C[0]) * (s4[u].x - C[0]) +
                       (s4[u].y - C[1]) * (s4[u].y - C[1])) +
                      (s4[u].z - C[2]) * (s4[u].z - C[2]) - s4[u].w);
          float key = base + qx[0];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const float du = (D[k][0] * ux + D[k][1] * uy) + D[k][2] * uz;
            key = fminf(key, (base + qx[1 + k]) - 2.f * du);
          }
          // non-owners rank after the owners (key 0) whatever the rounding
          ins(PE::pack(PE::key16(fmaxf(key, 1.17549435e-38f)), id, soff),
              j > 0);
        }
      }
    }
    // only owners have key 0 (the others are clamped above it) and none
    // is ever pushed out, so they hold the first n_own slots
    om = n_own >= 32 ? ~0u : (1u << n_own) - 1u;
  }
  omask[t] = om;
  auto id_at = [&](int q) {
    return lst[q] != PE::kEmpty ? PE::id(lst[q], soff) : -1;
  };
#pragma unroll
  for (int k = 0; k < kOut; k += 4)  // 16-byte stores of the row
    *reinterpret_cast<int4*>(kt + k) =
        make_int4(id_at(k), id_at(k + 1), id_at(k + 2), id_at(k + 3));
  const double ext =
      fmax(fmax(I.dom[3] - I.dom[0], I.dom[4] - I.dom[1]), I.dom[5] - I.dom[2]);
  const double inv_ext = ext > 0 ? 1.0 / ext : 0.0;
  const float inv_r = rho2f > 0.f ? 1.f / sqrtf(rho2f) : 0.f;
  float f[kIn];
#pragma unroll
  for (int q = 0; q < kIn; ++q) f[q] = 0.f;
  f[0] = (float)((C[0] - I.dom[0]) * inv_ext);
  f[1] = (float)((C[1] - I.dom[1]) * inv_ext);
  f[2] = (float)((C[2] - I.dom[2]) * inv_ext);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (lst[k] != PE::kEmpty) {
      const double4 sp = S.site[PE::id(lst[k], soff)];
      f[3 + 5 * k + 0] = (float)(sp.x - C[0]) * inv_r;
      f[3 + 5 * k + 1] = (float)(sp.y - C[1]) * inv_r;
      f[3 + 5 * k + 2] = (float)(sp.z - C[2]) * inv_r;
      f[3 + 5 * k + 3] = (float)sqrt(sp.w) * inv_r;
      f[3 + 5 * k + 4] = PE::key(lst[k]) * (inv_r * inv_r);
    }
  }
#pragma unroll
  for (int q = 0; q < kIn; q += 8) {  // 16-byte stores of the row
    uint4 v;
    v.x = (uint32_t)f2bf(f[q]) | ((uint32_t)f2bf(f[q + 1]) << 16);
    v.y = (uint32_t)f2bf(f[q + 2]) | ((uint32_t)f2bf(f[q + 3]) << 16);
    v.z = (uint32_t)f2bf(f[q + 4]) | ((uint32_t)f2bf(f[q + 5]) << 16);
    v.w = (uint32_t)f2bf(f[q + 6]) | ((uint32_t)f2bf(f[q + 7]) << 16);
    *reinterpret_cast<uint4*>(x + q) = v;
  }
}
// 16-bit instance-local ids: compiled for 5 resident wavefronts per SIMD
// (its few spills sit outside the ring loop); the 64-bit entries of larger
// instances for 4
__global__ __launch_bounds__(64, 5) void k_knn_features(
    DevMesh M, DevSites S, const InstInfo* inst, const int32_t* tet_inst,
    const int32_t* vowner, int32_t* knn, uint16_t* X, uint32_t* omask) {
  knn_features<false>(M, S, inst, tet_inst, vowner, knn, X, omask);
}
__global__ __launch_bounds__(64, 4) void k_knn_features_wide(
    DevMesh M, DevSites S, const InstInfo* inst, const int32_t* tet_inst,
    const int32_t* vowner, int32_t* knn, uint16_t* X, uint32_t* omask) {
  knn_features<true>(M, S, inst, tet_inst, vowner, knn, X, omask);
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

// Inference: selection mask per row (bit k: logit_k > thr); optional
// logits.  Eight waves share the weights, staged once per block in LDS
// (rows padded against bank conflicts); each wave runs kInferTiles 16-row
// tiles at a time, so that every weight fragment read from LDS feeds
// kInferTiles independent MFMA chains.
constexpr int kS1 = kIn + 8, kS2 = kHid + 8;
constexpr int kInferWaves = 8;
constexpr int kInferTiles = 2;

// The layers run transposed (Z^T = W^T X^T: weights as the A operand,
// activations as the B operand), so a layer's output fragment in a lane --
// 4 consecutive features of tet l % 16 -- is exactly the next layer's B
// fragment in the same lane: activations stay in registers, and only the
// weights live in LDS (58 KB, two blocks per CU).
__global__ __launch_bounds__(64 * kInferWaves) void k_mlp_infer(
    MlpDev P, const uint16_t* X, int64_t n, float thr, uint32_t* sel,
    float* logits) {
  constexpr int T = kInferTiles;
  __shared__ uint16_t s_w1[kHid * kS1];
  __shared__ uint16_t s_w2[kHid * kS2];
  __shared__ uint16_t s_w3[kOut * kS2];
  for (int q = threadIdx.x; q < kHid * kIn; q += 64 * kInferWaves)
    s_w1[(q / kIn) * kS1 + q % kIn] = P.w1t[q];
  for (int q = threadIdx.x; q < kHid * kHid; q += 64 * kInferWaves)
    s_w2[(q / kHid) * kS2 + q % kHid] = P.w2t[q];
  for (int q = threadIdx.x; q < kOut * kHid; q += 64 * kInferWaves)
    s_w3[(q / kHid) * kS2 + q % kHid] = P.w3t[q];
  __syncthreads();
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = l % 16, rq = 4 * (l / 16);
  const int64_t n_tiles = (n + 15) / 16;
  // lane l: features 16 nb + rq + i of tet row0 + col
  auto act = [](f32x4 c, const float* b) {
    const u32x2 w = {f2bf2(fmaxf(c[0] + b[0], 0.f), fmaxf(c[1] + b[1], 0.f)),
                     f2bf2(fmaxf(c[2] + b[2], 0.f), fmaxf(c[3] + b[3], 0.f))};
    return __builtin_bit_cast(bf16x4, w);
  };
  for (int64_t tile0 = ((int64_t)blockIdx.x * kInferWaves + wv) * T;
       tile0 < n_tiles; tile0 += (int64_t)gridDim.x * kInferWaves * T) {
    // the biases are read where they are used (L1 hits): an empty asm
    // statement hides their addresses' invariance, so the compiler does not
    // keep 72 of them in registers across the loop (it spilled them)
    const float *b1 = P.b1, *b2 = P.b2, *b3 = P.b3;
    asm volatile("" : "+s"(b1), "+s"(b2), "+s"(b3));
    int64_t my_row[T];
    bf16x4 x[T][3];
#pragma unroll
    for (int t = 0; t < T; ++t) {
      my_row[t] = (tile0 + t) * 16 + col;
      load_x(X, my_row[t] < n ? my_row[t] : -1, l, x[t]);
    }
    bf16x4 h1[T][8];
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
      f32x4 acc[T];
#pragma unroll
      for (int t = 0; t < T; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 3; ++kb) {
        const bf16x4 w = ld_bf4(&s_w1[(16 * nb + col) * kS1 + 16 * kb + rq]);
#pragma unroll
        for (int t = 0; t < T; ++t) acc[t] = mfma(w, x[t][kb], acc[t]);
      }
#pragma unroll
      for (int t = 0; t < T; ++t) h1[t][nb] = act(acc[t], b1 + 16 * nb + rq);
    }
    bf16x4 h2[T][8];
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
      f32x4 acc[T];
#pragma unroll
      for (int t = 0; t < T; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 8; ++kb) {
        const bf16x4 w = ld_bf4(&s_w2[(16 * nb + col) * kS2 + 16 * kb + rq]);
#pragma unroll
        for (int t = 0; t < T; ++t) acc[t] = mfma(w, h1[t][kb], acc[t]);
      }
#pragma unroll
      for (int t = 0; t < T; ++t) h2[t][nb] = act(acc[t], b2 + 16 * nb + rq);
    }
    f32x4 z[T][2];
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
      f32x4 acc[T];
#pragma unroll
      for (int t = 0; t < T; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 8; ++kb) {
        const bf16x4 w = ld_bf4(&s_w3[(16 * nb + col) * kS2 + 16 * kb + rq]);
#pragma unroll
        for (int t = 0; t < T; ++t) acc[t] = mfma(w, h2[t][kb], acc[t]);
      }
#pragma unroll
      for (int t = 0; t < T; ++t) {
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[t][i] += b3[16 * nb + rq + i];
        z[t][nb] = acc[t];
      }
    }
#pragma unroll
    for (int t = 0; t < T; ++t) {
      // proposal mask of tet my_row: this lane's 8 bits, OR over the four
      // lanes of the tet (l, l ^ 16, l ^ 32, l ^ 48)
      uint32_t m = 0;
#pragma unroll
      for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          m |= (uint32_t)(z[t][nb][i] > thr) << (16 * nb + rq + i);
      m |= (uint32_t)__shfl_xor((int)m, 16);
      m |= (uint32_t)__shfl_xor((int)m, 32);
      if (l < 16 && my_row[t] < n) sel[my_row[t]] = m;
      if (logits && my_row[t] < n) {
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
#pragma unroll
          for (int i = 0; i < 4; ++i)
            logits[my_row[t] * kOut + 16 * nb + rq + i] = z[t][nb][i];
      }
    }
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
// per 16x16 output block, blockIdx.y = chunk.  The waves of the first input
// block also write the chunk's column sums of DZ (the bias gradient
// partials) to part_b[chunk][b_off + out].
constexpr int kBiasN = 2 * kHid + kOut;
__global__ __launch_bounds__(64) void k_mlp_wgrad(const uint16_t* A,
                                                  const uint16_t* DZ,
                                                  int32_t in, int32_t out,
                                                  int32_t tiles_per_chunk,
                                                  int32_t n_tiles,
                                                  float* part, float* part_b,
                                                  int32_t b_off) {
  const int l = threadIdx.x;
  const int nbo = out / 16;
  const int mb = blockIdx.x / nbo, nb = blockIdx.x % nbo;
  const int chunk = blockIdx.y;
  const int col = l % 16, rq = 4 * (l / 16);
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  const int t0 = chunk * tiles_per_chunk;
  const int t1 = min(n_tiles, t0 + tiles_per_chunk);
  for (int tile = t0; tile < t1; ++tile) {
    const int64_t r0 = (int64_t)tile * 16 + rq;
    bf16x4 a, b;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a[i] = (short)A[(r0 + i) * in + 16 * mb + col];
      b[i] = (short)DZ[(r0 + i) * out + 16 * nb + col];
      bsum += bf2f((uint16_t)b[i]);
    }
    acc = mfma(a, b, acc);
  }
  float* p = part + (int64_t)chunk * in * out;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    p[(int64_t)(16 * mb + rq + i) * out + 16 * nb + col] = acc[i];
  if (mb == 0) {
    bsum += __shfl_xor(bsum, 16);
    bsum += __shfl_xor(bsum, 32);
    if (l < 16) part_b[(int64_t)chunk * kBiasN + b_off + 16 * nb + col] = bsum;
  }
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
    const float* part2, const float* part3, const float* part_b,
    int32_t n_chunks, AdamArgs a, float* grad_out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNParam) return;
  const float* src;
  int64_t stride;
  if (p < kOffB1) {
    src = part1 + p;
    stride = kIn * kHid;
  } else if (p < kOffW2) {
    src = part_b + (p - kOffB1);
    stride = kBiasN;
  } else if (p < kOffB2) {
    src = part2 + (p - kOffW2);
    stride = kHid * kHid;
  } else if (p < kOffW3) {
    src = part_b + kHid + (p - kOffB2);
    stride = kBiasN;
  } else if (p < kOffB3) {
    src = part3 + (p - kOffW3);
    stride = kHid * kOut;
  } else {
    src = part_b + 2 * kHid + (p - kOffB3);
    stride = kBiasN;
  }
  float g = 0.f;
  for (int c = 0; c < n_chunks; ++c) g += src[(int64_t)c * stride];
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

// selected proposals (owners always) that pass the reference relation
// (voronoi.cu:154-193): their slot mask replaces sel, their count is
// written; only such spheres are clipped, as on the plain path
__global__ __launch_bounds__(64) void k_learn_count(int32_t n_tet, DevMesh M,
                                                    DevSites S,
                                                    const int32_t* knn,
                                                    uint32_t* sel,
                                                    const uint32_t* omask,
                                                    int64_t* count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tet) return;
  TetLoad L;
  load_tet(M, t, L);
  const uint32_t s = sel[t] | omask[t];
  uint32_t pass = 0;
  int n = 0;
  // only the selected slots are read (slots past the last proposal hold -1)
  for (uint32_t rest = s; rest; rest &= rest - 1) {
    const int k = __ffs(rest) - 1;
    const int id = knn[(int64_t)t * kOut + k];
    if (id < 0) break;
    uint32_t it = 0;
    if (relation(S, L, id, &it)) {
      pass |= 1u << k;
      ++n;
    }
  }
  sel[t] = pass;
  count[t] = n;
}

// (a captured run: the candidate lists hold cap entries; a tet whose
// candidates pass them sets *overflow and writes those below cap, so that
// every slot a later launch reads holds a pair of this run)
__global__ __launch_bounds__(64) void k_learn_fill(int32_t n_tet,
                                                   const int32_t* knn,
                                                   const uint32_t* sel,
                                                   const int64_t* off,
                                                   Cand* cand,
                                                   int32_t* rel_ids,
                                                   int64_t cap,
                                                   int32_t* overflow) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  // (no early return: the whole wavefront takes part in the reduction and,
  // on the small-set path, in the staging below)
  const bool live = t < n_tet;
  const uint32_t m = live ? sel[t] : 0u;
  const int n = __popc(m);
  int n_wave = n;  // the wavefront's largest set
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) n_wave = max(n_wave, __shfl_xor(n_wave, d));
  if (n_wave <= 8) {
    // small sets (the usual case): ids inserted one by one into 8 sorted
    // registers, each slot updated in place to the median of its
    // predecessor, the new id and itself; the block's lists lie back to
    // back from off[t0], so they are staged in LDS and written out by the
    // whole block as one contiguous run of records
    __shared__ int32_t s_seed[64 * 8];
    __shared__ int32_t s_tet[64 * 8];
    int32_t w[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) w[k] = 0x7fffffff;
    for (uint32_t rest = m; rest; rest &= rest - 1) {
      const int32_t y = knn[(int64_t)t * kOut + (__ffs(rest) - 1)];
      const int32_t x = y >= 0 ? y : 0x7fffffff;
#pragma unroll
      for (int k = 7; k > 0; --k) w[k] = slot_median(x, w[k], w[k - 1]);
      w[0] = min(x, w[0]);
    }
    const int t0 = blockIdx.x * blockDim.x;  // < n_tet: the block exists
    const int64_t o0 = off[t0];
    const int64_t total = off[min(t0 + (int)blockDim.x, n_tet)] - o0;
    if (live) {
      const int64_t o = off[t];
      if (o + n > cap) *overflow = 1;
      const int lo = (int)(o - o0);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < n) {
          s_seed[lo + k] = w[k];
          s_tet[lo + k] = t;
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
      const int64_t dst = o0 + e;
      if (dst < cap) {
        Cand c;
        c.tet = s_tet[e];
        c.seed = s_seed[e];
        c.m0 = 0;
        c.m1 = 0;
        cand[dst] = c;
        rel_ids[dst] = c.seed;
      }
    }
    return;
  }
  if (!live) return;
  const int64_t o = off[t];
  if (o + n > cap) *overflow = 1;
  // selected ids ascending: bitonic sorting network on 32 registers
  int32_t v[kOut];
#pragma unroll
  for (int k = 0; k < kOut; ++k) {
    int32_t id = 0x7fffffff;
    if ((m >> k) & 1u) {  // only the selected slots are read
      const int32_t x = knn[(int64_t)t * kOut + k];
      if (x >= 0) id = x;
    }
    v[k] = id;
  }
#pragma unroll
  for (int size = 2; size <= kOut; size <<= 1)
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1)
#pragma unroll
      for (int i = 0; i < kOut; ++i) {
        const int j = i ^ stride;
        if (j > i) {
          const bool up = (i & size) == 0;
          const int32_t a = v[i], b = v[j];
          const bool sw = up ? a > b : a < b;
          v[i] = sw ? b : a;
          v[j] = sw ? a : b;
        }
      }
#pragma unroll
  for (int k = 0; k < kOut; ++k) {
    if (k < n && o + k < cap) {
      Cand c;
      c.tet = t;
      c.seed = v[k];
      c.m0 = 0;
      c.m1 = 0;
      cand[o + k] = c;
      rel_ids[o + k] = v[k];
    }
  }
}

// tets of missing pairs left after the last closure round
__global__ __launch_bounds__(256) void k_flag_miss(int64_t n,
                                                   const uint64_t* miss,
                                                   uint8_t* flag,
                                                   const int32_t* n_dev) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t m = n_dev ? min(n, (int64_t)*n_dev) : n;
  if (q < m) flag[miss[q] >> 32] = 1;
}

// sorted keys -> first occurrences (1) / repeats (0); with n_dev only the
// first min(*n_dev, n) keys count (a captured run sorts a padded list)
__global__ __launch_bounds__(256) void k_unique_mark(int64_t n,
                                                     const uint64_t* keys,
                                                     int32_t* mark,
                                                     const int32_t* n_dev) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q > n) return;
  const int64_t m = n_dev ? min(n, (int64_t)*n_dev) : n;
  mark[q] = q < m && (q == 0 || keys[q] != keys[q - 1]) ? 1 : 0;
}

// unique missing pairs: the relation decides which become candidates
// (appended at cand_base in key order); all become members
__global__ __launch_bounds__(256) void k_closure_split(
    int64_t n, const uint64_t* keys, const int32_t* mark, const int32_t* pos,
    DevMesh M, DevSites S, uint64_t* ukeys, uint8_t* pass) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n || !mark[q]) return;
  const uint64_t k = keys[q];
  const int64_t u = pos[q];
  ukeys[u] = k;
  TetLoad L;
  load_tet(M, (int)(k >> 32), L);
  uint32_t it = 0;
  pass[u] = relation(S, L, (int)(k & 0xffffffffu), &it) ? 1 : 0;
}

// (a captured run: n_u on the device, at most cap new candidates)
__global__ __launch_bounds__(256) void k_closure_fill(
    int64_t n_u, const uint64_t* ukeys, const uint8_t* pass,
    const int32_t* ppos, Cand* cand, int64_t cand_base, uint8_t* reopened,
    const int32_t* n_u_dev, int64_t cap, int32_t* overflow) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n_u || (n_u_dev && u >= *n_u_dev) || !pass[u]) return;
  if (ppos[u] >= cap) {
    *overflow = 1;
    return;
  }
  reopened[ukeys[u] >> 32] = 1;
  Cand c;
  c.tet = (int)(ukeys[u] >> 32);
  c.seed = (int)(ukeys[u] & 0xffffffffu);
  c.m0 = 0;
  c.m1 = 0;
  cand[cand_base + ppos[u]] = c;
}

// out[q] = p[q] for q < *n_dev, 0 for *n_dev <= q <= n (n >= *n_dev: the
// host sizes the launch by an upper bound of the device count)
__global__ __launch_bounds__(256) void k_pass32(int64_t n,
                                                const int32_t* n_dev,
                                                const uint8_t* p,
                                                int32_t* out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q <= n) out[q] = q < *n_dev ? p[q] : 0;
}

// final flags (failed, dropped or cell-less tets) and the count of tets
// re-run in any way, added per block to one of 64 counters (blocks
// spread their atomics over them)
__global__ __launch_bounds__(256) void k_flag_final(int32_t n_tet,
                                                    uint8_t* flag,
                                                    const uint8_t* has_cell,
                                                    const uint8_t* reopened,
                                                    int32_t* flag32,
                                                    int32_t* n_reopened64) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  bool re = false;
  if (t < n_tet) {
    const uint8_t f = flag[t] || !has_cell[t];
    flag[t] = f;
    flag32[t] = f;
    re = f || reopened[t];
  }
  const int c = __syncthreads_count(re);
  if (threadIdx.x == 0 && c) atomicAdd(&n_reopened64[blockIdx.x & 63], c);
}

// tets re-run in any way (closure or full relation)
__global__ __launch_bounds__(256) void k_flag_list(int32_t n_tet,
                                                   const int32_t* flag32,
                                                   const int32_t* pos,
                                                   int32_t* list, int64_t cap,
                                                   int32_t* overflow) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tet || !flag32[t]) return;
  if (pos[t] >= cap) {
    *overflow = 1;
    return;
  }
  list[pos[t]] = t;
}

// Fallback: the reference relation of each listed tet against every sphere
// of its instance; one wave per tet, sites in ascending order.
// (a captured run: the list length on the device, candidates written
// below cand_end only)
__global__ __launch_bounds__(64) void k_brute(
    int32_t n_list, const int32_t* list, DevMesh M, DevSites S,
    const InstInfo* inst, const int32_t* tet_inst, const int64_t* off,
    int64_t* count, Cand* cand, int64_t cand_base, const int32_t* n_list_dev,
    int64_t cand_end) {
  const int f = blockIdx.x;
  if (f >= n_list || (n_list_dev && f >= *n_list_dev)) return;
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
    const int64_t at = o + n + __popcll(b & ((1ull << lane) - 1));
    if (off && r && at < cand_end) {
      Cand c;
      c.tet = t;
      c.seed = s;
      c.m0 = 0;
      c.m1 = 0;
      cand[at] = c;
    }
    n += __popcll(b);
  }
  if (count && lane == 0) count[f] = n;
}

// Output order.  A = first-round candidates [0, n1) in (tet, id) order
// (CSR a_off / a_ids); B = the candidates added after, sorted by key.  A
// candidate of a tet re-run on the full relation (flag) is left out, as
// are B entries before n_brute of such tets.  Each tet's output is A and B
// merged by id.
__global__ __launch_bounds__(256) void k_b_keys(int64_t n1, int64_t n_all,
                                                int64_t n_brute,
                                                const Cand* cand,
                                                const uint8_t* flag,
                                                uint64_t* keys, int32_t* vals) {
  const int64_t c = n1 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_all) return;
  const Cand cd = cand[c];
  const bool out = c < n_brute && flag[cd.tet];
  keys[c - n1] =
      out ? ~0ull : (((uint64_t)(uint32_t)cd.tet << 32) | (uint32_t)cd.seed);
  vals[c - n1] = (int32_t)c;
}

__device__ __forceinline__ int64_t lower_u64(const uint64_t* keys, int64_t n,
                                             uint64_t k) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < k)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_out_counts(
    int32_t n_tet, const uint8_t* flag, const uint8_t* reopened,
    const int64_t* a_off, const uint64_t* bkeys, int64_t n_b,
    int64_t* count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > n_tet) return;
  if (t == n_tet) {
    count[t] = 0;
    return;
  }
  const int64_t na = flag[t] ? 0 : a_off[t + 1] - a_off[t];
  int64_t nb = 0;
  if (flag[t] || reopened[t]) {  // only such tets have added candidates
    const uint64_t k0 = (uint64_t)(uint32_t)t << 32;
    nb = lower_u64(bkeys, n_b, k0 + (1ull << 32)) - lower_u64(bkeys, n_b, k0);
  }
  count[t] = na + nb;
}

// (a captured run: n1 on the device; positions at or past perm_cap, which
// the host then finds over capacity, are not written)
__global__ __launch_bounds__(256) void k_place_a(
    int64_t n1, const Cand* cand, const uint8_t* flag,
    const uint8_t* reopened, const int64_t* a_off, const uint64_t* bkeys,
    int64_t n_b, const int64_t* out_off, int32_t* perm, const int64_t* n1_dev,
    int64_t perm_cap) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n1 || (n1_dev && c >= *n1_dev)) return;
  const Cand cd = cand[c];
  if (flag[cd.tet]) return;
  int64_t before_b = 0;
  if (reopened[cd.tet]) {  // only such tets have added candidates
    const uint64_t k0 = (uint64_t)(uint32_t)cd.tet << 32;
    before_b = lower_u64(bkeys, n_b, k0 | (uint32_t)cd.seed) -
               lower_u64(bkeys, n_b, k0);
  }
  const int64_t at = out_off[cd.tet] + (c - a_off[cd.tet]) + before_b;
  if (at < perm_cap) perm[at] = (int32_t)c;
}

__global__ __launch_bounds__(256) void k_place_b(
    int64_t n_b, const uint64_t* bkeys, const int32_t* bvals,
    const uint8_t* flag, const int64_t* a_off, const int32_t* a_ids,
    const int64_t* out_off, int32_t* perm, int64_t perm_cap) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_b) return;
  const uint64_t k = bkeys[r];
  if (k == ~0ull) return;  // excluded (sorted after every real key)
  const int t = (int)(k >> 32);
  const int32_t j = (int32_t)(k & 0xffffffffu);
  const int64_t start = lower_u64(bkeys, n_b, (uint64_t)(uint32_t)t << 32);
  int64_t before_a = 0;
  if (!flag[t]) {
    int64_t lo = a_off[t], hi = a_off[t + 1];
    const int64_t base = lo;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (a_ids[mid] < j)
        lo = mid + 1;
      else
        hi = mid;
    }
    before_a = lo - base;
  }
  const int64_t at = out_off[t] + (r - start) + before_a;
  if (at < perm_cap) perm[at] = bvals[r];
}

// relation CSR ids of the output order (a captured run: the count on the
// device)
__global__ __launch_bounds__(256) void k_out_ids(int64_t n, const Cand* cand,
                                                 const int32_t* perm,
                                                 int32_t* ids,
                                                 const int64_t* n_dev) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n && (!n_dev || r < *n_dev)) ids[r] = cand[perm[r]].seed;
}

// ------------------------------------------ captured learned pipeline
// Small helpers of the captured learned run, whose counts stay on the
// device: lists are sized by capacities and padded, and every count that
// passes its capacity sets *overflow (the host then runs eagerly and
// captures again).

// a range's count n_src (clamped to cap, *overflow past it) and its end
// base + count
template <typename T>
__global__ void k_range_count(const T* n_src, int64_t base, int64_t cap,
                              int64_t* rel, int64_t* abs_end,
                              int32_t* overflow) {
  if (threadIdx.x != 0) return;
  int64_t n = (int64_t)*n_src;
  if (n > cap) {
    *overflow = 1;
    n = cap;
  }
  *rel = n;
  *abs_end = base + n;
}

// The end of a captured closure round: its new candidates' count *n_new
// (clamped to cap_new) and end base + count into rel / abs_end, and the
// merged pair list ext += the round's n_u unique pairs (ukeys) with ~0 past
// them, its length into *n_ext_out (*overflow past cap_ext or cap_new).
__global__ __launch_bounds__(256) void k_round_close(
    const int32_t* n_new, int64_t base, int64_t cap_new, int64_t* rel,
    int64_t* abs_end, uint64_t* ext, int64_t cap_ext, const uint64_t* ukeys,
    const int32_t* n_u, const int64_t* n_ext_in, int64_t* n_ext_out,
    int32_t* overflow) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ne = *n_ext_in, nu = *n_u;
  if (i == 0) {
    int64_t n = *n_new;
    if (n > cap_new) {
      *overflow = 1;
      n = cap_new;
    }
    *rel = n;
    *abs_end = base + n;
    int64_t tot = ne + nu;
    if (tot > cap_ext) {
      *overflow = 1;
      tot = cap_ext;
    }
    *n_ext_out = tot;
  }
  if (i >= cap_ext || i < ne) return;
  ext[i] = i < ne + nu ? ukeys[i - ne] : ~0ull;
}

// after an overflow (whose results are discarded) the first-round CSR is
// cut to the capacity, so that no later launch reads past its lists
__global__ __launch_bounds__(256) void k_clamp_offsets(int64_t* off,
                                                       int64_t n, int64_t cap,
                                                       const int32_t* overflow) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t <= n && *overflow) off[t] = min(off[t], cap);
}

// A captured miss list of nm misses from a launch of *cap_n candidates: the
// eager path's list holds max(4096, 2 * *cap_n) of them, this one
// lim = min(cap, that); *n_eff = min(nm, lim) are listed (the rest of
// their tets were sent to the full relation by the certificate, as the
// eager path does), entries past them are padded with ~0, and *overflow
// is set (with overflow non-null) when this list cut misses the eager one
// would hold.
__global__ __launch_bounds__(256) void k_pad_miss(uint64_t* p, int64_t cap,
                                                  const int32_t* nm,
                                                  const int64_t* cap_n,
                                                  int32_t* n_eff,
                                                  int32_t* overflow) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t capc = max((int64_t)4096, 2 * *cap_n);
  const int64_t m = min((int64_t)*nm, min(cap, capc));
  if (i == 0) {
    *n_eff = (int32_t)m;
    if (overflow && (int64_t)*nm > cap && capc > cap) *overflow = 1;
  }
  if (i < cap && i >= m) p[i] = ~0ull;
}

// *overflow = 1 when *n > 0 (misses left after the last captured round)
__global__ void k_flag_if_positive(const int32_t* n, int32_t* overflow) {
  if (threadIdx.x == 0 && *n > 0) *overflow = 1;
}

// Candidate ranges of a captured learned run: closure round r at
// [base[r], base[r + 1]), the fallback from base[n] on.
struct RegionTable {
  int64_t base[8];
  int32_t n;
};

// Output keys of the added candidates of a captured run: closure round r
// holds its candidates at [base[r], + rel[2 r]), the fallback at
// [base[n], + rel[2 n]) (rel: the device counts); slots past a count are
// excluded (~0), and so are closure candidates of tets re-run on the full
// relation, as in k_b_keys.
__global__ __launch_bounds__(256) void k_b_keys_ranges(
    RegionTable reg, int64_t end, const int64_t* rel, const Cand* cand,
    const uint8_t* flag, uint64_t* keys, int32_t* vals) {
  const int64_t c = reg.base[0] + (int64_t)blockIdx.x * blockDim.x +
                    threadIdx.x;
  if (c >= end) return;
  int r = 0;
  while (r < reg.n && c >= reg.base[r + 1]) ++r;
  const bool closure = r < reg.n;
  const bool valid = c - reg.base[r] < rel[2 * r];
  uint64_t k = ~0ull;
  if (valid) {
    const Cand cd = cand[c];
    if (!(closure && flag[cd.tet]))
      k = ((uint64_t)(uint32_t)cd.tet << 32) | (uint32_t)cd.seed;
  }
  keys[c - reg.base[0]] = k;
  vals[c - reg.base[0]] = (int32_t)c;
}

}  // namespace rpd
