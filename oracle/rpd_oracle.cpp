// ORACLE -- test infrastructure only.
//
// A plain serial CPU restatement of LibMAT's restricted-power-diagram (RPD)
// hot path, used by tests/, by __graft_entry__.smoke() and by the
// cpu_baseline leg of bench.py as the checker.  The product never links,
// loads or calls this file.
//
// What it restates (reference file:line):
//  * tet-sphere relation (tet_sphere_relations_dev, src/rpd3d/voronoi.cu:
//    154-193): sphere i relates to tet t iff i is flagged and for every
//    neighbour m of i some tet corner v has pd(v,m) > pd(v,i);
//    tet_knn lists are ascending site ids (voronoi.cu:284-317);
//  * per-(tet, seed) convex-cell clip (clipped_voro_cell_test_GPU_param_tet,
//    src/rpd3d/convex_cell.cu:1166-1337): ConvexCell ctor (:116-214),
//    new_plane (:561-592), conflict test (:437-500), vertex swap loop
//    (:706-721), compute_boundary (:618-678), new_edge (:604-616),
//    new_vertex/find_edge_id (:511-554), is_vertex_perturb (:274-316);
//    capacities _MAX_P_ 64, _MAX_T_ 96, _MAX_E_ 152
//    (src/rpd3d_base/voronoi_common.h:56-63); any non-success status drops
//    the cell (:1279-1283, voronoi.cu:749).
//
// Two arithmetic modes:
//  * mode 0 ("X", canonical): planes in binary64 from binary64 inputs; the
//    conflict sign det4x4(p1,p2,p3,e) > 0 is decided exactly (semi-static
//    filter with its own bound, then exact big-integer determinant);
//    needs_perturb iff the normals' 3x3 determinant is exactly 0.
//  * mode 1 ("R", reference arithmetic): inputs rounded to binary32, planes
//    in binary32, det4x4 in binary64 from the float coefficients with the
//    float-equality shortcut, no filter; needs_perturb iff the binary32
//    3x3 determinant is 0.  Used for the CPU baseline and to report how
//    often float arithmetic disagrees with exact signs.
//
// Geometry (volume, centroid, facet areas) is computed in binary64 from the
// mode-X planes by an independent method: each active face is ordered into
// its vertex cycle (voronoi_defs.cxx:108-188 logic), facet areas by Newell's
// formula, volume and centroid by a fan of tetrahedra from the first
// vertex.
//
// What pins each part (tests/test_oracle_pins.py):
//  * exact_det4_sign: rational arithmetic (fractions.Fraction) on random,
//    rank-deficient and near-degenerate matrices; the reference's own exact
//    expansion determinant (src/CPU/Delaunay_psm.cpp:9539) built into
//    oracle/_ref from the mounted sources;
//  * planes, bisectors, related(): closed forms (one sphere owns every tet;
//    two spheres split at their power plane with analytic volumes; eight
//    octant spheres whose bisectors lie on grid planes) and brute force
//    (clipping by every sphere equals clipping by the neighbour lists);
//  * the clip and cell_geometry: the partition of every tet (volumes sum to
//    the tet volume), and the reference's own FP64 VBW::ConvexCell
//    (src/CPU/convex_cell.cpp) clipping the same planes;
//  * mode 1 (reference float arithmetic): parity unpinned -- the binary32
//    bits of the CUDA reference cannot be produced here; it only serves as
//    the CPU baseline and to count disagreements with mode 0.
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr int kMaxP = 64, kMaxT = 96, kMaxE = 152;
constexpr uint8_t kEnd = 255;

enum Status {
  st_triangle_overflow = 0,
  st_vertex_overflow = 1,
  st_inconsistent_boundary = 2,
  st_success = 4,
  st_no_intersection = 6,
  st_edge_overflow = 7,
  st_needs_perturb = 8,
};

// ----------------------------------------------------- big integers
struct Big {
  int sign = 0;                 // -1, 0, +1
  std::vector<uint32_t> mag;    // little endian

  static Big from_shifted(int64_t m, int shift) {
    Big b;
    if (m == 0) return b;
    b.sign = m < 0 ? -1 : 1;
    uint64_t a = (uint64_t)(m < 0 ? -m : m);
    int limbs = shift / 32, bits = shift % 32;
    b.mag.assign(limbs, 0);
    unsigned __int128 v = (unsigned __int128)a << bits;
    while (v) {
      b.mag.push_back((uint32_t)v);
      v >>= 32;
    }
    return b;
  }
  static int cmp_mag(const Big& a, const Big& b) {
    if (a.mag.size() != b.mag.size())
      return a.mag.size() < b.mag.size() ? -1 : 1;
    for (size_t i = a.mag.size(); i-- > 0;)
      if (a.mag[i] != b.mag[i]) return a.mag[i] < b.mag[i] ? -1 : 1;
    return 0;
  }
  void trim() {
    while (!mag.empty() && mag.back() == 0) mag.pop_back();
    if (mag.empty()) sign = 0;
  }
  static Big mul(const Big& a, const Big& b) {
    Big r;
    if (!a.sign || !b.sign) return r;
    r.sign = a.sign * b.sign;
    r.mag.assign(a.mag.size() + b.mag.size(), 0);
    for (size_t i = 0; i < a.mag.size(); ++i) {
      uint64_t carry = 0;
      for (size_t j = 0; j < b.mag.size(); ++j) {
        uint64_t cur = (uint64_t)a.mag[i] * b.mag[j] + r.mag[i + j] + carry;
        r.mag[i + j] = (uint32_t)cur;
        carry = cur >> 32;
      }
      size_t k = i + b.mag.size();
      while (carry) {
        uint64_t cur = (uint64_t)r.mag[k] + carry;
        r.mag[k] = (uint32_t)cur;
        carry = cur >> 32;
        ++k;
      }
    }
    r.trim();
    return r;
  }
  static Big add(const Big& a, const Big& b) {
    if (!a.sign) return b;
    if (!b.sign) return a;
    Big r;
    if (a.sign == b.sign) {
      r.sign = a.sign;
      size_t n = std::max(a.mag.size(), b.mag.size());
      r.mag.assign(n + 1, 0);
      uint64_t carry = 0;
      for (size_t i = 0; i < n; ++i) {
        uint64_t cur = carry;
        if (i < a.mag.size()) cur += a.mag[i];
        if (i < b.mag.size()) cur += b.mag[i];
        r.mag[i] = (uint32_t)cur;
        carry = cur >> 32;
      }
      r.mag[n] = (uint32_t)carry;
      r.trim();
      return r;
    }
    int c = cmp_mag(a, b);
    if (c == 0) return r;
    const Big& big = c > 0 ? a : b;
    const Big& sml = c > 0 ? b : a;
    r.sign = big.sign;
    r.mag.assign(big.mag.size(), 0);
    int64_t borrow = 0;
    for (size_t i = 0; i < big.mag.size(); ++i) {
      int64_t cur = (int64_t)big.mag[i] - borrow -
                    (i < sml.mag.size() ? (int64_t)sml.mag[i] : 0);
      borrow = cur < 0;
      if (cur < 0) cur += (1LL << 32);
      r.mag[i] = (uint32_t)cur;
    }
    r.trim();
    return r;
  }
};

// exact sign of det of a 4x4 matrix of doubles a[r][c] (Leibniz formula on
// row-scaled integer matrices)
int exact_det4_sign(const double a[4][4]) {
  int64_t m[4][4];
  int ex[4][4];
  Big B[4][4];
  for (int r = 0; r < 4; ++r) {
    int emin = 1 << 30;
    for (int c = 0; c < 4; ++c) {
      if (a[r][c] == 0.0) {
        m[r][c] = 0;
        ex[r][c] = 0;
        continue;
      }
      int e;
      double f = std::frexp(a[r][c], &e);
      m[r][c] = (int64_t)std::ldexp(f, 53);
      ex[r][c] = e - 53;
      emin = std::min(emin, ex[r][c]);
    }
    for (int c = 0; c < 4; ++c)
      B[r][c] = m[r][c] ? Big::from_shifted(m[r][c], ex[r][c] - emin) : Big();
  }
  static const int perm[24][4] = {
      {0, 1, 2, 3}, {0, 1, 3, 2}, {0, 2, 1, 3}, {0, 2, 3, 1}, {0, 3, 1, 2},
      {0, 3, 2, 1}, {1, 0, 2, 3}, {1, 0, 3, 2}, {1, 2, 0, 3}, {1, 2, 3, 0},
      {1, 3, 0, 2}, {1, 3, 2, 0}, {2, 0, 1, 3}, {2, 0, 3, 1}, {2, 1, 0, 3},
      {2, 1, 3, 0}, {2, 3, 0, 1}, {2, 3, 1, 0}, {3, 0, 1, 2}, {3, 0, 2, 1},
      {3, 1, 0, 2}, {3, 1, 2, 0}, {3, 2, 0, 1}, {3, 2, 1, 0}};
  Big acc;
  for (int p = 0; p < 24; ++p) {
    int inv = 0;
    for (int i = 0; i < 4; ++i)
      for (int j = i + 1; j < 4; ++j) inv += perm[p][i] > perm[p][j];
    Big t = B[0][perm[p][0]];
    for (int r = 1; r < 4 && t.sign; ++r) t = Big::mul(t, B[r][perm[p][r]]);
    if (!t.sign) continue;
    if (inv & 1) t.sign = -t.sign;
    acc = Big::add(acc, t);
  }
  return acc.sign;
}

int exact_det3_sign(const double a[3][3]) {
  double b[4][4] = {{a[0][0], a[0][1], a[0][2], 0},
                    {a[1][0], a[1][1], a[1][2], 0},
                    {a[2][0], a[2][1], a[2][2], 0},
                    {0, 0, 0, 1}};
  return exact_det4_sign(b);
}

// reference-order determinants, template over the scalar
template <class T>
T det2(T a11, T a12, T a21, T a22) {
  return a11 * a22 - a12 * a21;
}
template <class T>
T det3(T a11, T a12, T a13, T a21, T a22, T a23, T a31, T a32, T a33) {
  return a11 * det2(a22, a23, a32, a33) - a21 * det2(a12, a13, a32, a33) +
         a31 * det2(a12, a13, a22, a23);
}
template <class T>
T det4(T a11, T a12, T a13, T a14, T a21, T a22, T a23, T a24, T a31, T a32,
       T a33, T a34, T a41, T a42, T a43, T a44) {
  T m12 = a21 * a12 - a11 * a22;
  T m13 = a31 * a12 - a11 * a32;
  T m14 = a41 * a12 - a11 * a42;
  T m23 = a31 * a22 - a21 * a32;
  T m24 = a41 * a22 - a21 * a42;
  T m34 = a41 * a32 - a31 * a42;
  T m123 = m23 * a13 - m13 * a23 + m12 * a33;
  T m124 = m24 * a13 - m14 * a23 + m12 * a43;
  T m134 = m34 * a13 - m14 * a33 + m13 * a43;
  T m234 = m34 * a23 - m24 * a33 + m23 * a43;
  return (m234 * a14 - m134 * a24 + m124 * a34 - m123 * a44);
}

struct P4 {
  double x, y, z, w;
};

// permanent of the reference-order det4x4 expression (every subtraction
// turned into an addition of absolute values): the sum of |monomials|
inline double perm4(const P4& p1, const P4& p2, const P4& p3, const P4& e) {
  double a11 = std::fabs(p1.x), a12 = std::fabs(p2.x), a13 = std::fabs(p3.x),
         a14 = std::fabs(e.x);
  double a21 = std::fabs(p1.y), a22 = std::fabs(p2.y), a23 = std::fabs(p3.y),
         a24 = std::fabs(e.y);
  double a31 = std::fabs(p1.z), a32 = std::fabs(p2.z), a33 = std::fabs(p3.z),
         a34 = std::fabs(e.z);
  double a41 = std::fabs(p1.w), a42 = std::fabs(p2.w), a43 = std::fabs(p3.w),
         a44 = std::fabs(e.w);
  double m12 = a21 * a12 + a11 * a22;
  double m13 = a31 * a12 + a11 * a32;
  double m14 = a41 * a12 + a11 * a42;
  double m23 = a31 * a22 + a21 * a32;
  double m24 = a41 * a22 + a21 * a42;
  double m34 = a41 * a32 + a31 * a42;
  double m123 = m23 * a13 + m13 * a23 + m12 * a33;
  double m124 = m24 * a13 + m14 * a23 + m12 * a43;
  double m134 = m34 * a13 + m14 * a33 + m13 * a43;
  double m234 = m34 * a23 + m24 * a33 + m23 * a43;
  return m234 * a14 + m134 * a24 + m124 * a34 + m123 * a44;
}

// ------------------------------------------------------------- inputs
struct Input {
  int32_t n_vert, n_tet, n_site;
  const double* verts;    // 3*n_vert
  const int32_t* tets;    // 4*n_tet
  const int32_t* v_adj;   // n_vert
  const int32_t* e_adj;   // 6*n_tet
  const int32_t* f_adj;   // 4*n_tet
  const int32_t* f_id;    // 4*n_tet
  const double* sites;    // 4*n_site (x,y,z,w=r^2)
  const uint32_t* flags;  // n_site
  const int64_t* knn_off; // n_site+1
  const int32_t* knn_ids;
};

struct Counters {
  int64_t n_det = 0, n_bisector = 0, n_exact = 0, n_vout = 0;
};

// ---------------------------------------------------------- the cell
template <int kMode>
struct Cell {
  const Input* in;
  int seed, tet;
  int status;
  int nb_v, nb_p, nb_e, nb_r;
  uint8_t first_boundary;
  uint8_t ver[kMaxT][4];
  P4 plane[kMaxP];           // arithmetic planes (float-rounded in mode 1)
  int32_t src[kMaxP];        // >=0 neighbour site; -1-k tet face k
  float h[kMaxP];
  int32_t id2[kMaxP][2];
  uint8_t edge[kMaxE][3];
  uint8_t bnext[kMaxP];
  Counters* cnt;

  static double rnd(double v) { return kMode == 1 ? (double)(float)v : v; }

  P4 site(int s) const {
    const double* p = in->sites + 4 * (int64_t)s;
    return {p[0], p[1], p[2], p[3]};
  }

  void init(const Input* input, int t, int s, Counters* c) {
    in = input;
    tet = t;
    seed = s;
    cnt = c;
    status = st_success;
    first_boundary = kEnd;
    for (int i = 0; i < kMaxP; ++i) bnext[i] = kEnd;
    static const int fl[4][3] = {{2, 1, 3}, {0, 2, 3}, {1, 0, 3}, {0, 1, 2}};
    const int32_t* tv = in->tets + 4 * (int64_t)t;
    for (int k = 0; k < 4; ++k) {
      const double* a = in->verts + 3 * (int64_t)tv[fl[k][0]];
      const double* b = in->verts + 3 * (int64_t)tv[fl[k][1]];
      const double* cc = in->verts + 3 * (int64_t)tv[fl[k][2]];
      if (kMode == 1) {
        float ax = (float)a[0], ay = (float)a[1], az = (float)a[2];
        float ux = (float)b[0] - ax, uy = (float)b[1] - ay,
              uz = (float)b[2] - az;
        float vx = (float)cc[0] - ax, vy = (float)cc[1] - ay,
              vz = (float)cc[2] - az;
        float nx = uy * vz - uz * vy, ny = uz * vx - ux * vz,
              nz = ux * vy - uy * vx;
        float d = -(ax * nx + ay * ny + az * nz);
        plane[k] = {nx, ny, nz, d};
      } else {
        double ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
        double vx = cc[0] - a[0], vy = cc[1] - a[1], vz = cc[2] - a[2];
        double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz,
               nz = ux * vy - uy * vx;
        double d = -(a[0] * nx + a[1] * ny + a[2] * nz);
        plane[k] = {nx, ny, nz, d};
      }
      src[k] = -1 - k;
      h[k] = (float)in->f_adj[4 * (int64_t)t + k];
      id2[k][0] = in->f_id[4 * (int64_t)t + k];
      id2[k][1] = -1;
    }
    nb_p = 4;
    static const uint8_t vt[4][3] = {{1, 3, 2}, {0, 2, 3}, {0, 3, 1}, {0, 1, 2}};
    for (int k = 0; k < 4; ++k) {
      for (int j = 0; j < 3; ++j) ver[k][j] = vt[k][j];
      ver[k][3] = (uint8_t)in->v_adj[tv[k]];
    }
    nb_v = 4;
    static const uint8_t et[6][2] = {{2, 3}, {1, 3}, {1, 2},
                                     {0, 3}, {0, 2}, {0, 1}};
    for (int k = 0; k < 6; ++k) {
      edge[k][0] = et[k][0];
      edge[k][1] = et[k][1];
      edge[k][2] = (uint8_t)in->e_adj[6 * (int64_t)t + k];
    }
    nb_e = 6;
  }

  bool conflict(const uint8_t* v, const P4& e) {
    cnt->n_det++;
    const P4 &p1 = plane[v[0]], &p2 = plane[v[1]], &p3 = plane[v[2]];
    if (kMode == 1) {
      auto same = [&](const P4& p) {
        return (float)e.x == (float)p.x && (float)e.y == (float)p.y &&
               (float)e.z == (float)p.z && (float)e.w == (float)p.w;
      };
      if (same(p1) || same(p2) || same(p3)) return false;
      double d = det4<double>(p1.x, p2.x, p3.x, e.x, p1.y, p2.y, p3.y, e.y,
                              p1.z, p2.z, p3.z, e.z, p1.w, p2.w, p3.w, e.w);
      return d > 0.0;
    }
    double d = det4<double>(p1.x, p2.x, p3.x, e.x, p1.y, p2.y, p3.y, e.y,
                            p1.z, p2.z, p3.z, e.z, p1.w, p2.w, p3.w, e.w);
    // every monomial of the reference-order det4x4 passes through at most
    // 9 roundings: |err| <= gamma_9 * permanent, gamma_9 < 1.0e-15
    double pm = perm4(p1, p2, p3, e);
    if (std::isfinite(pm) && pm < 1e280) {
      double eps = 1.1e-15 * pm + 1e-290;
      if (d > eps) return true;
      if (d < -eps) return false;
    }
    cnt->n_exact++;
    double a[4][4] = {{p1.x, p2.x, p3.x, e.x},
                      {p1.y, p2.y, p3.y, e.y},
                      {p1.z, p2.z, p3.z, e.z},
                      {p1.w, p2.w, p3.w, e.w}};
    return exact_det4_sign(a) > 0;
  }

  bool w_is_zero(int a, int b, int c) {
    const P4 &p1 = plane[a], &p2 = plane[b], &p3 = plane[c];
    if (kMode == 1) {
      float w = det3<float>((float)p1.x, (float)p1.y, (float)p1.z, (float)p2.x,
                            (float)p2.y, (float)p2.z, (float)p3.x, (float)p3.y,
                            (float)p3.z);
      return w == 0.f;
    }
    double m[3][3] = {{p1.x, p1.y, p1.z}, {p2.x, p2.y, p2.z},
                      {p3.x, p3.y, p3.z}};
    return exact_det3_sign(m) == 0;
  }

  int new_plane(int j) {
    if (nb_p >= kMaxP) {
      status = st_vertex_overflow;
      return -1;
    }
    cnt->n_bisector++;
    P4 S = site(seed), B = site(j);
    if (kMode == 1) {
      float sx = (float)S.x, sy = (float)S.y, sz = (float)S.z;
      float sw = (float)S.w, bx = (float)B.x, by = (float)B.y,
            bz = (float)B.z, bw = (float)B.w;
      float dx = sx - bx, dy = sy - by, dz = sz - bz;
      float ax = sx + bx, ay = sy + by, az = sz + bz;
      float dot = ax * dx + ay * dy + az * dz;
      dot = dot + (bw - sw);
      plane[nb_p] = {dx, dy, dz, -dot / 2.f};
    } else {
      double dx = S.x - B.x, dy = S.y - B.y, dz = S.z - B.z;
      double ax = S.x + B.x, ay = S.y + B.y, az = S.z + B.z;
      double dot = ax * dx + ay * dy + az * dz;
      dot = dot + (B.w - S.w);
      plane[nb_p] = {dx, dy, dz, -dot / 2.0};
    }
    src[nb_p] = j;
    h[nb_p] = 1.f;
    id2[nb_p][0] = std::min(seed, j);
    id2[nb_p][1] = std::max(seed, j);
    return nb_p++;
  }

  int find_edge(int a, int b) const {
    int lo = std::min(a, b), hi = std::max(a, b);
    for (int e = 0; e < nb_e; ++e)
      if (edge[e][0] == lo && edge[e][1] == hi) return e;
    return -1;
  }

  void new_edge(int a, int b) {
    if (nb_e >= kMaxE) {
      status = st_edge_overflow;
      return;
    }
    float m = std::max(h[a], h[b]);
    edge[nb_e][0] = (uint8_t)std::min(a, b);
    edge[nb_e][1] = (uint8_t)std::max(a, b);
    edge[nb_e][2] = (uint8_t)m;
    ++nb_e;
  }

  void new_vertex(int i, int j, int k) {
    if (nb_v + 1 >= kMaxT) {
      status = st_triangle_overflow;
      return;
    }
    int e1 = find_edge(i, j), e2 = find_edge(i, k), e3 = find_edge(j, k);
    if (e1 < 0 || e2 < 0 || e3 < 0) {
      status = st_inconsistent_boundary;
      return;
    }
    uint8_t adj =
        std::max(std::max(edge[e1][2], edge[e2][2]), edge[e3][2]);
    ver[nb_v][0] = (uint8_t)i;
    ver[nb_v][1] = (uint8_t)j;
    ver[nb_v][2] = (uint8_t)k;
    ver[nb_v][3] = adj;
    ++nb_v;
  }

  void swap_ver(int a, int b) {
    for (int q = 0; q < 4; ++q) std::swap(ver[a][q], ver[b][q]);
  }

  void compute_boundary() {
    for (int i = 0; i < kMaxP; ++i) bnext[i] = kEnd;
    first_boundary = kEnd;
    int iter = 0;
    int t = nb_v;
    while (nb_r > 0) {
      if (iter++ > 65535) {
        status = st_inconsistent_boundary;
        return;
      }
      bool in_border[3], next_opp[3];
      for (int e = 0; e < 3; ++e) in_border[e] = bnext[ver[t][e]] != kEnd;
      for (int e = 0; e < 3; ++e)
        next_opp[e] = bnext[ver[t][(e + 1) % 3]] == ver[t][e];
      bool simple = true;
      for (int e = 0; e < 3; ++e)
        if (!next_opp[e] && !next_opp[(e + 1) % 3] && in_border[(e + 1) % 3])
          simple = false;
      if (!next_opp[0] && !next_opp[1] && !next_opp[2]) {
        if (first_boundary == kEnd) {
          for (int e = 0; e < 3; ++e)
            bnext[ver[t][e]] = ver[t][(e + 1) % 3];
          first_boundary = ver[t][0];
        } else {
          simple = false;
        }
      }
      if (!simple) {
        ++t;
        if (t == nb_v + nb_r) t = nb_v;
        continue;
      }
      for (int e = 0; e < 3; ++e)
        if (!next_opp[e]) bnext[ver[t][e]] = ver[t][(e + 1) % 3];
      for (int e = 0; e < 3; ++e)
        if (next_opp[e] && next_opp[(e + 1) % 3]) {
          uint8_t p = ver[t][(e + 1) % 3];
          if (first_boundary == p) first_boundary = bnext[p];
          bnext[p] = kEnd;
        }
      swap_ver(t, nb_v + nb_r - 1);
      t = nb_v;
      --nb_r;
    }
  }

  void clip(int j) {
    int cur = new_plane(j);
    if (status != st_success) return;
    P4 e = plane[cur];
    nb_r = 0;
    int i = 0;
    while (i < nb_v) {
      if (conflict(ver[i], e)) {
        --nb_v;
        swap_ver(i, nb_v);
        ++nb_r;
      } else {
        ++i;
      }
    }
    if (nb_r == 0) {
      --nb_p;
      return;
    }
    if (nb_v == 0) {
      status = st_no_intersection;
      return;
    }
    compute_boundary();
    if (status != st_success) return;
    if (first_boundary == kEnd) return;
    uint8_t cir = first_boundary;
    do {
      new_edge(cur, cir);
      if (status != st_success) return;
      cir = bnext[cir];
    } while (cir != first_boundary);
    cir = first_boundary;
    do {
      int a = cur, b = cir, c = bnext[cir];
      new_vertex(a, b, c);
      if (w_is_zero(a, b, c) && status == st_success)
        status = st_needs_perturb;
      if (status != st_success) return;
      cir = bnext[cir];
    } while (cir != first_boundary);
  }
};

// ------------------------------------------------------------ geometry
// binary64 planes of a cell (mode X construction), independent of mode
struct Geo {
  double vol, cx, cy, cz;
  double area[kMaxP];
};

template <int kMode>
void cell_geometry(const Cell<kMode>& c, Geo& g) {
  // rebuild binary64 planes
  P4 pl[kMaxP];
  static const int fl[4][3] = {{2, 1, 3}, {0, 2, 3}, {1, 0, 3}, {0, 1, 2}};
  const int32_t* tv = c.in->tets + 4 * (int64_t)c.tet;
  for (int p = 0; p < c.nb_p; ++p) {
    if (c.src[p] < 0) {
      int k = -1 - c.src[p];
      const double* a = c.in->verts + 3 * (int64_t)tv[fl[k][0]];
      const double* b = c.in->verts + 3 * (int64_t)tv[fl[k][1]];
      const double* cc = c.in->verts + 3 * (int64_t)tv[fl[k][2]];
      double ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
      double vx = cc[0] - a[0], vy = cc[1] - a[1], vz = cc[2] - a[2];
      double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz,
             nz = ux * vy - uy * vx;
      pl[p] = {nx, ny, nz, -(a[0] * nx + a[1] * ny + a[2] * nz)};
    } else {
      P4 S = c.site(c.seed), B = c.site(c.src[p]);
      double dx = S.x - B.x, dy = S.y - B.y, dz = S.z - B.z;
      double ax = S.x + B.x, ay = S.y + B.y, az = S.z + B.z;
      double dot = ax * dx + ay * dy + az * dz + (B.w - S.w);
      pl[p] = {dx, dy, dz, -dot / 2.0};
    }
  }
  // vertex positions: intersection of 3 planes by Cramer's rule
  double px[kMaxT], py[kMaxT], pz[kMaxT];
  for (int t = 0; t < c.nb_v; ++t) {
    const P4 &a = pl[c.ver[t][0]], &b = pl[c.ver[t][1]], &d = pl[c.ver[t][2]];
    double D = det3<double>(a.x, a.y, a.z, b.x, b.y, b.z, d.x, d.y, d.z);
    double Dx = det3<double>(-a.w, a.y, a.z, -b.w, b.y, b.z, -d.w, d.y, d.z);
    double Dy = det3<double>(a.x, -a.w, a.z, b.x, -b.w, b.z, d.x, -d.w, d.z);
    double Dz = det3<double>(a.x, a.y, -a.w, b.x, b.y, -b.w, d.x, d.y, -d.w);
    px[t] = Dx / D;
    py[t] = Dy / D;
    pz[t] = Dz / D;
  }
  g.vol = 0;
  double mx = 0, my = 0, mz = 0;
  for (int p = 0; p < kMaxP; ++p) g.area[p] = 0;
  const double ox = px[0], oy = py[0], oz = pz[0];
  for (int p = 0; p < c.nb_p; ++p) {
    // vertices on plane p with the local slot of p
    int tv_[kMaxT], lp[kMaxT], n = 0;
    for (int t = 0; t < c.nb_v; ++t)
      for (int k = 0; k < 3; ++k)
        if (c.ver[t][k] == p) {
          tv_[n] = t;
          lp[n] = k;
          ++n;
          break;
        }
    if (n < 3) continue;
    // order the cycle: next of vertex i is the vertex j whose plane at slot
    // (lp_j + 2) % 3 equals the plane of i at slot (lp_i + 1) % 3
    int cyc[kMaxT], m = 0, cur = 0;
    std::vector<char> used(n, 0);
    while (m < n) {
      cyc[m++] = tv_[cur];
      used[cur] = 1;
      int nxt_plane = c.ver[tv_[cur]][(lp[cur] + 1) % 3];
      int nxt = -1;
      for (int j = 0; j < n; ++j)
        if (!used[j] && c.ver[tv_[j]][(lp[j] + 2) % 3] == nxt_plane) {
          nxt = j;
          break;
        }
      if (nxt < 0) break;
      cur = nxt;
    }
    // Newell area vector and fan volume from vertex 0
    double ax = 0, ay = 0, az = 0;
    for (int q = 0; q < m; ++q) {
      int a = cyc[q], b = cyc[(q + 1) % m];
      ax += (py[a] - py[b]) * (pz[a] + pz[b]);
      ay += (pz[a] - pz[b]) * (px[a] + px[b]);
      az += (px[a] - px[b]) * (py[a] + py[b]);
    }
    g.area[p] = 0.5 * std::sqrt(ax * ax + ay * ay + az * az);
    int a0 = cyc[0];
    for (int q = 1; q + 1 < m; ++q) {
      int a = cyc[q], b = cyc[q + 1];
      double ux = px[a0] - ox, uy = py[a0] - oy, uz = pz[a0] - oz;
      double vx = px[a] - ox, vy = py[a] - oy, vz = pz[a] - oz;
      double wx = px[b] - ox, wy = py[b] - oy, wz = pz[b] - oz;
      double v6 = ux * (vy * wz - vz * wy) - uy * (vx * wz - vz * wx) +
                  uz * (vx * wy - vy * wx);
      g.vol += v6;
      mx += v6 * (ox + px[a0] + px[a] + px[b]);
      my += v6 * (oy + py[a0] + py[a] + py[b]);
      mz += v6 * (oz + pz[a0] + pz[a] + pz[b]);
    }
  }
  double sgn = g.vol < 0 ? -1.0 : 1.0;
  if (g.vol != 0) {
    g.cx = mx / (4.0 * g.vol);
    g.cy = my / (4.0 * g.vol);
    g.cz = mz / (4.0 * g.vol);
  } else {
    g.cx = g.cy = g.cz = 0;
  }
  g.vol = sgn * g.vol / 6.0;
}

// ------------------------------------------------------------ results
struct CellOut {
  int32_t tet, seed;
  uint8_t status, nb_v, nb_p, nb_e;
  uint32_t ver[kMaxT];      // p0 | p1<<8 | p2<<16 | adj<<24
  int32_t src[kMaxP];       // >=0 site, -1-k tet face k
  uint32_t edge[kMaxE];     // a | b<<8 | adj<<16
  double geo[4];            // volume, centroid
  double area[kMaxP];
};

struct Result {
  std::vector<int64_t> rel_off;   // per requested tet
  std::vector<int32_t> rel_ids;
  std::vector<CellOut> cells;     // successful cells, (tet order, site asc)
  std::vector<int32_t> dropped_status_count;  // per status code 0..8
  Counters cnt;
};

bool related(const Input& in, int t, int i) {
  if (in.flags[i] != 1u) return false;
  const int32_t* tv = in.tets + 4 * (int64_t)t;
  const double* si = in.sites + 4 * (int64_t)i;
  double pdi[4];
  for (int k = 0; k < 4; ++k) {
    const double* v = in.verts + 3 * (int64_t)tv[k];
    double dx = si[0] - v[0], dy = si[1] - v[1], dz = si[2] - v[2];
    pdi[k] = dx * dx + dy * dy + dz * dz - si[3];
  }
  for (int64_t q = in.knn_off[i]; q < in.knn_off[i + 1]; ++q) {
    int m = in.knn_ids[q];
    if (m < 0) continue;
    const double* sm = in.sites + 4 * (int64_t)m;
    bool ok = false;
    for (int k = 0; k < 4 && !ok; ++k) {
      const double* v = in.verts + 3 * (int64_t)tv[k];
      double dx = sm[0] - v[0], dy = sm[1] - v[1], dz = sm[2] - v[2];
      double pdm = dx * dx + dy * dy + dz * dz - sm[3];
      ok = pdm > pdi[k];
    }
    if (!ok) return false;
  }
  return true;
}

template <int kMode>
void run_cells(const Input& in, const std::vector<int32_t>& tet_list,
               bool brute_force, int n_threads, Result& res,
               const int64_t* given_off = nullptr,
               const int32_t* given_ids = nullptr) {
  const int64_t nt = (int64_t)tet_list.size();
  std::vector<std::vector<int32_t>> rel(nt);
  std::vector<std::vector<CellOut>> out(nt);
  std::vector<Counters> cnts(std::max(1, n_threads));
  std::vector<std::vector<int32_t>> drops(std::max(1, n_threads),
                                          std::vector<int32_t>(9, 0));
  std::atomic<int64_t> next{0};
  auto worker = [&](int tid) {
    Cell<kMode>* c = new Cell<kMode>();
    std::vector<int32_t> all_sites;
    if (brute_force) {
      all_sites.resize(in.n_site);
      for (int s = 0; s < in.n_site; ++s) all_sites[s] = s;
    }
    for (;;) {
      int64_t q = next.fetch_add(1);
      if (q >= nt) break;
      int t = tet_list[q];
      std::vector<int32_t> cands;
      if (given_off) {
        cands.assign(given_ids + given_off[q], given_ids + given_off[q + 1]);
      } else {
        for (int i = 0; i < in.n_site; ++i) {
          bool r = brute_force ? in.flags[i] == 1u : related(in, t, i);
          if (r) cands.push_back(i);
        }
      }
      for (int i : cands) {
        rel[q].push_back(i);
        c->init(&in, t, i, &cnts[tid]);
        if (brute_force) {
          for (int s : all_sites) {
            if (s == i) continue;
            c->clip(s);
            if (c->status != st_success) break;
          }
        } else {
          for (int64_t k = in.knn_off[i]; k < in.knn_off[i + 1]; ++k) {
            int j = in.knn_ids[k];
            if (j == -1) break;
            c->clip(j);
            if (c->status != st_success) break;
          }
        }
        if (c->status != st_success) {
          drops[tid][c->status]++;
          continue;
        }
        CellOut o;
        std::memset(&o, 0, sizeof(o));
        o.tet = t;
        o.seed = i;
        o.status = (uint8_t)c->status;
        o.nb_v = (uint8_t)c->nb_v;
        o.nb_p = (uint8_t)c->nb_p;
        o.nb_e = (uint8_t)c->nb_e;
        for (int v = 0; v < c->nb_v; ++v)
          o.ver[v] = c->ver[v][0] | (c->ver[v][1] << 8) |
                     (c->ver[v][2] << 16) | ((uint32_t)c->ver[v][3] << 24);
        for (int p = 0; p < c->nb_p; ++p) o.src[p] = c->src[p];
        for (int e = 0; e < c->nb_e; ++e)
          o.edge[e] = c->edge[e][0] | (c->edge[e][1] << 8) |
                      (c->edge[e][2] << 16);
        Geo g;
        cell_geometry(*c, g);
        cnts[tid].n_vout += c->nb_v;
        o.geo[0] = g.vol;
        o.geo[1] = g.cx;
        o.geo[2] = g.cy;
        o.geo[3] = g.cz;
        for (int p = 0; p < kMaxP; ++p) o.area[p] = g.area[p];
        out[q].push_back(o);
      }
    }
    delete c;
  };
  int nthr = std::max(1, n_threads);
  std::vector<std::thread> pool;
  for (int t = 0; t < nthr; ++t) pool.emplace_back(worker, t);
  for (auto& th : pool) th.join();
  res.rel_off.assign(nt + 1, 0);
  for (int64_t q = 0; q < nt; ++q) {
    res.rel_off[q + 1] = res.rel_off[q] + (int64_t)rel[q].size();
    res.rel_ids.insert(res.rel_ids.end(), rel[q].begin(), rel[q].end());
    res.cells.insert(res.cells.end(), out[q].begin(), out[q].end());
  }
  res.dropped_status_count.assign(9, 0);
  for (auto& d : drops)
    for (int s = 0; s < 9; ++s) res.dropped_status_count[s] += d[s];
  for (auto& c : cnts) {
    res.cnt.n_det += c.n_det;
    res.cnt.n_bisector += c.n_bisector;
    res.cnt.n_exact += c.n_exact;
    res.cnt.n_vout += c.n_vout;
  }
}

}  // namespace

// ---------------------------------------------------------------- C ABI
extern "C" {

typedef struct {
  int32_t n_vert, n_tet, n_site;
  const double* verts;
  const int32_t* tets;
  const int32_t* v_adj;
  const int32_t* e_adj;
  const int32_t* f_adj;
  const int32_t* f_id;
  const double* sites;
  const uint32_t* flags;
  const int64_t* knn_off;
  const int32_t* knn_ids;
} orc_input;

typedef struct orc_result orc_result;
struct orc_result {
  Result r;
};

// Runs relation + clip for the tets in tet_list (or all tets if tet_list is
// NULL).  mode 0 = exact binary64, 1 = reference float arithmetic.
// brute_force = 1 clips every flagged sphere by every other sphere.
orc_result* orc_run(const orc_input* in, const int32_t* tet_list,
                    int32_t n_list, int32_t mode, int32_t brute_force,
                    int32_t n_threads) {
  Input I{in->n_vert, in->n_tet,  in->n_site, in->verts,
          in->tets,   in->v_adj,  in->e_adj,  in->f_adj,
          in->f_id,   in->sites,  in->flags,  in->knn_off,
          in->knn_ids};
  std::vector<int32_t> tl;
  if (tet_list) {
    tl.assign(tet_list, tet_list + n_list);
  } else {
    tl.resize(in->n_tet);
    for (int t = 0; t < in->n_tet; ++t) tl[t] = t;
  }
  auto* res = new orc_result();
  if (mode == 1)
    run_cells<1>(I, tl, brute_force != 0, n_threads, res->r);
  else
    run_cells<0>(I, tl, brute_force != 0, n_threads, res->r);
  return res;
}

// Clip only: candidate lists per requested tet are given (CSR over
// tet_list), the relation is not evaluated.
orc_result* orc_clip(const orc_input* in, const int32_t* tet_list,
                     int32_t n_list, const int64_t* cand_off,
                     const int32_t* cand_ids, int32_t mode,
                     int32_t n_threads) {
  Input I{in->n_vert, in->n_tet,  in->n_site, in->verts,
          in->tets,   in->v_adj,  in->e_adj,  in->f_adj,
          in->f_id,   in->sites,  in->flags,  in->knn_off,
          in->knn_ids};
  std::vector<int32_t> tl(tet_list, tet_list + n_list);
  auto* res = new orc_result();
  if (mode == 1)
    run_cells<1>(I, tl, false, n_threads, res->r, cand_off, cand_ids);
  else
    run_cells<0>(I, tl, false, n_threads, res->r, cand_off, cand_ids);
  return res;
}

void orc_free(orc_result* r) { delete r; }
int64_t orc_n_cells(const orc_result* r) { return (int64_t)r->r.cells.size(); }
int64_t orc_n_rel(const orc_result* r) { return (int64_t)r->r.rel_ids.size(); }
int64_t orc_cell_bytes(void) { return (int64_t)sizeof(CellOut); }
void orc_get_relation(const orc_result* r, int64_t* off, int32_t* ids) {
  std::copy(r->r.rel_off.begin(), r->r.rel_off.end(), off);
  std::copy(r->r.rel_ids.begin(), r->r.rel_ids.end(), ids);
}
void orc_get_cells(const orc_result* r, void* dst) {
  std::memcpy(dst, r->r.cells.data(), r->r.cells.size() * sizeof(CellOut));
}
void orc_get_drops(const orc_result* r, int32_t* counts9) {
  for (int s = 0; s < 9; ++s) counts9[s] = r->r.dropped_status_count[s];
}
void orc_get_counters(const orc_result* r, int64_t* c4) {
  c4[0] = r->r.cnt.n_det;
  c4[1] = r->r.cnt.n_bisector;
  c4[2] = r->r.cnt.n_exact;
  c4[3] = r->r.cnt.n_vout;
}

// exact sign of a 4x4 determinant given row-major doubles (test hook)
int orc_exact_det4_sign(const double* a16) {
  double a[4][4];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) a[r][c] = a16[4 * r + c];
  return exact_det4_sign(a);
}

}  // extern "C"
