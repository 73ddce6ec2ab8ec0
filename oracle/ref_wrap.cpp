// ORACLE cross-check (test infrastructure only): C-ABI wrappers around the
// reference's own CPU sources, compiled in place from the mounted
// reference by oracle/ref.mk into oracle/_ref/libref_psm.so:
//   src/CPU/Delaunay_psm.cpp  exact expansion determinant
//                             (sign_of_expansion_determinant, :9539) and
//                             the regular weighted Delaunay "BPOW"
//                             (:20398-20548, :22084-22110)
//   src/CPU/convex_cell.cpp   FP64 VBW::ConvexCell clipping with volume,
//                             barycenter and facet areas (:169-600)
#include <cmath>
#include <cstdint>
#include <vector>

// The expansion class is only declared inside the PSM source, so that
// source is compiled as part of this translation unit (from where it lies).
#include "Delaunay_psm.cpp"
#include "convex_cell.h"

extern "C" {

int ref_init() {
  GEO::initialize();
  return 0;
}

// exact sign of det of the row-major 4x4 matrix
int ref_det4_sign(const double* a) {
  GEO::expansion* e[16];
  for (int k = 0; k < 16; ++k) {
    e[k] = GEO::expansion::new_expansion_on_heap(1);
    e[k]->assign(a[k]);
  }
  GEO::Sign s = GEO::sign_of_expansion_determinant(
      *e[0], *e[1], *e[2], *e[3], *e[4], *e[5], *e[6], *e[7], *e[8], *e[9],
      *e[10], *e[11], *e[12], *e[13], *e[14], *e[15]);
  for (int k = 0; k < 16; ++k) GEO::expansion::delete_expansion_on_heap(e[k]);
  return (int)s;
}

// regular triangulation of weighted points (x, y, z, w): 1-ring
// neighbours per point (empty for hidden points); CSR into off/ids
// (ids capacity cap); returns the number of ids or -1 if cap is short.
int64_t ref_rt_neighbors(int32_t n, const double* xyzw, int64_t* off,
                         int32_t* ids, int64_t cap) {
  double W = 0;
  for (int i = 0; i < n; ++i) W = std::max(W, xyzw[4 * i + 3]);
  std::vector<double> v4(4 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    v4[4 * i + 0] = xyzw[4 * i + 0];
    v4[4 * i + 1] = xyzw[4 * i + 1];
    v4[4 * i + 2] = xyzw[4 * i + 2];
    v4[4 * i + 3] = std::sqrt(W - xyzw[4 * i + 3]);
  }
  GEO::Delaunay_var del = GEO::Delaunay::create(4, "BPOW");
  del->set_keeps_infinite(true);
  del->set_stores_neighbors(true);
  del->set_vertices((GEO::index_t)n, v4.data());
  int64_t tot = 0;
  off[0] = 0;
  GEO::vector<GEO::index_t> nb;
  for (int i = 0; i < n; ++i) {
    nb.clear();
    del->get_neighbors((GEO::index_t)i, nb);
    std::vector<int32_t> L;
    for (GEO::index_t j : nb)
      if (j < (GEO::index_t)n && (int)j != i) L.push_back((int32_t)j);
    std::sort(L.begin(), L.end());
    L.erase(std::unique(L.begin(), L.end()), L.end());
    for (int32_t j : L) {
      if (tot >= cap) return -1;
      ids[tot++] = j;
    }
    off[i + 1] = tot;
  }
  return tot;
}

// VBW convex cell: bounding box, then the planes in order (keep side
// a x + b y + c z + d >= 0 in the library's convention, checked by the
// tests).  out: volume, barycenter (3), then one facet area per plane.
int ref_vbw_cell(const double* box6, int32_t n_planes, const double* planes,
                 double* out) {
  VBW::ConvexCell C;
  C.init_with_box(box6[0], box6[1], box6[2], box6[3], box6[4], box6[5]);
  const VBW::index_t first = C.nb_v();
  for (int p = 0; p < n_planes; ++p)
    C.clip_by_plane(VBW::make_vec4(planes[4 * p], planes[4 * p + 1],
                                   planes[4 * p + 2], planes[4 * p + 3]));
  C.compute_geometry();
  out[0] = C.volume();
  VBW::vec3 b = C.barycenter();
  out[1] = b.x;
  out[2] = b.y;
  out[3] = b.z;
  for (int p = 0; p < n_planes; ++p)
    out[4 + p] = (first + p < C.nb_v()) ? C.facet_area(first + p) : 0.0;
  return (int)first;
}

}  // extern "C"
