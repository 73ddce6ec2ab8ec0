// Device-side data layout and kernel launchers of the RPD engine (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../common/rpd_predicates.h"

namespace rpd {

// Per-tet metadata packed in 16 bytes:
//   x: f_adj[0..3] (u8 each)      y: v_adj of the 4 vertices (u8 each)
//   z: e_adj[0..3] (u8 each)      w: e_adj[4..5] (u8 each), 2 bytes spare
// The reference stores these counts in uchar fields of the cell
// (ver.w, edge.z) so they are truncated to 8 bits the same way.
struct TetMeta {
  uint32_t x, y, z, w;
};

// One candidate cell (tet, seed) with the bitmask of neighbour planes that
// may cut it (bit k = k-th neighbour of seed in its list; lists longer than
// 128 treat positions >= 128 as active).
struct Cand {
  int32_t tet;
  int32_t seed;
  uint64_t m0, m1;
};

struct DevMesh {
  int32_t n_vert, n_tet;
  const double4* vert;   // x, y, z, 0
  const int4* tet_v;
  const int4* tet_fid;
  const TetMeta* tet_meta;
};

struct DevSites {
  int32_t n_site;
  const double4* site;   // x, y, z, w = r^2
  const uint32_t* flags;
  const int64_t* knn_off;
  const int32_t* knn_ids;   // neighbours in ascending site id (clip order)
  const int32_t* knn_near;  // the same lists (nearest bisector first for
                            // the relation, list order for the learned
                            // path), bit 31 set for unflagged sites
  const double4* knn_site_near;  // the spheres of knn_near, inline
  const double4* knn_bis;  // power bisector of each knn_ids entry (zero
                           // for a -1 entry), in knn_ids order
  const float4* knn_bisf;  // the same rounded to binary32 for the plane
                           // masks (NaN for components of magnitude
                           // outside [1e-30, 1e30])
};

// capacities of a clip tier
template <int V_, int P_, int E_>
struct Caps {
  static constexpr int V = V_, P = P_, E = E_;
};
using SmallCaps = Caps<32, 16, 48>;
using WideCaps = Caps<48, 24, 72>;   // small tier for larger cells
using TightCaps = Caps<16, 12, 32>;  // small tier for small cells
using MidCaps = Caps<24, 14, 40>;
using LargeCaps = Caps<96, 64, 152>;

// Output slot of one cell (indexed by cell or by re-run index).  Its words
// hold the edge records at 0 (appended as the clip creates edges; none is
// ever removed), then the final vertices and plane sources right after the
// last edge, so a cell's payload is one span [0, nb_e + nb_v + nb_p) at
// the start of its slot; the stride rounds V + P + E up to 16 words (64
// bytes), so the span starts on a 64-byte boundary.
inline constexpr int32_t slot_stride_words(int V, int P, int E) {
  return (V + P + E + 15) & ~15;
}
struct SlotView {
  int32_t stride_words;   // slot_stride_words(V, P, E)
  int32_t stride_areas;   // P
  uint32_t* hdr;          // nb_v | nb_p<<8 | nb_e<<16 | status<<24
  uint32_t* words;
  double* geom;           // 4 per slot
  double* areas;
};

enum : uint32_t {
  kStatusRerunLarge = 32,   // internal: small tier overflowed
  kStatusRerunExact = 33,   // internal: filter undecided, exact re-run
};

}  // namespace rpd
