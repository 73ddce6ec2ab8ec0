// Test driver (host only, no GPU): reads a cell file (the format of
// oracle/adj_oracle.cpp), writes the cells' faces as a binary geometry with
// the product's save_convex_cells_bgeo (all faces, and boundary faces only)
// and the triangle fans of get_one_convex_cell_faces_const as text (one
// line per triangle: site, site across, three point indices).
// usage: geo_dump CELLS MAX_SURF_FID OUT_ALL.bgeo OUT_BOUNDARY.bgeo OUT_TRI
#include <cstdio>
#include <cstdlib>

#include "cell_reader.h"
#include "libmat_rpd.h"

int main(int argc, char** argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: geo_dump CELLS MAX_SURF_FID ALL BND TRI\n");
    return 2;
  }
  try {
    std::vector<ConvexCellHost> cells = cellfile::read_cells(argv[1]);
    const int max_sf = std::atoi(argv[2]);
    libmat_rpd::save_convex_cells_bgeo(argv[3], cells, max_sf, false);
    libmat_rpd::save_convex_cells_bgeo(argv[4], cells, max_sf, true);
    std::vector<cfloat3> pts;
    std::vector<std::vector<unsigned>> tri;
    std::vector<int> site, adj;
    for (const auto& c : cells)
      get_one_convex_cell_faces_const(c, pts, tri, site, adj, true, max_sf,
                                      false);
    FILE* f = std::fopen(argv[5], "w");
    if (!f) throw std::runtime_error("cannot open output");
    for (size_t q = 0; q < tri.size(); ++q)
      std::fprintf(f, "%d %d %u %u %u\n", site[q], adj[q], tri[q][0],
                   tri[q][1], tri[q][2]);
    std::fclose(f);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "geo_dump: %s\n", e.what());
    return 1;
  }
  return 0;
}
