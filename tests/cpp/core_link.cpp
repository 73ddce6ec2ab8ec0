// A program that compiles its own ConvexCellHost methods (as LibMAT does
// with voronoi_defs.cxx) and links librpd_amd_core.so for the RPD entry
// point: the library defines none of the methods, so each has exactly one
// definition in the program.
#include <cstdio>

#include "libmat_rpd.h"

static int g_calls = 0;
cfloat4 ConvexCellHost::compute_vertex_coordinates(cuchar3, bool) const {
  ++g_calls;
  return cfloat4{0, 0, 0, 1};
}
void ConvexCellHost::print_info() const { ++g_calls; }
void ConvexCellHost::reload_active() { ++g_calls; }
void ConvexCellHost::reload_pc_explicit() { ++g_calls; }
Scalar ConvexCellHost::cal_cell_euler() { return ++g_calls; }
Scalar ConvexCellHost::cal_halfplane_facet_euler(int) { return ++g_calls; }
void ConvexCellHost::print_cell_detail_euler() const { ++g_calls; }
bool is_convex_cell_valid(const ConvexCellHost&) { return ++g_calls > 0; }

int main() {
  ConvexCellHost c;
  c.reload_active();
  // the entry point is taken from the library (not called: no GPU needed)
  auto* f = &compute_clipped_voro_diagram_GPU;
  std::printf("core link ok %d %d\n", g_calls, f != nullptr);
  return 0;
}
