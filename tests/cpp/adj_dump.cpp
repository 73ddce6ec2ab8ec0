// Test driver (host only, no GPU): reads a cell file (the format of
// oracle/adj_oracle.cpp), builds the ConvexCellHost records, runs the
// product's update_power_cells_adjacency, cal_cell_euler and
// cal_halfplane_facet_euler, and writes the canonical text dump that the
// oracle's orc_adjacency writes for the same input.
// usage: adj_dump CELLS ES2FE|- MAX_SURF_FID OUT
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "cell_reader.h"
#include "libmat_rpd.h"

namespace {

void put_set(FILE* f, const std::set<int>& s) {
  std::fputc('{', f);
  bool first = true;
  for (int x : s) {
    std::fprintf(f, first ? "%d" : " %d", x);
    first = false;
  }
  std::fputc('}', f);
}
void put_pf(FILE* f, const rpd_v2int& p) {
  std::fprintf(f, "(%a %a %a %d)", p.first[0], p.first[1], p.first[2],
               p.second);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: adj_dump CELLS ES2FE|- MAX_SURF_FID OUT\n");
    return 2;
  }
  try {
    std::vector<ConvexCellHost> cells = cellfile::read_cells(argv[1]);
    std::map<aint3, aint3> es2fe;
    if (std::strcmp(argv[2], "-") != 0) {
      FILE* f = std::fopen(argv[2], "rb");
      if (!f) throw std::runtime_error("cannot open feature-edge file");
      int32_t n = 0;
      cellfile::rd(f, &n, 1);
      for (int q = 0; q < n; ++q) {
        int32_t r[6];
        cellfile::rd(f, r, 6);
        es2fe[aint3{{r[0], r[1], r[2]}}] = aint3{{r[3], r[4], r[5]}};
      }
      std::fclose(f);
    }
    const int max_surf_fid = std::atoi(argv[3]);
    auto adj = update_power_cells_adjacency(cells, max_surf_fid, &es2fe);
    FILE* f = std::fopen(argv[4], "w");
    if (!f) throw std::runtime_error("cannot open output");
    for (auto& c : cells) {
      std::fprintf(f, "ce %d %a\n", c.id, c.cal_cell_euler());
      for (int p = 0; p < c.nb_p; ++p) {
        if (c.active_clipping_planes[p] <= 0) continue;
        const cint2 h = c.clip_id2_const(p);
        if (h.y == -1) continue;
        const int nb = h.x == c.voro_id ? h.y : h.x;
        std::fprintf(f, "fe %d %d %a\n", c.id, nb,
                     c.cal_halfplane_facet_euler(nb));
      }
    }
    for (const auto& kv : adj) {
      const PowerCellAdjacency& p = kv.second;
      std::fprintf(f, "S %d\ncells ", kv.first);
      put_set(f, p.cell_ids);
      std::fprintf(f, "\ntets ");
      put_set(f, p.tet_ids);
      std::fputc('\n', f);
      for (const auto& x : p.cell_neighbors) {
        std::fprintf(f, "nbr %d ", x.first);
        put_set(f, x.second);
        std::fputc('\n', f);
      }
      for (const auto& x : p.cell_to_tfids) {
        std::fprintf(f, "tfid %d ", x.first);
        put_set(f, x.second);
        std::fputc('\n', f);
      }
      for (const auto& x : p.cell_to_surfv2fid) {
        std::fprintf(f, "surf %d", x.first);
        for (const auto& y : x.second) put_pf(f, y);
        std::fputc('\n', f);
      }
      for (const auto& x : p.facet_neigh_to_cells) {
        std::fprintf(f, "facet %d ", x.first);
        put_set(f, x.second);
        std::fputc('\n', f);
      }
      for (const auto& x : p.e_to_cells) {
        std::fprintf(f, "edge %d %d ", x.first[0], x.first[1]);
        put_set(f, x.second);
        std::fputc('\n', f);
      }
      for (const auto& x : p.edge_2endvertices) {
        std::fprintf(f, "e2v %d %d", x.first[0], x.first[1]);
        for (const auto& y : x.second)
          std::fprintf(f, " [%d %d %d %d]", y[0][0], y[0][1], y[1][0],
                       y[1][1]);
        std::fputc('\n', f);
      }
      for (const auto& x : p.vertex_2id)
        std::fprintf(f, "v2id %d %d %d %d %d\n", x.first[0], x.first[1],
                     x.second[0], x.second[1], x.second[2]);
      for (const auto& x : p.vertex_2pos) {
        std::fprintf(f, "v2pos %d %d ", x.first[0], x.first[1]);
        put_pf(f, x.second);
        std::fputc('\n', f);
      }
      for (const auto& x : p.se_covered_lvids)
        std::fprintf(f, "se %d %d %d %d %d\n", x[0], x[1], x[2], x[3], x[4]);
      for (const auto& x : p.ce_covered_lvids)
        std::fprintf(f, "ce5 %d %d %d %d %d\n", x[0], x[1], x[2], x[3], x[4]);
      for (const auto& x : p.se_line_endpos)
        std::fprintf(f, "endpos %d %d %d %d %a %a %a\n", x.first[0],
                     x.first[1], x.first[2], x.first[3], x.second[0],
                     x.second[1], x.second[2]);
      std::fprintf(f, "cc");
      for (const auto& g : p.cc_cells) {
        std::fputc(' ', f);
        put_set(f, g);
      }
      std::fprintf(f, "\nccsurf");
      for (const auto& g : p.cc_surf_v2fids) {
        std::fprintf(f, " <");
        for (const auto& y : g) put_pf(f, y);
        std::fputc('>', f);
      }
      std::fputc('\n', f);
      for (const auto& x : p.facet_cc_cells) {
        std::fprintf(f, "fcc %d", x.first);
        for (const auto& g : x.second) {
          std::fputc(' ', f);
          put_set(f, g);
        }
        std::fputc('\n', f);
      }
      for (const auto& x : p.facet_cc_surf_v2fids) {
        std::fprintf(f, "fccsurf %d", x.first);
        for (const auto& g : x.second) {
          std::fprintf(f, " <");
          for (const auto& y : g) put_pf(f, y);
          std::fputc('>', f);
        }
        std::fputc('\n', f);
      }
      for (const auto& x : p.edge_cc_cells) {
        std::fprintf(f, "ecc %d %d", x.first[0], x.first[1]);
        for (const auto& g : x.second) {
          std::fputc(' ', f);
          put_set(f, g);
        }
        std::fputc('\n', f);
      }
    }
    std::fclose(f);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "adj_dump: %s\n", e.what());
    return 1;
  }
  return 0;
}
