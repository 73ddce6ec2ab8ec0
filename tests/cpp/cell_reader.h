// Test helper (host only): the cell file of oracle/adj_oracle.cpp read into
// ConvexCellHost records.
#pragma once
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "libmat_rpd.h"

namespace cellfile {

template <class T>
void rd(FILE* f, T* p, size_t n) {
  if (n && std::fread(p, sizeof(T), n, f) != n)
    throw std::runtime_error("cell file truncated");
}

std::vector<ConvexCellHost> read_cells(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) throw std::runtime_error("cannot open cell file");
  char magic[4];
  int32_t version = 0, n = 0;
  rd(f, magic, 4);
  rd(f, &version, 1);
  rd(f, &n, 1);
  if (std::memcmp(magic, "RPDC", 4) != 0 || version != 1)
    throw std::runtime_error("bad cell file");
  std::vector<ConvexCellHost> cells(n);
  for (auto& c : cells) {
    int32_t h[8];
    rd(f, h, 8);
    c.id = h[0];
    c.voro_id = h[1];
    c.tet_id = h[2];
    c.thread_id = h[3];
    c.status = (Status)h[4];
    c.nb_v = (uchar)h[5];
    c.nb_p = (uchar)h[6];
    c.nb_e = (uchar)h[7];
    c.is_active = true;
    rd(f, &c.weight, 1);
    rd(f, &c.cell_vol, 1);
    for (int v = 0; v < c.nb_v; ++v) rd(f, &c.ver_data_trans[v].x, 4);
    for (int p = 0; p < c.nb_p; ++p) rd(f, &c.clip_data_trans[p].x, 5);
    for (int p = 0; p < c.nb_p; ++p) rd(f, &c.clip_id2_data_trans[p].x, 2);
    for (int e = 0; e < c.nb_e; ++e) rd(f, &c.edge_data[e].x, 3);
  }
  std::fclose(f);
  return cells;
}

}  // namespace cellfile
