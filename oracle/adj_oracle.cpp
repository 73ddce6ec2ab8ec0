// ORACLE (test infrastructure only): serial restatement of the per-sphere
// power-cell adjacency that LibMAT derives from the convex cells, and of
// the host-side cell methods it reads.  Only tests/ call it, as the checker
// of the product's update_power_cells_adjacency and ConvexCellHost methods.
//
// Follows, step by step:
//   ConvexCellHost::compute_vertex_coordinates  voronoi_defs.cxx:33-74
//   ConvexCellHost::reload_active               voronoi_defs.cxx:76-106
//   ConvexCellHost::reload_pc_explicit          voronoi_defs.cxx:108-188
//   ConvexCellHost::cal_cell_euler              voronoi_defs.cxx:190-221
//   ConvexCellHost::cal_halfplane_facet_euler   voronoi_defs.cxx:223-267
//   get_cell_v2surffid, convert_e_lfids_to_lvids rpd_update.cxx:20-70
//   get_all_voro_info                           rpd_update.cxx:80-340
//   get_CC_given_neighbors                      include/common_cxx.h:447-492
//   update_pc_cc_info, update_pc_facet_cc_info,
//   update_pc_edge_cc_info (+ surface groups)   rpd_update.cxx:409-521
//   update_power_cells (cells grouped by sphere) rpd_update.cxx:568-640
//
// Input: a cell file (binary, little-endian): "RPDC", int32 version (1),
// int32 n_cells; per cell int32 id, voro_id, tet_id, thread_id, status,
// nb_v, nb_p, nb_e, float weight, cell_vol, then nb_v x 4 uint8 dual
// vertices (3 plane indices, #adjacent cells), nb_p x 5 float planes
// (a, b, c, d, h), nb_p x 2 int32 plane ids, nb_e x 3 uint8 edges (2 plane
// indices, #adjacent cells).  Optional feature-edge file: int32 n, then n x
// 6 int32 (tet, local fid min, local fid max, type, fe id, line id).
// Output: the canonical text dump written by write_dump() below (the
// product-side test driver tests/cpp/adj_dump.cpp writes the same format).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <queue>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

// RPD_ORACLE_PLANT (tests only): plants one plausible mistake so a test can
// show that the comparison catches it -- "euler_double": the vertex terms
// of the Euler sums as binary64 quotients; "face_walk": face cycles walked
// from their last vertex; "cc_largest": components grown from the largest
// unvisited element
int planted(const char* what) {
  const char* e = std::getenv("RPD_ORACLE_PLANT");
  return e && std::strcmp(e, what) == 0;
}

using i2 = std::array<int, 2>;
using i3 = std::array<int, 3>;
using i4 = std::array<int, 4>;
using i5 = std::array<int, 5>;
using v3 = std::array<double, 3>;
using posfid = std::pair<v3, int>;

struct OCell {
  int id, voro_id, tet_id, thread_id, status, nb_v, nb_p, nb_e;
  float weight, cell_vol;
  std::vector<std::array<uint8_t, 4>> ver;
  std::vector<std::array<float, 5>> clip;
  std::vector<i2> id2;
  std::vector<std::array<uint8_t, 3>> edge;
  // derived
  std::vector<int> act_p, act_e;
  std::vector<std::array<float, 3>> pts;
  std::vector<std::vector<int>> faces;
  std::vector<int> lf2active;
};

template <class T>
void rd(FILE* f, T* p, size_t n) {
  if (n && std::fread(p, sizeof(T), n, f) != n)
    throw std::runtime_error("cell file truncated");
}

std::vector<OCell> read_cells(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) throw std::runtime_error("cannot open cell file");
  char magic[4];
  int32_t version = 0, n = 0;
  rd(f, magic, 4);
  rd(f, &version, 1);
  rd(f, &n, 1);
  if (std::memcmp(magic, "RPDC", 4) != 0 || version != 1)
    throw std::runtime_error("bad cell file");
  std::vector<OCell> cells(n);
  for (auto& c : cells) {
    int32_t h[8];
    rd(f, h, 8);
    c.id = h[0];
    c.voro_id = h[1];
    c.tet_id = h[2];
    c.thread_id = h[3];
    c.status = h[4];
    c.nb_v = h[5];
    c.nb_p = h[6];
    c.nb_e = h[7];
    rd(f, &c.weight, 1);
    rd(f, &c.cell_vol, 1);
    c.ver.resize(c.nb_v);
    c.clip.resize(c.nb_p);
    c.id2.resize(c.nb_p);
    c.edge.resize(c.nb_e);
    for (auto& v : c.ver) rd(f, v.data(), 4);
    for (auto& p : c.clip) rd(f, p.data(), 5);
    for (auto& p : c.id2) rd(f, p.data(), 2);
    for (auto& e : c.edge) rd(f, e.data(), 3);
  }
  std::fclose(f);
  return cells;
}

// ---- cell methods (voronoi_defs.cxx)
float det2(float a, float b, float c, float d) { return a * d - b * c; }
float det3(float a11, float a12, float a13, float a21, float a22, float a23,
           float a31, float a32, float a33) {
  return a11 * det2(a22, a23, a32, a33) - a21 * det2(a12, a13, a32, a33) +
         a31 * det2(a12, a13, a22, a23);
}

std::array<float, 3> vertex_xyz(const OCell& c, int v) {
  const auto& a = c.clip[c.ver[v][0]];
  const auto& b = c.clip[c.ver[v][1]];
  const auto& d = c.clip[c.ver[v][2]];
  // columns (x, y, z, w) of the three planes; Cramer's rule with the w
  // column moved in, as compute_vertex_coordinates does
  const float X = -det3(a[3], a[1], a[2], b[3], b[1], b[2], d[3], d[1], d[2]);
  const float Y = -det3(a[0], a[3], a[2], b[0], b[3], b[2], d[0], d[3], d[2]);
  const float Z = -det3(a[0], a[1], a[3], b[0], b[1], b[3], d[0], d[1], d[3]);
  const float W = det3(a[0], a[1], a[2], b[0], b[1], b[2], d[0], d[1], d[2]);
  return {X / W, Y / W, Z / W};
}

void derive(OCell& c) {
  // active planes: carried by a vertex; active edges: their two planes
  // share at least two vertices
  c.act_p.assign(c.nb_p + 1, 0);
  std::vector<std::set<int>> pv(c.nb_p + 1);
  for (int v = 0; v < c.nb_v; ++v)
    for (int k = 0; k < 3; ++k) {
      c.act_p[c.ver[v][k]]++;
      pv[c.ver[v][k]].insert(v);
    }
  c.act_e.assign(c.nb_e + 1, 0);
  for (int e = 0; e < c.nb_e; ++e) {
    const int p = c.edge[e][0], q = c.edge[e][1];
    if (c.act_p[p] <= 0 || c.act_p[q] <= 0) continue;
    int common = 0;
    for (int v : pv[p]) common += pv[q].count(v) ? 1 : 0;
    if (common >= 2) c.act_e[e]++;
  }
  // explicit faces: each active plane's vertices in cycle order; the
  // successor of (t, slot) is the first vertex whose plane at slot' + 2
  // equals t's plane at slot + 1
  c.pts.clear();
  for (int v = 0; v < c.nb_v; ++v) c.pts.push_back(vertex_xyz(c, v));
  c.faces.clear();
  c.lf2active.assign(c.nb_p, -1);
  for (int p = 0; p < c.nb_p; ++p) {
    if (c.act_p[p] <= 0) continue;
    std::vector<int> tv, tl;
    for (int v = 0; v < c.nb_v; ++v)
      for (int k = 0; k < 3; ++k)
        if (c.ver[v][k] == p) {
          tv.push_back(v);
          tl.push_back(k);
          break;
        }
    std::vector<int> face;
    size_t i = planted("face_walk") ? tv.size() - 1 : 0;
    while (face.size() < tv.size()) {
      const int want = c.ver[tv[i]][(tl[i] + 1) % 3];
      size_t j = 0;
      for (; j < tv.size(); ++j)
        if (c.ver[tv[j]][(tl[j] + 2) % 3] == want) break;
      if (j == tv.size()) throw std::runtime_error("open face cycle");
      face.push_back(tv[i]);
      i = j;
    }
    c.lf2active[p] = (int)c.faces.size();
    c.faces.push_back(face);
  }
}

double cell_euler(const OCell& c) {
  double sv = 0, se = 0, sf = 0;
  const bool dbl = planted("euler_double");
  for (int v = 0; v < c.nb_v; ++v)
    sv += dbl ? 1.0 / (int)c.ver[v][3] : 1.f / (int)c.ver[v][3];
  for (int p = 0; p < c.nb_p; ++p)
    if (c.act_p[p] > 0) sf += 1. / c.clip[p][4];
  for (int e = 0; e < c.nb_e; ++e)
    if (c.act_e[e] > 0) se += 1. / c.edge[e][2];
  const float eu = (float)(sv - se + sf);
  return eu;
}

double facet_euler(const OCell& c, int neigh) {
  int pid = -1;
  for (int p = 0; p < c.nb_p && pid < 0; ++p)
    if (c.act_p[p] > 0 && (c.id2[p][0] == neigh || c.id2[p][1] == neigh))
      pid = p;
  if (pid < 0) throw std::runtime_error("facet not found");
  double sv = 0, se = 0, sf = 1.f;
  for (int v = 0; v < c.nb_v; ++v)
    if (c.ver[v][0] == pid || c.ver[v][1] == pid || c.ver[v][2] == pid)
      sv += 1.f / (int)c.ver[v][3];
  for (int e = 0; e < c.nb_e; ++e)
    if ((c.edge[e][0] == pid || c.edge[e][1] == pid) && c.act_e[e] > 0)
      se += 1. / c.edge[e][2];
  const float fe = (float)(sv - se + sf);
  return fe;
}

// ---- per-sphere adjacency (rpd_update.cxx)
struct OPower {
  std::set<int> cell_ids, tet_ids;
  std::map<int, std::set<int>> cell_neighbors, cell_to_tfids;
  std::map<int, std::vector<posfid>> cell_to_surfv2fid;
  std::map<int, std::set<int>> facet_neigh_to_cells;
  std::map<i2, std::set<int>> e_to_cells;
  std::map<i2, std::vector<std::array<i2, 2>>> edge_2endvertices;
  std::map<i2, i3> vertex_2id;
  std::map<i2, posfid> vertex_2pos;
  std::set<i5> se_covered, ce_covered;
  std::map<i4, v3> se_line_endpos;
  std::vector<std::set<int>> cc_cells;
  std::vector<std::vector<posfid>> cc_surf;
  std::map<int, std::vector<std::set<int>>> facet_cc;
  std::map<int, std::vector<std::vector<posfid>>> facet_cc_surf;
  std::map<i2, std::vector<std::set<int>>> edge_cc;
};

int other(const OCell& c, const i2& h) { return h[0] == c.voro_id ? h[1] : h[0]; }

v3 pos_of(const OCell& c, int v) {
  const auto& p = c.pts[v];
  return v3{p[0], p[1], p[2]};
}

i2 shared_lvids(const OCell& c, int p, int q) {
  const std::vector<int>& a = c.faces.at(c.lf2active.at(p));
  const std::vector<int>& b = c.faces.at(c.lf2active.at(q));
  std::set<int> sa(a.begin(), a.end()), sb(b.begin(), b.end());
  std::vector<int> common;
  for (int v : sa)
    if (sb.count(v)) common.push_back(v);
  if (common.size() != 2) throw std::runtime_error("edge without 2 ends");
  return i2{common[0], common[1]};
}

int components(const std::set<int>& to_visit,
               const std::map<int, std::set<int>>& nb,
               std::vector<std::set<int>>& out) {
  // breadth first from the least unvisited element
  out.clear();
  std::set<int> left = to_visit, seen;
  while (!left.empty()) {
    std::set<int> one;
    std::queue<int> q;
    q.push(planted("cc_largest") ? *left.rbegin() : *left.begin());
    while (!q.empty()) {
      const int x = q.front();
      q.pop();
      if (seen.count(x)) continue;
      seen.insert(x);
      left.erase(x);
      one.insert(x);
      auto it = nb.find(x);
      if (it == nb.end()) continue;
      for (int y : it->second)
        if (to_visit.count(y)) q.push(y);
    }
    out.push_back(one);
  }
  return (int)out.size();
}

std::vector<std::vector<posfid>> group_surf(
    const std::map<int, std::vector<posfid>>& c2s,
    const std::vector<std::set<int>>& groups) {
  std::vector<std::vector<posfid>> out;
  for (const auto& g : groups) {
    std::vector<posfid> v;
    for (int cid : g) {
      auto it = c2s.find(cid);
      if (it == c2s.end()) continue;
      v.insert(v.end(), it->second.begin(), it->second.end());
    }
    out.push_back(v);
  }
  return out;
}

void sphere_info(const std::vector<const OCell*>& cells, OPower& pc,
                 int max_surf_fid, const std::map<i3, i3>& es2fe) {
  std::map<int, std::set<int>> tfid_to_cells;
  for (const OCell* cp : cells) {
    const OCell& c = *cp;
    const int cid = c.id;
    int lfa = 0;
    for (int p = 0; p < c.nb_p; ++p) {
      if (c.act_p[p] <= 0) continue;
      const i2& h = c.id2[p];
      if (h[1] != -1) {
        pc.facet_neigh_to_cells[other(c, h)].insert(cid);
      } else {
        if (h[0] <= max_surf_fid) {
          // float centroid of the face's vertices
          float sx = 0.f, sy = 0.f, sz = 0.f;
          const auto& face = c.faces.at(lfa);
          for (int lv : face) {
            sx = sx + c.pts[lv][0];
            sy = sy + c.pts[lv][1];
            sz = sz + c.pts[lv][2];
          }
          const float n = (float)face.size();
          pc.cell_to_surfv2fid[cid].push_back(
              posfid(v3{sx / n, sy / n, sz / n}, h[0]));
        }
        tfid_to_cells[h[0]].insert(cid);
        pc.cell_to_tfids[cid].insert(h[0]);
      }
      ++lfa;
    }
    for (int v = 0; v < c.nb_v; ++v) {
      std::set<int> seeds;
      int sf = -1;
      for (int k = 0; k < 3; ++k) {
        const i2& h = c.id2[c.ver[v][k]];
        if (h[1] != -1)
          seeds.insert(other(c, h));
        else if (h[0] <= max_surf_fid)
          sf = h[0];
      }
      if (seeds.size() < 2) continue;
      if (seeds.size() == 2) seeds.insert(-1);
      i3 key3{{-1, -1, -1}};
      int q = 0;
      for (int s : seeds) key3[q++] = s;
      std::sort(key3.begin(), key3.end());
      pc.vertex_2id[i2{cid, v}] = key3;
      pc.vertex_2pos[i2{cid, v}] = posfid(pos_of(c, v), sf);
    }
    for (int e = 0; e < c.nb_e; ++e) {
      if (c.act_e[e] <= 0) continue;
      const int p1 = c.edge[e][0], p2 = c.edge[e][1];
      const i2 &h1 = c.id2[p1], &h2 = c.id2[p2];
      if (h1[1] == -1 && h2[1] == -1) {
        i2 lf{{p1, p2}};
        std::sort(lf.begin(), lf.end());
        auto it = es2fe.find(i3{c.tet_id, lf[0], lf[1]});
        if (it == es2fe.end()) continue;
        const i2 lv = shared_lvids(c, lf[0], lf[1]);
        const i3& fe = it->second;
        const i5 k5{{cid, lv[0], lv[1], fe[2], fe[1]}};
        if (fe[0] == 1) pc.se_covered.insert(k5);
        else if (fe[0] == 2) pc.ce_covered.insert(k5);
        for (int x = 0; x < 2; ++x) {
          const int lvid = lv[x];
          for (int k = 0; k < 3; ++k) {
            const i2& hk = c.id2[c.ver[lvid][k]];
            if (hk[1] == -1) continue;
            pc.se_line_endpos[i4{cid, lvid, other(c, hk), fe[2]}] =
                pos_of(c, lvid);
            break;
          }
        }
        continue;
      }
      if (h1[1] == -1 || h2[1] == -1) continue;
      i2 eu{{other(c, h1), other(c, h2)}};
      std::sort(eu.begin(), eu.end());
      pc.e_to_cells[eu].insert(cid);
      const i2 ends = shared_lvids(c, p1, p2);
      const i2 k1{cid, ends[0]}, k2{cid, ends[1]};
      if (!pc.vertex_2id.count(k1) || !pc.vertex_2id.count(k2))
        throw std::runtime_error("power-cell edge at a convex-cell vertex");
      pc.edge_2endvertices[eu].push_back({k1, k2});
    }
  }
  for (const auto& kv : tfid_to_cells) {
    if (kv.second.size() == 1) continue;
    const int a = *kv.second.begin(), b = *kv.second.rbegin();
    pc.cell_neighbors[a].insert(b);
    pc.cell_neighbors[b].insert(a);
  }
  components(pc.cell_ids, pc.cell_neighbors, pc.cc_cells);
  pc.cc_surf = group_surf(pc.cell_to_surfv2fid, pc.cc_cells);
  for (const auto& kv : pc.facet_neigh_to_cells) {
    auto& g = pc.facet_cc[kv.first];
    components(kv.second, pc.cell_neighbors, g);
    std::sort(g.begin(), g.end(),
              [](const std::set<int>& a, const std::set<int>& b) {
                return a.size() < b.size();
              });
  }
  for (const auto& kv : pc.facet_cc)
    pc.facet_cc_surf[kv.first] = group_surf(pc.cell_to_surfv2fid, kv.second);
  for (const auto& kv : pc.e_to_cells)
    components(kv.second, pc.cell_neighbors, pc.edge_cc[kv.first]);
}

// ---- canonical dump
void put_set(FILE* f, const std::set<int>& s) {
  std::fputc('{', f);
  bool first = true;
  for (int x : s) {
    std::fprintf(f, first ? "%d" : " %d", x);
    first = false;
  }
  std::fputc('}', f);
}
void put_pf(FILE* f, const posfid& p) {
  std::fprintf(f, "(%a %a %a %d)", p.first[0], p.first[1], p.first[2],
               p.second);
}

void write_dump(FILE* f, const std::vector<OCell>& cells,
                const std::map<int, OPower>& pcs) {
  for (const auto& c : cells) {
    std::fprintf(f, "ce %d %a\n", c.id, cell_euler(c));
    for (int p = 0; p < c.nb_p; ++p)
      if (c.act_p[p] > 0 && c.id2[p][1] != -1)
        std::fprintf(f, "fe %d %d %a\n", c.id, other(c, c.id2[p]),
                     facet_euler(c, other(c, c.id2[p])));
  }
  for (const auto& kv : pcs) {
    const OPower& p = kv.second;
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
        std::fprintf(f, " [%d %d %d %d]", y[0][0], y[0][1], y[1][0], y[1][1]);
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
    for (const auto& x : p.se_covered)
      std::fprintf(f, "se %d %d %d %d %d\n", x[0], x[1], x[2], x[3], x[4]);
    for (const auto& x : p.ce_covered)
      std::fprintf(f, "ce5 %d %d %d %d %d\n", x[0], x[1], x[2], x[3], x[4]);
    for (const auto& x : p.se_line_endpos)
      std::fprintf(f, "endpos %d %d %d %d %a %a %a\n", x.first[0], x.first[1],
                   x.first[2], x.first[3], x.second[0], x.second[1],
                   x.second[2]);
    std::fprintf(f, "cc");
    for (const auto& g : p.cc_cells) {
      std::fputc(' ', f);
      put_set(f, g);
    }
    std::fprintf(f, "\nccsurf");
    for (const auto& g : p.cc_surf) {
      std::fprintf(f, " <");
      for (const auto& y : g) put_pf(f, y);
      std::fputc('>', f);
    }
    std::fputc('\n', f);
    for (const auto& x : p.facet_cc) {
      std::fprintf(f, "fcc %d", x.first);
      for (const auto& g : x.second) {
        std::fputc(' ', f);
        put_set(f, g);
      }
      std::fputc('\n', f);
    }
    for (const auto& x : p.facet_cc_surf) {
      std::fprintf(f, "fccsurf %d", x.first);
      for (const auto& g : x.second) {
        std::fprintf(f, " <");
        for (const auto& y : g) put_pf(f, y);
        std::fputc('>', f);
      }
      std::fputc('\n', f);
    }
    for (const auto& x : p.edge_cc) {
      std::fprintf(f, "ecc %d %d", x.first[0], x.first[1]);
      for (const auto& g : x.second) {
        std::fputc(' ', f);
        put_set(f, g);
      }
      std::fputc('\n', f);
    }
  }
}

}  // namespace

extern "C" int orc_adjacency(const char* cell_path, const char* es2fe_path,
                             int max_surf_fid, const char* out_path) {
  try {
    std::vector<OCell> cells = read_cells(cell_path);
    std::map<i3, i3> es2fe;
    if (es2fe_path && es2fe_path[0]) {
      FILE* f = std::fopen(es2fe_path, "rb");
      if (!f) return -2;
      int32_t n = 0;
      rd(f, &n, 1);
      for (int q = 0; q < n; ++q) {
        int32_t r[6];
        rd(f, r, 6);
        es2fe[i3{r[0], r[1], r[2]}] = i3{r[3], r[4], r[5]};
      }
      std::fclose(f);
    }
    for (auto& c : cells) derive(c);
    std::map<int, std::vector<const OCell*>> by_sphere;
    // a sphere's cells in ascending cell id (the order of PowerCell::
    // cell_ids, a std::set, in which update_power_cells collects them)
    for (const auto& c : cells) by_sphere[c.voro_id].push_back(&c);
    for (auto& kv : by_sphere)
      std::sort(kv.second.begin(), kv.second.end(),
                [](const OCell* a, const OCell* b) { return a->id < b->id; });
    std::map<int, OPower> pcs;
    for (auto& kv : by_sphere) {
      OPower& pc = pcs[kv.first];
      for (const OCell* c : kv.second) {
        pc.cell_ids.insert(c->id);
        pc.tet_ids.insert(c->tet_id);
      }
      sphere_info(kv.second, pc, max_surf_fid, es2fe);
    }
    FILE* f = std::fopen(out_path, "w");
    if (!f) return -3;
    write_dump(f, cells, pcs);
    std::fclose(f);
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "orc_adjacency: %s\n", e.what());
    return -1;
  }
}
