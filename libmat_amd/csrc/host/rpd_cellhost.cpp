// Methods of the drop-in ConvexCellHost (voronoi_defs.cxx:33-298) and the
// per-sphere power-cell adjacency (rpd_update.cxx:80-340, 409-521).  Kept
// apart from the run entry points so that librpd_amd_core.so, which a
// program that compiles LibMAT's own voronoi_defs.cxx links, defines none
// of these methods a second time.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <cstdio>
#include <cstring>
#include <iterator>
#include <mutex>
#include <thread>

#include "libmat_rpd.h"

namespace {

inline float det2f(float a11, float a12, float a21, float a22) {
  return a11 * a22 - a12 * a21;
}
inline float det3f(float a11, float a12, float a13, float a21, float a22,
                   float a23, float a31, float a32, float a33) {
  return a11 * det2f(a22, a23, a32, a33) - a21 * det2f(a12, a13, a32, a33) +
         a31 * det2f(a12, a13, a22, a23);
}

template <class T>
void set_intersection2(const std::set<T>& a, const std::set<T>& b,
                       std::set<T>& out) {
  out.clear();
  std::set_intersection(a.begin(), a.end(), b.begin(), b.end(),
                        std::inserter(out, out.begin()));
}

}  // namespace

// ------------------------------------------------------- ConvexCellHost
cfloat4 ConvexCellHost::compute_vertex_coordinates(cuchar3 t,
                                                   bool persp_divide) const {
  cfloat4 p1 = clip_trans_const(t.x), p2 = clip_trans_const(t.y),
          p3 = clip_trans_const(t.z);
  cfloat4 r;
  r.x = -det3f(p1.w, p1.y, p1.z, p2.w, p2.y, p2.z, p3.w, p3.y, p3.z);
  r.y = -det3f(p1.x, p1.w, p1.z, p2.x, p2.w, p2.z, p3.x, p3.w, p3.z);
  r.z = -det3f(p1.x, p1.y, p1.w, p2.x, p2.y, p2.w, p3.x, p3.y, p3.w);
  r.w = det3f(p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, p3.x, p3.y, p3.z);
  if (persp_divide) return cfloat4{r.x / r.w, r.y / r.w, r.z / r.w, 1};
  return r;
}

void ConvexCellHost::print_info() const {
  printf(
      "Cell of id: %d, voro_id: %d, sq_radius: %f, tet_id: %d, nb_v: %d, "
      "nb_p: %d, nb_e: %d, status: %d, thread_id: %d\n",
      id, voro_id, weight, tet_id, nb_v, nb_p, nb_e, (int)status, thread_id);
}

void ConvexCellHost::reload_active() {
  active_clipping_planes.assign(nb_p + 1, 0);
  std::map<int, std::set<int>> plane2v;
  for (int t = 0; t < nb_v; ++t)
    for (int k = 0; k < 3; ++k) {
      int p = ith_plane(t, k);
      active_clipping_planes[p]++;
      plane2v[p].insert(t);
    }
  active_edges.assign(nb_e + 1, 0);
  std::set<int> common;
  for (int e = 0; e < nb_e; ++e) {
    cuchar3 ed = edge_data[e];
    if (active_clipping_planes[ed.x] <= 0 || active_clipping_planes[ed.y] <= 0)
      continue;
    set_intersection2(plane2v.at(ed.x), plane2v.at(ed.y), common);
    if (common.size() < 2) continue;
    active_edges[e]++;
  }
  is_active_updated = true;
}

void ConvexCellHost::reload_pc_explicit() {
  pc_points.clear();
  for (int i = 0; i < nb_v; ++i) {
    cuchar4 v = ver_trans(i);
    cfloat4 p = compute_vertex_coordinates(cuchar3{v.x, v.y, v.z});
    if (std::isnan(p.x) || std::isnan(p.y) || std::isnan(p.z)) {
      is_vertex_null = true;
      return;
    }
    pc_points.push_back(cfloat3{p.x, p.y, p.z});
  }
  if (!is_active_updated) reload_active();
  pc_local_active_faces.clear();
  pc_lf2active_map.assign(nb_p, -1);
  int lf = 0;
  for (int plane = 0; plane < nb_p; ++plane) {
    if (active_clipping_planes[plane] <= 0) continue;
    std::vector<int> tab_v, tab_lp;
    for (int t = 0; t < nb_v; ++t)
      for (int k = 0; k < 3; ++k)
        if (ith_plane(t, k) == plane) {
          tab_v.push_back(t);
          tab_lp.push_back(k);
          break;
        }
    // walk the face: next of (t, slot) is the vertex whose plane at
    // (slot' + 2) % 3 equals t's plane at (slot + 1) % 3
    std::vector<int> face;
    int i = 0;
    std::vector<char> used(tab_v.size(), 0);
    while (face.size() < tab_v.size()) {
      face.push_back(tab_v[i]);
      used[i] = 1;
      int want = ith_plane(tab_v[i], (tab_lp[i] + 1) % 3);
      int nxt = -1;
      for (size_t j = 0; j < tab_v.size(); ++j)
        if (!used[j] && ith_plane(tab_v[j], (tab_lp[j] + 2) % 3) == want) {
          nxt = (int)j;
          break;
        }
      if (nxt < 0) break;
      i = nxt;
    }
    pc_local_active_faces.push_back(face);
    pc_lf2active_map[plane] = lf++;
  }
  is_pc_explicit = true;
}

// Euler sums with the reference's rounding (voronoi_defs.cxx:190-267):
// vertex terms are binary32 quotients 1.f / #adjacent cells, face and edge
// terms binary64 quotients, accumulated in binary64 (Scalar), and the
// result passes through a float
Scalar ConvexCellHost::cal_cell_euler() {
  Scalar sv = 0, sf = 0, se = 0;
  if (!is_active_updated) reload_active();
  for (int v = 0; v < nb_v; ++v) sv += 1.f / (int)ver_trans(v).w;
  for (int f = 0; f < nb_p; ++f)
    if (active_clipping_planes[f] > 0) sf += 1. / clip_trans(f).h;
  for (int e = 0; e < nb_e; ++e)
    if (active_edges[e] > 0) se += 1. / edge(e).z;
  euler = (float)(sv - se + sf);
  return euler;
}

Scalar ConvexCellHost::cal_halfplane_facet_euler(int neigh_id) {
  if (!is_active_updated) reload_active();
  int clip_id = -1;
  for (int f = 0; f < nb_p; ++f) {
    if (active_clipping_planes[f] <= 0) continue;
    cint2 id2 = clip_id2_const(f);
    if (id2.x == neigh_id || id2.y == neigh_id) {
      clip_id = f;
      break;
    }
  }
  if (clip_id < 0) throw std::runtime_error("halfplane facet not found");
  Scalar sv = 0, se = 0, sf = 1;
  for (int v = 0; v < nb_v; ++v) {
    cuchar4 t = ver_trans(v);
    if (t.x != clip_id && t.y != clip_id && t.z != clip_id) continue;
    sv += 1.f / (int)t.w;
  }
  for (int e = 0; e < nb_e; ++e) {
    cuchar3 ed = edge(e);
    if (ed.x != clip_id && ed.y != clip_id) continue;
    if (active_edges[e] <= 0) continue;
    se += 1. / ed.z;
  }
  const float facet_euler = (float)(sv - se + sf);
  return facet_euler;
}

void ConvexCellHost::print_cell_detail_euler() const {
  // (voronoi_defs.cxx:269-291) needs reload_active() first
  std::printf("--------- \n");
  for (int v = 0; v < nb_v; ++v)
    std::printf("[Detail] seed %d id %d has vertex %d with #adj_cells: %d \n",
                voro_id, id, v, (int)ver_data_trans[v].w);
  for (int e = 0; e < nb_e; ++e) {
    if (e >= (int)active_edges.size() || active_edges[e] <= 0) continue;
    std::printf("[Detail] seed %d id %d has edge %d: (%d, %d) with "
                "#adj_cells: %d\n",
                voro_id, id, e, (int)edge_data[e].x, (int)edge_data[e].y,
                (int)edge_data[e].z);
  }
  for (int f = 0; f < nb_p; ++f) {
    if (f >= (int)active_clipping_planes.size() ||
        active_clipping_planes[f] <= 0)
      continue;
    std::printf("[Detail] seed %d id %d has plane %d with #adj_cells: %f \n",
                voro_id, id, f, clip_data_trans[f].h);
  }
}

bool is_convex_cell_valid(const ConvexCellHost& cell) {
  return cell.status == success || cell.status == security_radius_not_reached;
}

// ------------------------------------------------------------ adjacency
int rpd_connected_components(const std::set<int>& to_visit,
                             const std::map<int, std::set<int>>& neighbors,
                             std::vector<std::set<int>>& out) {
  out.clear();
  std::set<int> seen;
  for (int start : to_visit) {  // ascending: components by least element
    if (seen.count(start)) continue;
    std::set<int> comp;
    std::vector<int> stack{start};
    seen.insert(start);
    while (!stack.empty()) {
      const int c = stack.back();
      stack.pop_back();
      comp.insert(c);
      auto it = neighbors.find(c);
      if (it == neighbors.end()) continue;
      for (int nb : it->second)
        if (to_visit.count(nb) && seen.insert(nb).second) stack.push_back(nb);
    }
    out.push_back(std::move(comp));
  }
  return (int)out.size();
}

namespace {

rpd_vec3 vertex_pos(const ConvexCellHost& c, int v) {
  const cuchar4 w = c.ver_trans_const(v);
  const cfloat4 p = c.compute_vertex_coordinates(cuchar3{w.x, w.y, w.z});
  return rpd_vec3{{(double)p.x, (double)p.y, (double)p.z}};
}

// the two local vertices shared by two active faces of an explicit cell
aint2 shared_vertices(const ConvexCellHost& c, int lf1, int lf2) {
  const auto& f1 = c.pc_local_active_faces.at(c.pc_lf2active_map.at(lf1));
  const auto& f2 = c.pc_local_active_faces.at(c.pc_lf2active_map.at(lf2));
  std::set<int> s1(f1.begin(), f1.end()), s2(f2.begin(), f2.end());
  std::vector<int> common;
  std::set_intersection(s1.begin(), s1.end(), s2.begin(), s2.end(),
                        std::back_inserter(common));
  if (common.size() != 2)
    throw std::runtime_error("adjacency: an active edge without two vertices");
  return aint2{{common[0], common[1]}};
}

std::vector<std::vector<rpd_v2int>> surf_groups(
    const std::map<int, std::vector<rpd_v2int>>& cell_to_surfv2fid,
    const std::vector<std::set<int>>& groups) {
  std::vector<std::vector<rpd_v2int>> out;
  for (const auto& g : groups) {
    std::vector<rpd_v2int> v;
    for (int cid : g) {
      auto it = cell_to_surfv2fid.find(cid);
      if (it != cell_to_surfv2fid.end())
        v.insert(v.end(), it->second.begin(), it->second.end());
    }
    out.push_back(std::move(v));
  }
  return out;
}

}  // namespace

std::map<int, PowerCellAdjacency> update_power_cells_adjacency(
    std::vector<ConvexCellHost>& cells, int max_surf_fid,
    const std::map<aint3, aint3>* tet_es2fe_map, int n_threads) {
  std::map<int, PowerCellAdjacency> pcs;
  for (auto& c : cells) {
    auto& pc = pcs[c.voro_id];
    pc.voro_id = c.voro_id;
    pc.cell_ids.insert(c.id);
    pc.tet_ids.insert(c.tet_id);
    if (!c.is_active_updated) c.reload_active();
    if (!c.is_pc_explicit) c.reload_pc_explicit();
  }
  std::vector<int> keys;
  for (auto& kv : pcs) keys.push_back(kv.first);
  std::map<int, int> id2pos;
  for (size_t q = 0; q < cells.size(); ++q) id2pos[cells[q].id] = (int)q;
  auto neigh_of = [](const ConvexCellHost& c, cint2 hp) {
    return hp.x == c.voro_id ? hp.y : hp.x;
  };
  auto work = [&](int sphere) {
    PowerCellAdjacency& pc = pcs.at(sphere);
    std::map<int, std::set<int>> tfid_to_cells;
    for (int cid : pc.cell_ids) {
      const ConvexCellHost& c = cells[id2pos.at(cid)];
      // facets: halfplanes by neighbour, tet faces (and their centroid on
      // the surface) by face id
      int lfa = 0;
      for (int p = 0; p < c.nb_p; ++p) {
        if (c.active_clipping_planes[p] <= 0) continue;
        cint2 hp = c.clip_id2_const(p);
        if (hp.y != -1) {
          pc.facet_neigh_to_cells[neigh_of(c, hp)].insert(cid);
        } else {
          if (hp.x <= max_surf_fid) {
            cfloat3 cen{0.f, 0.f, 0.f};
            const auto& face = c.pc_local_active_faces.at(lfa);
            for (int lv : face) {
              cen.x += c.pc_points.at(lv).x;
              cen.y += c.pc_points.at(lv).y;
              cen.z += c.pc_points.at(lv).z;
            }
            const float n = (float)face.size();
            pc.cell_to_surfv2fid[cid].push_back(rpd_v2int(
                rpd_vec3{{cen.x / n, cen.y / n, cen.z / n}}, hp.x));
          }
          tfid_to_cells[hp.x].insert(cid);
          pc.cell_to_tfids[cid].insert(hp.x);
        }
        ++lfa;
      }
      // vertices of the power cell: 3 neighbour seeds, or 2 and -1 on the
      // surface; convex-cell-only vertices are skipped
      for (int v = 0; v < c.nb_v; ++v) {
        std::set<int> seeds;
        int surf_fid = -1;
        for (int k = 0; k < 3; ++k) {
          cint2 hp = c.clip_id2_const(c.ith_plane_const(v, k));
          if (hp.y != -1)
            seeds.insert(neigh_of(c, hp));
          else if (hp.x <= max_surf_fid)
            surf_fid = hp.x;
        }
        if (seeds.size() < 2) continue;
        if (seeds.size() == 2) seeds.insert(-1);
        aint3 key3;
        int q = 0;
        for (int sd : seeds) key3[q++] = sd;
        std::sort(key3.begin(), key3.end());
        const aint2 key{{cid, v}};
        pc.vertex_2id[key] = key3;
        pc.vertex_2pos[key] = rpd_v2int(vertex_pos(c, v), surf_fid);
      }
      // edges: feature edges of the tet; power-cell edges between two
      // halfplanes with their end vertices
      for (int e = 0; e < c.nb_e; ++e) {
        if (c.active_edges[e] <= 0) continue;
        cuchar2 ef = c.edge_id2_const(e);
        cint2 h1 = c.clip_id2_const(ef.x), h2 = c.clip_id2_const(ef.y);
        if (h1.y == -1 && h2.y == -1 && tet_es2fe_map) {
          aint2 lfs{{ef.x, ef.y}};
          std::sort(lfs.begin(), lfs.end());
          auto it = tet_es2fe_map->find(aint3{{c.tet_id, lfs[0], lfs[1]}});
          if (it != tet_es2fe_map->end()) {
            const aint2 lv = shared_vertices(c, lfs[0], lfs[1]);
            const aint3& fe = it->second;  // <fe_type, fe_id, fe_line_id>
            const aint5 k5{{cid, lv[0], lv[1], fe[2], fe[1]}};
            if (fe[0] == RPD_EDGE_SE)
              pc.se_covered_lvids.insert(k5);
            else if (fe[0] == RPD_EDGE_CE)
              pc.ce_covered_lvids.insert(k5);
            // an end of the feature line inside the power cell: the first
            // halfplane at either end vertex names it
            for (int lvid : lv)
              for (int k = 0; k < 3; ++k) {
                cint2 hp = c.clip_id2_const(c.ith_plane_const(lvid, k));
                if (hp.y == -1) continue;
                pc.se_line_endpos[aint4{{cid, lvid, neigh_of(c, hp), fe[2]}}] =
                    vertex_pos(c, lvid);
                break;
              }
          }
          continue;
        }
        if (h1.y == -1 || h2.y == -1) continue;
        aint2 eu{{neigh_of(c, h1), neigh_of(c, h2)}};
        std::sort(eu.begin(), eu.end());
        pc.e_to_cells[eu].insert(cid);
        const aint2 ends = shared_vertices(c, ef.x, ef.y);
        const aint2 k1{{cid, ends[0]}}, k2{{cid, ends[1]}};
        if (!pc.vertex_2id.count(k1) || !pc.vertex_2id.count(k2))
          throw std::runtime_error(
              "adjacency: a power-cell edge ends at a non-power-cell vertex");
        pc.edge_2endvertices[eu].push_back({{k1, k2}});
      }
    }
    for (auto& kv : tfid_to_cells) {
      if (kv.second.size() != 2) continue;
      int a = *kv.second.begin(), b = *kv.second.rbegin();
      pc.cell_neighbors[a].insert(b);
      pc.cell_neighbors[b].insert(a);
    }
    // connected components (update_pc_cc_info, update_pc_facet_cc_info,
    // update_pc_edge_cc_info)
    rpd_connected_components(pc.cell_ids, pc.cell_neighbors, pc.cc_cells);
    pc.cc_surf_v2fids = surf_groups(pc.cell_to_surfv2fid, pc.cc_cells);
    for (auto& kv : pc.facet_neigh_to_cells) {
      auto& groups = pc.facet_cc_cells[kv.first];
      rpd_connected_components(kv.second, pc.cell_neighbors, groups);
      std::stable_sort(groups.begin(), groups.end(),
                       [](const std::set<int>& x, const std::set<int>& y) {
                         return x.size() < y.size();
                       });
      pc.facet_cc_surf_v2fids[kv.first] =
          surf_groups(pc.cell_to_surfv2fid, groups);
    }
    for (auto& kv : pc.e_to_cells)
      rpd_connected_components(kv.second, pc.cell_neighbors,
                               pc.edge_cc_cells[kv.first]);
  };
  int nt = n_threads > 0 ? n_threads : (int)std::thread::hardware_concurrency();
  nt = std::max(1, std::min<int>(nt, (int)keys.size()));
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&, t]() {
      for (size_t q = t; q < keys.size(); q += nt) work(keys[q]);
    });
  for (auto& th : pool) th.join();
  return pcs;
}

// ------------------------------------------------------------ export
// The faces of one cell appended to a polygon soup (io_cuda.cxx:21-141):
// every vertex of the cell becomes a point (binary32, from the cell's own
// planes), every active plane a polygon walked around its vertices --
// the next vertex holds the current one's plane (slot + 1) % 3 at its slot
// (slot + 2) % 3 -- or the fan (0, p, p + 1) of triangles; each polygon
// carries the cell's sphere and the sphere across it (-1 for a tet face).
// With is_boundary_only, only bisector faces and tet faces with an id below
// max_sf_fid are kept.  A cell with a non-finite vertex adds nothing.
void get_one_convex_cell_faces_const(
    const ConvexCellHost& c, std::vector<cfloat3>& points,
    std::vector<std::vector<unsigned>>& faces, std::vector<int>& face_sites,
    std::vector<int>& face_adj_sites, bool is_triangle, int max_sf_fid,
    bool is_boundary_only) {
  if (faces.size() != face_sites.size() ||
      faces.size() != face_adj_sites.size())
    throw std::runtime_error("faces and their site lists differ in length");
  const size_t row = points.size();
  std::vector<cfloat3> pts;
  for (int v = 0; v < c.nb_v; ++v) {
    const cuchar4 w = c.ver_trans_const(v);
    const cfloat4 p = c.compute_vertex_coordinates(cuchar3{w.x, w.y, w.z});
    if (std::isnan(p.x) || std::isnan(p.y) || std::isnan(p.z)) return;
    pts.push_back(cfloat3{p.x, p.y, p.z});
  }
  points.insert(points.end(), pts.begin(), pts.end());
  if (!c.is_active_updated)
    throw std::runtime_error("get_one_convex_cell_faces_const: "
                             "reload_active() first");
  for (int plane = 0; plane < c.nb_p; ++plane) {
    if (c.active_clipping_planes[plane] <= 0) continue;
    const cint2 hp = c.clip_id2_const(plane);
    if (is_boundary_only && !(hp.y != -1 || hp.x < max_sf_fid)) continue;
    std::vector<int> tab_v, tab_lp;
    for (int t = 0; t < c.nb_v; ++t) {
      const cuchar4 w = c.ver_trans_const(t);
      const int s = (int)w.x == plane ? 0
                    : (int)w.y == plane ? 1
                    : (int)w.z == plane ? 2 : -1;
      if (s < 0) continue;
      tab_v.push_back(t);
      tab_lp.push_back(s);
    }
    std::vector<int> cyc;
    std::vector<char> used(tab_v.size(), 0);
    int i = 0;
    while (cyc.size() < tab_v.size()) {
      cyc.push_back(tab_v[i]);
      used[i] = 1;
      const int want = c.ith_plane_const(tab_v[i], (tab_lp[i] + 1) % 3);
      int nxt = -1;
      for (size_t j = 0; j < tab_v.size(); ++j)
        if (!used[j] &&
            c.ith_plane_const(tab_v[j], (tab_lp[j] + 2) % 3) == want) {
          nxt = (int)j;
          break;
        }
      if (nxt < 0) break;
      i = nxt;
    }
    const int adj = hp.y != -1 ? (hp.x == c.voro_id ? hp.y : hp.x) : -1;
    const int n = (int)cyc.size();
    if (is_triangle) {
      for (int p = 1; p + 1 < n; ++p) {
        faces.push_back({(unsigned)(row + cyc[0]), (unsigned)(row + cyc[p]),
                         (unsigned)(row + cyc[(p + 1) % n])});
        face_sites.push_back(c.voro_id);
        face_adj_sites.push_back(adj);
      }
    } else {
      std::vector<unsigned> f(n);
      for (int p = 0; p < n; ++p) f[p] = (unsigned)(row + cyc[p]);
      faces.push_back(f);
      face_sites.push_back(c.voro_id);
      face_adj_sites.push_back(adj);
    }
  }
}

namespace {

// big-endian writes of the bgeo stream
struct BeOut {
  std::FILE* f;
  void u8(uint8_t x) { std::fputc(x, f); }
  void u16(uint16_t x) {
    u8((uint8_t)(x >> 8));
    u8((uint8_t)x);
  }
  void u32(uint32_t x) {
    u16((uint16_t)(x >> 16));
    u16((uint16_t)x);
  }
  void f32(float x) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    u32(b);
  }
  void str(const char* s) {
    const size_t n = std::strlen(s);
    u16((uint16_t)n);
    std::fwrite(s, 1, n, f);
  }
};

}  // namespace

std::array<float, 3> libmat_rpd::cell_barycenter(const ConvexCellHost& c) {
  // mean of the cell's vertices in binary32 (compute_cell_barycenter,
  // io_cuda.cxx:6-17)
  float x = 0, y = 0, z = 0;
  for (int v = 0; v < c.nb_v; ++v) {
    const cuchar4 w = c.ver_trans_const(v);
    const cfloat4 p = c.compute_vertex_coordinates(cuchar3{w.x, w.y, w.z});
    x = p.x + x;
    y = p.y + y;
    z = p.z + z;
  }
  return {x / c.nb_v, y / c.nb_v, z / c.nb_v};
}

void libmat_rpd::save_convex_cells_bgeo(
    const std::string& path, std::vector<ConvexCellHost>& cells,
    int max_sf_fid, bool is_boundary_only,
    const std::function<bool(const std::array<float, 3>&)>& skip_cell) {
  std::vector<cfloat3> points;
  std::vector<std::vector<unsigned>> faces;
  std::vector<int> sites, adj;
  for (auto& c : cells) {
    if (skip_cell && skip_cell(cell_barycenter(c))) continue;
    if (!c.is_active_updated) c.reload_active();
    get_one_convex_cell_faces_const(c, points, faces, sites, adj, false,
                                    max_sf_fid, is_boundary_only);
  }
  if (points.empty())
    throw std::runtime_error("save_convex_cells_bgeo: no points");
  std::FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("save_convex_cells_bgeo: cannot open " +
                                   path);
  BeOut o{f};
  // header of a version-5 binary geometry (io_utils.cpp:106-135): magic,
  // 'V', version, points, primitives, point groups, primitive groups,
  // point attributes beyond P, vertex attributes, primitive attributes,
  // detail attributes
  o.u32(0x4267656fu);  // "Bgeo"
  o.u8('V');
  o.u32(5);
  o.u32((uint32_t)points.size());
  o.u32((uint32_t)faces.size());
  o.u32(0);
  o.u32(0);
  o.u32(0);
  o.u32(0);
  o.u32(1);
  o.u32(0);
  // points: x, y, z, w = 1
  for (const cfloat3& p : points) {
    o.f32(p.x);
    o.f32(p.y);
    o.f32(p.z);
    o.f32(1.f);
  }
  // the primitive attribute "PrimAttr": one int (type 1), default 0
  o.str("PrimAttr");
  o.u16(1);
  o.u32(1);
  o.u32(0);
  // polygons: type 1, vertex count, closed '<', indices (16 bit up to
  // 65536 points, else 32 bit), then the attribute value
  const bool wide = points.size() > (1u << 16);
  for (size_t q = 0; q < faces.size(); ++q) {
    o.u32(1);
    o.u32((uint32_t)faces[q].size());
    o.u8('<');
    for (unsigned i : faces[q]) {
      if (wide)
        o.u32(i);
      else
        o.u16((uint16_t)i);
    }
    o.u32((uint32_t)sites[q]);
  }
  o.u8(0x00);
  o.u8(0xff);
  if (std::fclose(f) != 0)
    throw std::runtime_error("save_convex_cells_bgeo: write failed");
}
