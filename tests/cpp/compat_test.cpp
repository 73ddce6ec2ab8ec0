// Drop-in C++ surface test: ConvexCellHost layout (always) and, with
// --gpu, compute_clipped_voro_diagram_GPU / RPD3D_GPU / adjacency on the
// 10^3-grid, 500-sphere instance built through the C ABI of rpd_prep.h.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>

#include "libmat_rpd.h"
#include "rpd_prep.h"

#define CHECK(c)                                               \
  do {                                                         \
    if (!(c)) {                                                \
      std::printf("CHECK failed line %d: %s\n", __LINE__, #c); \
      return 1;                                                \
    }                                                          \
  } while (0)

// cell file (the format of oracle/adj_oracle.cpp) of the drop-in's cells
static bool write_cells(const char* path,
                        const std::vector<ConvexCellHost>& cells) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  const int32_t hdr[2] = {1, (int32_t)cells.size()};
  std::fwrite("RPDC", 1, 4, f);
  std::fwrite(hdr, 4, 2, f);
  for (const auto& c : cells) {
    const int32_t h[8] = {c.id,        c.voro_id, c.tet_id, c.thread_id,
                          (int)c.status, c.nb_v,  c.nb_p,   c.nb_e};
    std::fwrite(h, 4, 8, f);
    std::fwrite(&c.weight, 4, 1, f);
    std::fwrite(&c.cell_vol, 4, 1, f);
    for (int v = 0; v < c.nb_v; ++v) std::fwrite(&c.ver_data_trans[v].x, 1, 4, f);
    for (int p = 0; p < c.nb_p; ++p) std::fwrite(&c.clip_data_trans[p].x, 4, 5, f);
    for (int p = 0; p < c.nb_p; ++p)
      std::fwrite(&c.clip_id2_data_trans[p].x, 4, 2, f);
    for (int e = 0; e < c.nb_e; ++e) std::fwrite(&c.edge_data[e].x, 1, 3, f);
  }
  return std::fclose(f) == 0;
}

int main(int argc, char** argv) {
  // byte layout of the reference record (GCC 11, x86-64)
  CHECK(sizeof(ConvexCellHost) == 3776);
  CHECK(alignof(ConvexCellHost) == 32);
  CHECK(offsetof(ConvexCellHost, ver_data_trans) == 164);
  CHECK(offsetof(ConvexCellHost, clip_data_trans) == 576);
  CHECK(offsetof(ConvexCellHost, clip_id2_data_trans) == 2624);
  CHECK(offsetof(ConvexCellHost, edge_data) == 3136);
  CHECK(offsetof(ConvexCellHost, cell_vol) == 3748);
  std::printf("layout ok\n");
  if (argc < 2 || std::strcmp(argv[1], "--gpu") != 0) return 0;

  const int n = 10, nv = 11 * 11 * 11, nt = 6 * n * n * n;
  std::vector<double> vd(3 * nv);
  std::vector<int> tets(4 * nt);
  CHECK(rpd_synth_grid_mesh(n, 0.0, 0, vd.data(), tets.data()) == 0);
  std::vector<int> v_adj(nv), e_adj(6 * nt), f_adj(4 * nt), f_id(4 * nt);
  int n_sf = 0;
  CHECK(rpd_tet_adjacency(nv, nt, tets.data(), v_adj.data(), e_adj.data(),
                          f_adj.data(), f_id.data(), &n_sf) == 0);
  std::vector<double> sph(4 * 500);
  CHECK(rpd_synth_spheres(500, 0, 0.02, 0.08, 0, sph.data()) == 0);

  // reference-format mesh: float AoS vertices, triangular e_adjs
  RpdTetMesh mesh;
  mesh.tet_vertices.assign(vd.begin(), vd.end());
  mesh.tet_indices = tets;
  mesh.v_adjs = v_adj;
  mesh.f_adjs = f_adj;
  mesh.f_ids = f_id;
  mesh.e_adjs.assign((size_t)nv * (1 + nv) / 2 + 1, -1);
  static const int el[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
  for (int t = 0; t < nt; ++t)
    for (int k = 0; k < 6; ++k) {
      long a = tets[4 * t + el[k][0]], b = tets[4 * t + el[k][1]];
      long lo = std::min(a, b), hi = std::max(a, b);
      long idx = (lo + 1) * nv - lo * (lo + 1) / 2 - (nv - hi);
      mesh.e_adjs[idx] = e_adj[6 * t + k];
    }
  std::vector<RpdSphere> spheres(500);
  for (int i = 0; i < 500; ++i) {
    spheres[i].id = i;
    spheres[i].center = {sph[4 * i], sph[4 * i + 1], sph[4 * i + 2]};
    spheres[i].radius = sph[4 * i + 3];
  }
  RPD3D_GPU rpd;
  rpd.init(&mesh, nullptr, nullptr);
  rpd.set_spheres(&spheres);
  rpd.calculate();
  const auto& cells = rpd.powercells;
  // --gpu --dump PATH: the cells of this first calculate() to PATH
  if (argc >= 4 && std::strcmp(argv[2], "--dump") == 0)
    CHECK(write_cells(argv[3], cells));
  std::printf("powercells %zu n_site %d site_k %d\n", cells.size(), rpd.n_site,
              rpd.site_k);
  CHECK(!cells.empty());
  double vol = 0;
  for (size_t q = 0; q < cells.size(); ++q) {
    CHECK(cells[q].id == (int)q);
    CHECK(is_convex_cell_valid(cells[q]));
    CHECK(cells[q].voro_id >= 0 && cells[q].voro_id < 500);
    vol += cells[q].cell_vol;
  }
  std::printf("total volume %.9f\n", vol);
  CHECK(std::fabs(vol - 1.0) < 1e-5);

  // ---- a tet subset of the resident mesh against the reference's partial
  // mesh (the tets' rows of indices / f_adjs / f_ids, in list order) run by
  // compute_clipped_voro_diagram_GPU: the same cells, thread ids counting
  // the partial tets; tet ids original vs partial positions
  {
    std::vector<int> sub;
    for (int t = nt - 1; t >= 0; t -= 7) sub.push_back(t);
    std::vector<int> pidx, pfadj, pfid;
    for (int t : sub)
      for (int j = 0; j < 4; ++j) {
        pidx.push_back(tets[4 * t + j]);
        pfadj.push_back(f_adj[4 * t + j]);
        pfid.push_back(f_id[4 * t + j]);
      }
    std::vector<float> site = rpd.site, vol_a, vol_b;
    std::vector<ConvexCellHost> a = compute_clipped_voro_diagram_GPU(
        0, mesh.tet_vertices, pidx, mesh.v2tets, mesh.v_adjs, mesh.e_adjs,
        pfadj, pfid, site, rpd.n_site, rpd.site_weights, rpd.site_flags,
        rpd.site_knn, rpd.site_k, vol_a, true);
    libmat_rpd::ResidentMesh rm(0);
    rm.upload(mesh.tet_vertices, mesh.tet_indices, mesh.v_adjs, mesh.e_adjs,
              mesh.f_adjs, mesh.f_ids);
    std::vector<ConvexCellHost> b =
        rm.run(rpd.site, rpd.n_site, rpd.site_weights, rpd.site_flags,
               rpd.site_knn, rpd.site_k, vol_b, &sub);
    CHECK(!a.empty() && a.size() == b.size());
    std::map<std::array<int, 2>, const ConvexCellHost*> by;
    for (const auto& c : b) by[{c.tet_id, c.voro_id}] = &c;
    int same = 0;
    for (const auto& c : a) {
      auto it = by.find({sub.at(c.tet_id), c.voro_id});
      CHECK(it != by.end());
      const ConvexCellHost& d = *it->second;
      CHECK(c.thread_id == d.thread_id && c.nb_v == d.nb_v &&
            c.nb_p == d.nb_p && c.nb_e == d.nb_e && c.weight == d.weight &&
            c.cell_vol == d.cell_vol);
      CHECK(std::memcmp(c.ver_data_trans, d.ver_data_trans,
                        c.nb_v * sizeof(c.ver_data_trans[0])) == 0);
      CHECK(std::memcmp(c.clip_data_trans, d.clip_data_trans,
                        c.nb_p * sizeof(c.clip_data_trans[0])) == 0);
      CHECK(std::memcmp(c.clip_id2_data_trans, d.clip_id2_data_trans,
                        c.nb_p * sizeof(c.clip_id2_data_trans[0])) == 0);
      CHECK(std::memcmp(c.edge_data, d.edge_data,
                        c.nb_e * sizeof(c.edge_data[0])) == 0);
      ++same;
    }
    // per-site volume sums (in binary64 over cells in another order)
    CHECK(vol_a.size() == vol_b.size());
    for (size_t i = 0; i < vol_a.size(); ++i)
      CHECK(std::fabs(vol_a[i] - vol_b[i]) <= 1e-6f * std::fabs(vol_a[i]));
    std::printf("resident subset: %d cells of %zu tets identical to the "
                "partial-mesh run\n", same, sub.size());
  }
  // Euler characteristic of every (convex) cell is 1 with the reference's
  // adjacency weights summed over all cells of a tet ... per cell: V-E+F
  // with unit weights is 2 for a convex polytope
  std::vector<ConvexCellHost> cc = cells;
  int bad_euler = 0;
  for (auto& c : cc) {
    c.reload_active();
    int V = c.nb_v, E = 0, F = 0;
    for (int e = 0; e < c.nb_e; ++e) E += c.active_edges[e] > 0;
    for (int p = 0; p < c.nb_p; ++p) F += c.active_clipping_planes[p] > 0;
    bad_euler += (V - E + F) != 2;
  }
  std::printf("cells with V-E+F != 2: %d\n", bad_euler);
  CHECK(bad_euler == 0);
  auto adj = update_power_cells_adjacency(cc, n_sf - 1);
  // facet symmetry: sphere a has a halfplane facet with b in tet t iff b
  // has one with a in t
  std::set<std::array<int, 3>> fac;
  for (auto& c : cc)
    for (int p = 0; p < c.nb_p; ++p) {
      if (c.active_clipping_planes[p] <= 0) continue;
      cint2 h = c.clip_id2_const(p);
      if (h.y == -1) continue;
      int nb = h.x == c.voro_id ? h.y : h.x;
      fac.insert({c.tet_id, c.voro_id, nb});
    }
  int asym = 0;
  for (auto& f : fac) asym += !fac.count({f[0], f[2], f[1]});
  std::printf("adjacency spheres %zu, facets %zu, asymmetric %d\n",
              adj.size(), fac.size(), asym);
  CHECK(asym == 0);

  // ---- feature edges: the 12 cube edges are sharp; every tet edge on one
  // maps <tet, lfid min, lfid max> (the two faces at the edge, face k
  // opposite vertex k) -> <SE, cube edge, cube edge>
  auto on_bnd = [](double x) { return x == 0.0 || x == 1.0; };
  auto cube_edge = [&](const double* a, const double* b) {
    for (int ax = 0; ax < 3; ++ax) {
      const int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
      if (on_bnd(a[o1]) && on_bnd(a[o2]) && a[o1] == b[o1] && a[o2] == b[o2])
        return ax * 4 + 2 * (int)a[o1] + (int)a[o2];
    }
    return -1;
  };
  std::map<aint3, aint3> es2fe;
  for (int t = 0; t < nt; ++t)
    for (int k = 0; k < 6; ++k) {
      const int i = el[k][0], j = el[k][1];
      const int ce = cube_edge(&vd[3 * tets[4 * t + i]], &vd[3 * tets[4 * t + j]]);
      if (ce < 0) continue;
      int f[2], q = 0;
      for (int l = 0; l < 4; ++l)
        if (l != i && l != j) f[q++] = l;
      es2fe[aint3{{t, f[0], f[1]}}] = aint3{{RPD_EDGE_SE, ce, ce}};
    }
  auto adj2 = update_power_cells_adjacency(cc, n_sf - 1, &es2fe);
  std::map<int, const ConvexCellHost*> byid;
  for (auto& c : cc) byid[c.id] = &c;
  auto pos_of = [&](int cid, int lv) {
    const ConvexCellHost& c = *byid.at(cid);
    cuchar4 w = c.ver_trans_const(lv);
    cfloat4 p = c.compute_vertex_coordinates(cuchar3{w.x, w.y, w.z});
    return std::array<double, 3>{{p.x, p.y, p.z}};
  };
  auto bdist = [](const std::array<double, 3>& p) {
    double d = 1e9;
    for (int a = 0; a < 3; ++a)
      d = std::min(d, std::min(std::fabs(p[a]), std::fabs(1.0 - p[a])));
    return d;
  };
  auto on_cube_edge = [](const std::array<double, 3>& p, int ce) {
    const int ax = ce / 4, o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
    return std::fabs(p[o1] - (double)((ce >> 1) & 1)) < 1e-5 &&
           std::fabs(p[o2] - (double)(ce & 1)) < 1e-5;
  };
  auto unite = [](const std::vector<std::set<int>>& groups, std::set<int>& u) {
    size_t n = 0;
    for (auto& g : groups) {
      u.insert(g.begin(), g.end());
      n += g.size();
    }
    return n == u.size();  // disjoint
  };
  size_t n_surf = 0, n_se = 0, n_end = 0, n_e2v = 0;
  for (auto& kv : adj2) {
    const PowerCellAdjacency& a = kv.second;
    std::set<int> u;
    CHECK(!a.cc_cells.empty() && unite(a.cc_cells, u) && u == a.cell_ids);
    CHECK(a.cc_surf_v2fids.size() == a.cc_cells.size());
    for (auto& f : a.facet_neigh_to_cells) {
      std::set<int> uf;
      CHECK(unite(a.facet_cc_cells.at(f.first), uf) && uf == f.second);
    }
    for (auto& e : a.e_to_cells) {
      std::set<int> ue;
      CHECK(unite(a.edge_cc_cells.at(e.first), ue) && ue == e.second);
    }
    for (auto& c2 : a.cell_to_surfv2fid)
      for (auto& v2 : c2.second) {
        CHECK(v2.second >= 0 && v2.second <= n_sf - 1);
        CHECK(bdist(v2.first) < 1e-5);
        ++n_surf;
      }
    for (auto& vp : a.vertex_2pos) {
      CHECK(a.vertex_2id.count(vp.first));
      if (vp.second.second != -1) CHECK(bdist(vp.second.first) < 1e-5);
    }
    for (auto& e2 : a.edge_2endvertices)
      for (auto& ends : e2.second) {
        for (auto& key : ends) {
          const aint3& ids = a.vertex_2id.at(key);
          CHECK(std::count(ids.begin(), ids.end(), e2.first[0]) == 1);
          CHECK(std::count(ids.begin(), ids.end(), e2.first[1]) == 1);
        }
        ++n_e2v;
      }
    for (auto& se : a.se_covered_lvids) {
      CHECK(on_cube_edge(pos_of(se[0], se[1]), se[3]));
      CHECK(on_cube_edge(pos_of(se[0], se[2]), se[3]));
      ++n_se;
    }
    CHECK(a.ce_covered_lvids.empty());
    for (auto& ep : a.se_line_endpos) {
      CHECK(on_cube_edge(ep.second, ep.first[3]));
      ++n_end;
    }
  }
  std::printf("surface face centroids %zu, edge end pairs %zu, sharp-edge "
              "pieces %zu, sharp-line ends %zu\n", n_surf, n_e2v, n_se, n_end);
  CHECK(n_surf > 0 && n_e2v > 0 && n_se > 0 && n_end > 0);

  // ---- calculate_partial: move 20 spheres, add one, delete one; the
  // merged cells must equal a full recompute, cell by cell (arrays are in
  // clip order, which follows site ids, so compare active planes + volume)
  rpd.update_spheres_power_cells();
  int itr = 1, added = 0;
  for (int i = 0; i < 500; i += 25) {
    if (spheres[i].is_deleted) continue;
    spheres[i].center[0] = std::min(0.95, spheres[i].center[0] + 0.01);
    spheres[i].center[1] = std::max(0.05, spheres[i].center[1] - 0.007);
    spheres[i].itr_cnt = itr;
    ++added;
  }
  RpdSphere extra;
  extra.id = 500;
  extra.center = {0.51, 0.49, 0.5};
  extra.radius = 0.06;
  extra.itr_cnt = itr;
  spheres.push_back(extra);
  ++added;
  int removed = -1;
  for (int i = 3; i < 500; ++i)
    if (!spheres[i].is_deleted && spheres[i].is_rt_valid) {
      spheres[i].is_deleted = true;
      removed = i;
      break;
    }
  CHECK(removed >= 0);
  rpd.calculate_partial(itr, added);
  CHECK(itr == 2 && added == 0);
  std::vector<RpdSphere> spheres2 = spheres;
  RPD3D_GPU full;
  full.init(&mesh, nullptr, nullptr);
  full.set_spheres(&spheres2);
  full.calculate();
  auto canon = [](std::vector<ConvexCellHost> v) {
    std::map<std::array<int, 2>, std::pair<std::set<std::array<int, 2>>, float>>
        m;
    for (auto& c : v) {
      c.reload_active();
      auto& e = m[{c.voro_id, c.tet_id}];
      for (int p = 0; p < c.nb_p; ++p)
        if (c.active_clipping_planes[p] > 0) {
          // (min, max) of the site ids, remapped to sphere ids: the pair's
          // orientation follows the site numbering of the run
          cint2 h = c.clip_id2_const(p);
          if (h.y != -1)
            e.first.insert({std::min(h.x, h.y), std::max(h.x, h.y)});
          else
            e.first.insert({h.x, h.y});
        }
      e.second = c.cell_vol;
    }
    return m;
  };
  auto mp = canon(rpd.powercells), mf = canon(full.powercells);
  int diff = 0;
  for (auto& kv : mf) {
    auto it = mp.find(kv.first);
    if (it == mp.end() || it->second.first != kv.second.first ||
        std::fabs(it->second.second - kv.second.second) >
            1e-5f * std::max(1e-6f, kv.second.second)) {
      if (++diff <= 4 && it != mp.end()) {
        std::printf("  cell (sphere %d, tet %d): vol %.9g vs %.9g planes",
                    kv.first[0], kv.first[1], it->second.second,
                    kv.second.second);
        for (auto& h : it->second.first) std::printf(" %d/%d", h[0], h[1]);
        std::printf(" | full");
        for (auto& h : kv.second.first) std::printf(" %d/%d", h[0], h[1]);
        std::printf("\n");
      }
    }
  }
  std::printf("partial: %zu cells, full: %zu cells, differing %d, "
              "removed sphere %d has cells %d\n",
              mp.size(), mf.size(), diff, removed,
              (int)spheres[removed].pcell.cell_ids.size());
  CHECK(mp.size() == mf.size() && diff == 0);
  CHECK(spheres[removed].pcell.cell_ids.empty());
  CHECK(!spheres[500].pcell.cell_ids.empty());
  std::printf("gpu ok\n");
  return 0;
}
