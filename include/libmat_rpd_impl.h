// Template methods of RPD3D_GPU_T (included by libmat_rpd.h).
#pragma once
#include <algorithm>
#include <cmath>

template <class TM, class SM, class PM, class SP>
void RPD3D_GPU_T<TM, SM, PM, SP>::mark_valid_spheres(
    std::set<int>& valid_sphere_ids) {
  // validity and 1-rings of the live spheres (the regular-triangulation
  // step of generate_RT_CGAL_and_mark_valid_spheres,
  // triangulation.cxx:62-144, without the 8 bounding-box points)
  std::vector<SP>& S = *all_medial_spheres;
  const int n = (int)S.size();
  double box[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
  const std::vector<float>& tv = tet_mesh->tet_vertices;
  for (size_t q = 0; q + 2 < tv.size(); q += 3)
    for (int a = 0; a < 3; ++a) {
      box[a] = std::min(box[a], (double)tv[q + a]);
      box[3 + a] = std::max(box[3 + a], (double)tv[q + a]);
    }
  const double ext =
      std::max(std::max(box[3] - box[0], box[4] - box[1]), box[5] - box[2]);
  for (int a = 0; a < 3; ++a) {
    box[a] -= 1e-6 * ext;
    box[3 + a] += 1e-6 * ext;
  }
  std::vector<std::array<double, 4>> xyzr(n);
  std::vector<char> live(n), valid;
  for (int i = 0; i < n; ++i) {
    xyzr[i] = {(double)S[i].center[0], (double)S[i].center[1],
               (double)S[i].center[2], (double)S[i].radius};
    live[i] = !S[i].is_deleted;
  }
  std::vector<std::set<int>> nbrs;
  libmat_rpd::power_neighbors(xyzr, live, box, valid, nbrs);
  valid_sphere_ids.clear();
  for (int i = 0; i < n; ++i) {
    SP& m = S[i];
    m.site_id = -1;
    m.is_rt_valid = valid[i] != 0;
    if (m.is_rt_valid) valid_sphere_ids.insert(m.id);
    m.rt_change_status = (!m.is_rt_prev_valid && m.is_rt_valid)   ? 1
                         : (m.is_rt_prev_valid && !m.is_rt_valid) ? -1
                                                                  : 0;
    m.is_rt_prev_valid = m.is_rt_valid;
    m.rt_neigh_ids_prev = m.rt_neigh_ids;
    m.rt_neigh_ids.clear();
    for (int j : nbrs[i]) m.rt_neigh_ids.insert(S[j].id);
  }
}

template <class TM, class SM, class PM, class SP>
int RPD3D_GPU_T<TM, SM, PM, SP>::partial_sites_and_neighbors(
    const std::set<int>& given, std::vector<int>& map_site2msphere,
    std::set<int>& spheres_and_1rings) {
  // sites of (given spheres + 1-ring + 2-ring) in first-encounter order
  // over the ascending map keys, and their neighbour matrix
  // (get_RT_partial_spheres_and_neighbors, triangulation.cxx:442-549)
  std::vector<SP>& S = *all_medial_spheres;
  map_site2msphere.clear();
  spheres_and_1rings.clear();
  if (given.empty()) return -1;
  std::map<int, std::set<int>> rings;
  auto add_rings = [&](const std::set<int>& from, bool collect) {
    for (int sid : from) {
      if (!S.at(sid).is_rt_valid) continue;  // not a vertex of the RT
      if (collect) spheres_and_1rings.insert(sid);
      for (int nb : S[sid].rt_neigh_ids) {
        rings[sid].insert(nb);
        rings[nb].insert(sid);
        if (collect) spheres_and_1rings.insert(nb);
      }
    }
  };
  add_rings(given, true);
  const std::set<int> ring1 = spheres_and_1rings;
  add_rings(ring1, false);
  // (the reference assigns MedialSphere::site_id on a by-value copy of the
  // spheres, so the caller's spheres keep site_id == -1; same here)
  std::vector<int> site_id_of(S.size(), -1);
  auto site_of = [&](int sid) {
    if (site_id_of[sid] != -1) return site_id_of[sid];
    site_id_of[sid] = (int)map_site2msphere.size();
    map_site2msphere.push_back(S[sid].id);
    return site_id_of[sid];
  };
  int num_neigh_max = 0;
  std::vector<std::set<int>> site_neighbors(rings.size());
  for (auto& kv : rings) {
    const int s0 = site_of(kv.first);
    num_neigh_max = std::max(num_neigh_max, (int)kv.second.size());
    for (int nb : kv.second) {
      const int s1 = site_of(nb);
      site_neighbors[s0].insert(s1);
      site_neighbors[s1].insert(s0);
    }
  }
  const int ns = (int)map_site2msphere.size();
  site_knn.assign((size_t)(num_neigh_max + 1) * ns, -1);
  for (int s = 0; s < ns; ++s) {
    int r = 0;
    for (int nb : site_neighbors[s]) site_knn[(size_t)r++ * ns + s] = nb;
  }
  return num_neigh_max;
}

template <class TM, class SM, class PM, class SP>
void RPD3D_GPU_T<TM, SM, PM, SP>::calculate() {
  std::set<int> valid_sphere_ids;
  mark_valid_spheres(valid_sphere_ids);
  std::vector<int> map_site2msphere;
  std::set<int> spheres_and_1rings;
  const int k = partial_sites_and_neighbors(valid_sphere_ids, map_site2msphere,
                                            spheres_and_1rings);
  if (map_site2msphere.empty() || k == -1) return;
  this->site_k = k;
  load_partial_spheres_to_sites(map_site2msphere, spheres_and_1rings);

  // device RPD over all tets of the resident mesh, ids back to spheres,
  // merge
  std::vector<ConvexCellHost> cells =
      gpu().run(site, n_site, site_weights, site_flags, site_knn, site_k,
                site_cell_vol);
  const std::vector<int> map_tet_new2orig;
  update_convex_cells_voro_and_tet_ids(map_site2msphere, map_tet_new2orig,
                                       cells, is_debug);
  powercells = merge_convex_cells(valid_sphere_ids, spheres_and_1rings,
                                  powercells, cells, is_debug);
  num_itr_rpd++;
}

template <class TM, class SM, class PM, class SP>
void RPD3D_GPU_T<TM, SM, PM, SP>::load_changed_spheres(
    int num_itr_global, std::set<int>& changed, std::set<int>& invalid) {
  // medial_sphere.cxx:1223-1282: spheres created in this iteration (and
  // their previous neighbours), spheres that became valid, and the valid
  // facet neighbours of spheres that became invalid
  changed.clear();
  invalid.clear();
  const std::vector<SP>& S = *all_medial_spheres;
  for (const SP& m : S) {
    if (m.is_rt_valid && m.itr_cnt == num_itr_global) {
      changed.insert(m.id);
      for (int nb : m.rt_neigh_ids_prev) changed.insert(nb);
    } else if (m.rt_change_status == 1) {
      changed.insert(m.id);
    } else if (m.rt_change_status == -1) {
      invalid.insert(m.id);
    }
  }
  for (int iid : invalid)
    for (const auto& kv : S.at(iid).pcell.facet_neigh_to_cells)
      if (S.at(kv.first).is_rt_valid) changed.insert(kv.first);
}

template <class TM, class SM, class PM, class SP>
void RPD3D_GPU_T<TM, SM, PM, SP>::load_partial_tet_given_spheres(
    const std::set<int>& given_spheres, std::vector<int>& partial_tet_indices,
    std::vector<int>& map_tet_new2orig, std::vector<int>& partial_tet_fids,
    std::vector<int>& partial_tet_f_adjs) {
  // rpd_api.cxx:482-535: the tets of the given spheres' previous cells, in
  // first-encounter order
  partial_tet_indices.clear();
  map_tet_new2orig.clear();
  partial_tet_fids.clear();
  partial_tet_f_adjs.clear();
  std::set<int> visited;
  for (int sid : given_spheres) {
    for (int t : all_medial_spheres->at(sid).pcell.tet_ids) {
      if (!visited.insert(t).second) continue;
      map_tet_new2orig.push_back(t);
      for (int j = 0; j < 4; ++j) {
        partial_tet_indices.push_back(tet_mesh->tet_indices.at(4 * t + j));
        partial_tet_fids.push_back(tet_mesh->f_ids.at(4 * t + j));
        partial_tet_f_adjs.push_back(tet_mesh->f_adjs.at(4 * t + j));
      }
    }
  }
}

template <class TM, class SM, class PM, class SP>
void RPD3D_GPU_T<TM, SM, PM, SP>::calculate_partial(int& num_itr_global,
                                                    int& num_sphere_added,
                                                    const bool all_tets) {
  // rpd_api.cxx:147-313
  if (num_sphere_added == 0) return;
  std::set<int> valid_sphere_ids;
  mark_valid_spheres(valid_sphere_ids);
  std::set<int> spheres_changed, spheres_invalid;
  load_changed_spheres(num_itr_global, spheres_changed, spheres_invalid);
  if (spheres_changed.empty()) return;
  std::vector<int> map_site2msphere;
  std::set<int> spheres_and_1rings;
  const int k = partial_sites_and_neighbors(spheres_changed, map_site2msphere,
                                            spheres_and_1rings);
  if (map_site2msphere.empty() || k == -1) return;
  this->site_k = k;
  load_partial_spheres_to_sites(map_site2msphere, spheres_and_1rings);

  // the partial tets are clipped in place on the resident mesh: the
  // cells come back with their original tet ids, so no tet id remap
  std::vector<ConvexCellHost> cells;
  if (all_tets) {
    cells = gpu().run(site, n_site, site_weights, site_flags, site_knn,
                      site_k, site_cell_vol);
  } else {
    // the previous cells of removed spheres must be re-covered too
    std::set<int> target = spheres_and_1rings;
    target.insert(spheres_invalid.begin(), spheres_invalid.end());
    std::vector<int> tet_indices, tet_fids, tet_f_adjs, partial_tets;
    load_partial_tet_given_spheres(target, tet_indices, partial_tets,
                                   tet_fids, tet_f_adjs);
    cells = gpu().run(site, n_site, site_weights, site_flags, site_knn,
                      site_k, site_cell_vol, &partial_tets);
  }
  const std::vector<int> map_tet_new2orig;
  update_convex_cells_voro_and_tet_ids(map_site2msphere, map_tet_new2orig,
                                       cells, is_debug);
  powercells = merge_convex_cells(valid_sphere_ids, spheres_and_1rings,
                                  powercells, cells, is_debug);
  update_spheres_power_cells();
  num_itr_global++;
  num_sphere_added = 0;
}

namespace rpd_detail {
// points of the adjacency into the caller's vector type (LibMAT: Vector3)
template <class V>
struct vec3_from {
  static V make(const rpd_vec3& p) { return V(p[0], p[1], p[2]); }
};
template <>
struct vec3_from<rpd_vec3> {
  static rpd_vec3 make(const rpd_vec3& p) { return p; }
};
template <class P>
P v2int_from(const rpd_v2int& x) {
  return P(vec3_from<typename P::first_type>::make(x.first), x.second);
}
template <class DV>
void copy_v2ints(const std::vector<rpd_v2int>& src, DV& dst) {
  dst.clear();
  for (const auto& x : src)
    dst.push_back(v2int_from<typename DV::value_type>(x));
}
template <class DVV>
void copy_v2int_groups(const std::vector<std::vector<rpd_v2int>>& src,
                       DVV& dst) {
  dst.clear();
  for (const auto& g : src) {
    dst.emplace_back();
    copy_v2ints(g, dst.back());
  }
}
template <class PC>
void fill_power_cell(PowerCellAdjacency& a, PC& pc) {
  pc.cell_ids = std::move(a.cell_ids);
  pc.tet_ids = std::move(a.tet_ids);
  pc.cell_neighbors = std::move(a.cell_neighbors);
  pc.cell_to_tfids = std::move(a.cell_to_tfids);
  pc.facet_neigh_to_cells = std::move(a.facet_neigh_to_cells);
  pc.e_to_cells = std::move(a.e_to_cells);
  pc.vertex_2id = std::move(a.vertex_2id);
  pc.edge_2endvertices = std::move(a.edge_2endvertices);
  pc.se_covered_lvids = std::move(a.se_covered_lvids);
  pc.ce_covered_lvids = std::move(a.ce_covered_lvids);
  pc.cc_cells = std::move(a.cc_cells);
  pc.facet_cc_cells = std::move(a.facet_cc_cells);
  pc.edge_cc_cells = std::move(a.edge_cc_cells);
  pc.cell_to_surfv2fid.clear();
  for (const auto& kv : a.cell_to_surfv2fid)
    copy_v2ints(kv.second, pc.cell_to_surfv2fid[kv.first]);
  pc.vertex_2pos.clear();
  for (const auto& kv : a.vertex_2pos)
    pc.vertex_2pos[kv.first] =
        v2int_from<typename std::decay_t<decltype(pc.vertex_2pos)>::
                       mapped_type>(kv.second);
  pc.se_line_endpos.clear();
  for (const auto& kv : a.se_line_endpos)
    pc.se_line_endpos[kv.first] = vec3_from<typename std::decay_t<
        decltype(pc.se_line_endpos)>::mapped_type>::make(kv.second);
  copy_v2int_groups(a.cc_surf_v2fids, pc.cc_surf_v2fids);
  pc.facet_cc_surf_v2fids.clear();
  for (const auto& kv : a.facet_cc_surf_v2fids)
    copy_v2int_groups(kv.second, pc.facet_cc_surf_v2fids[kv.first]);
}
template <class PC>
void clear_power_cell(PC& pc) {
  PowerCellAdjacency empty;
  fill_power_cell(empty, pc);
}
// TetMesh::tet_es2fe_map when the mesh type has one
template <class TM>
auto es2fe_of(const TM* m, int) -> decltype(&m->tet_es2fe_map) {
  return &m->tet_es2fe_map;
}
template <class TM>
const std::map<aint3, aint3>* es2fe_of(const TM*, long) {
  return nullptr;
}
}  // namespace rpd_detail

template <class TM, class SM, class PM, class SP>
void RPD3D_GPU_T<TM, SM, PM, SP>::update_spheres_power_cells() {
  // the adjacency part of update_power_cells (rpd_update.cxx:568-640):
  // every sphere's cells, tets, cell/facet/edge/vertex adjacency; a live
  // sphere without cells is marked deleted, as there
  int max_fid = max_surf_fid;
  if (max_fid < 0) {  // largest face id on the mesh boundary
    const std::vector<int>& fa = tet_mesh->f_adjs;
    for (size_t q = 0; q < fa.size(); ++q)
      if (fa[q] == 1) max_fid = std::max(max_fid, tet_mesh->f_ids[q]);
  }
  std::map<int, PowerCellAdjacency> adj = update_power_cells_adjacency(
      powercells, max_fid, rpd_detail::es2fe_of(tet_mesh, 0));
  for (SP& m : *all_medial_spheres) {
    auto it = adj.find(m.id);
    auto& pc = m.pcell;
    pc.voro_id = m.id;
    if (it == adj.end()) {
      rpd_detail::clear_power_cell(pc);
      if (!m.is_deleted) m.is_deleted = true;
      continue;
    }
    rpd_detail::fill_power_cell(it->second, pc);
  }
}

template <class TM, class SM, class PM, class SP>
void RPD3D_GPU_T<TM, SM, PM, SP>::load_partial_spheres_to_sites(
    const std::vector<int>& map_site2msphere,
    const std::set<int>& spheres_and_1rings) {
  if (map_site2msphere.empty() || spheres_and_1rings.empty()) return;
  n_site = (int)map_site2msphere.size();
  site.assign(3 * (size_t)n_site, 0.f);
  site_weights.assign(n_site, 0.f);
  site_flags.assign(n_site, 0u);
  for (int i = 0; i < n_site; ++i) {
    const SP& m = all_medial_spheres->at(map_site2msphere[i]);
    site[i] = (float)m.center[0];
    site[i + n_site] = (float)m.center[1];
    site[i + 2 * n_site] = (float)m.center[2];
    const float r = (float)m.radius;
    site_weights[i] = r * r;
    site_flags[i] = spheres_and_1rings.count(m.id) ? is_selected : no_flag;
  }
}

template <class TM, class SM, class PM, class SP>
void RPD3D_GPU_T<TM, SM, PM, SP>::update_convex_cells_voro_and_tet_ids(
    const std::vector<int>& map_site2msphere,
    const std::vector<int>& map_tet_new2orig,
    std::vector<ConvexCellHost>& cells, bool) {
  if (map_site2msphere.empty() && map_tet_new2orig.empty()) return;
  for (ConvexCellHost& c : cells) {
    if (!map_site2msphere.empty()) {
      c.voro_id = map_site2msphere.at(c.voro_id);
      for (int i = 0; i < c.nb_p; ++i) {
        cint2& id2 = c.clip_id2_data_trans[i];
        if (id2.y == -1) continue;
        id2.x = map_site2msphere.at(id2.x);
        id2.y = map_site2msphere.at(id2.y);
      }
    }
    if (!map_tet_new2orig.empty()) c.tet_id = map_tet_new2orig.at(c.tet_id);
  }
}

template <class TM, class SM, class PM, class SP>
std::vector<ConvexCellHost> RPD3D_GPU_T<TM, SM, PM, SP>::merge_convex_cells(
    const std::set<int>& valid_sphere_ids,
    const std::set<int>& spheres_and_1rings,
    const std::vector<ConvexCellHost>& prev,
    const std::vector<ConvexCellHost>& cur, bool) {
  std::vector<ConvexCellHost> merged;
  merged.reserve(prev.size() + cur.size());
  for (const ConvexCellHost& c : cur) {
    if (!spheres_and_1rings.count(c.voro_id)) continue;
    merged.push_back(c);
    merged.back().id = (int)merged.size() - 1;
  }
  for (const ConvexCellHost& c : prev) {
    if (!valid_sphere_ids.count(c.voro_id)) continue;
    if (spheres_and_1rings.count(c.voro_id)) continue;
    merged.push_back(c);
    merged.back().id = (int)merged.size() - 1;
  }
  return merged;
}
