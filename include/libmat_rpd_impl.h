// Template methods of RPD3D_GPU_T (included by libmat_rpd.h).
#pragma once
#include <algorithm>
#include <cmath>

template <class TM, class SM, class PM, class SP>
void RPD3D_GPU_T<TM, SM, PM, SP>::calculate() {
  // ---- validity and 1-rings of the live spheres (the regular-triangulation
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
  std::set<int> valid_sphere_ids;
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

  // ---- sites of (given spheres + 1-ring + 2-ring) in first-encounter
  // order over the map keys, and their neighbour matrix
  // (get_RT_partial_spheres_and_neighbors, triangulation.cxx:442-549)
  std::map<int, std::set<int>> rings;
  std::set<int> spheres_and_1rings;
  auto add_rings = [&](const std::set<int>& given, bool collect) {
    for (int sid : given) {
      if (!S[sid].is_rt_valid) continue;
      if (collect) spheres_and_1rings.insert(sid);
      for (int nb : S[sid].rt_neigh_ids) {
        rings[sid].insert(nb);
        rings[nb].insert(sid);
        if (collect) spheres_and_1rings.insert(nb);
      }
    }
  };
  add_rings(valid_sphere_ids, true);
  const std::set<int> ring1 = spheres_and_1rings;
  add_rings(ring1, false);
  // (the reference assigns MedialSphere::site_id on a by-value copy of the
  // spheres, so the caller's spheres keep site_id == -1; same here)
  std::vector<int> map_site2msphere;
  std::vector<int> site_id_of(n, -1);
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
  if (map_site2msphere.empty()) return;
  this->site_k = num_neigh_max;
  const int ns = (int)map_site2msphere.size();
  site_knn.assign((size_t)(num_neigh_max + 1) * ns, -1);
  for (int s = 0; s < ns; ++s) {
    int r = 0;
    for (int nb : site_neighbors[s]) site_knn[(size_t)r++ * ns + s] = nb;
  }
  load_partial_spheres_to_sites(map_site2msphere, spheres_and_1rings);

  // ---- device RPD over all tets, ids back to spheres, merge
  std::vector<ConvexCellHost> cells = compute_clipped_voro_diagram_GPU(
      num_itr_rpd, tet_mesh->tet_vertices, tet_mesh->tet_indices,
      tet_mesh->v2tets, tet_mesh->v_adjs, tet_mesh->e_adjs, tet_mesh->f_adjs,
      tet_mesh->f_ids, site, n_site, site_weights, site_flags, site_knn,
      site_k, site_cell_vol, site_is_transposed, 1, (int)S.size());
  const std::vector<int> map_tet_new2orig;
  update_convex_cells_voro_and_tet_ids(map_site2msphere, map_tet_new2orig,
                                       cells, is_debug);
  powercells = merge_convex_cells(valid_sphere_ids, spheres_and_1rings,
                                  powercells, cells, is_debug);
  num_itr_rpd++;
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
