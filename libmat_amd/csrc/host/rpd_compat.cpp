// C++ drop-in surface of the RPD (include/libmat_rpd.h) over the C ABI.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iterator>
#include <map>
#include <mutex>
#include <thread>

#include "libmat_rpd.h"
#include "rpd_capi.h"
#include "rpd_prep.h"

static_assert(sizeof(ConvexCellHost) == 3776, "ConvexCellHost layout");
static_assert(alignof(ConvexCellHost) == 32, "ConvexCellHost alignment");

namespace {

std::mutex g_mu;
// one context per device, for the device current on the calling thread
std::map<int, rpd_ctx*> g_ctx;

rpd_ctx* shared_ctx() {
  const int dev = rpd_current_device();
  if (dev < 0) throw std::runtime_error("RPD: no usable GPU");
  rpd_ctx*& c = g_ctx[dev];
  if (!c) {
    int rc = rpd_ctx_create(dev, &c);
    if (rc != 0) {
      c = nullptr;
      throw std::runtime_error("RPD: no usable GPU (rpd_ctx_create " +
                               std::to_string(rc) + ")");
    }
  }
  return c;
}

void check(rpd_ctx* ctx, int rc, const char* what) {
  if (rc != 0)
    throw std::runtime_error(std::string(what) + ": " + rpd_last_error(ctx));
}

// triangular-matrix edge index of get_edge_idx (convex_cell.h:46-60):
// sum_{j<=vmin} (n - j) - (n - vmax)
inline int64_t tri_edge_idx(int64_t v1, int64_t v2, int64_t n) {
  int64_t lo = std::min(v1, v2), hi = std::max(v1, v2);
  return (lo + 1) * n - lo * (lo + 1) / 2 - (n - hi);
}

// the drop-in's run configuration
rpd_run_cfg dropin_cfg() {
  rpd_run_cfg cfg{};
  cfg.cells_per_wave = 8;
  cfg.use_active_mask = 1;
  cfg.compute_geometry = 1;
  cfg.small_tier = 1;
  return cfg;
}

void check_site_arrays(const std::vector<float>& site, int n_site,
                       const std::vector<float>& site_weights,
                       const std::vector<uint>& site_flags,
                       const std::vector<int>& site_knn, int site_k) {
  if (n_site < 0 || (int64_t)site.size() < 3LL * n_site ||
      (int64_t)site_weights.size() < n_site ||
      (int64_t)site_flags.size() < n_site ||
      (int64_t)site_knn.size() < (int64_t)site_k * n_site)
    throw std::runtime_error("RPD: site arrays too short");
}

// the spheres as the C ABI takes them: x, y, z, w = r^2 in binary64,
// neighbour lists as CSR (the column-major site_knn, -1 ends a list)
struct SiteArrays {
  std::vector<double> sites;
  std::vector<uint32_t> flags;
  std::vector<int64_t> knn_off;
  std::vector<int32_t> knn_ids;
};

SiteArrays site_arrays(const std::vector<float>& site, int n_site,
                       const std::vector<float>& site_weights,
                       const std::vector<uint>& site_flags,
                       const std::vector<int>& site_knn, int site_k) {
  SiteArrays a;
  a.sites.resize(4LL * n_site);
  for (int i = 0; i < n_site; ++i) {
    a.sites[4 * i + 0] = site[i];
    a.sites[4 * i + 1] = site[i + n_site];
    a.sites[4 * i + 2] = site[i + 2 * n_site];
    a.sites[4 * i + 3] = site_weights[i];
  }
  a.flags.assign(site_flags.begin(), site_flags.begin() + n_site);
  a.knn_off.assign(n_site + 1, 0);
  for (int i = 0; i < n_site; ++i) {
    for (int r = 0; r < site_k; ++r) {
      int m = site_knn[(int64_t)r * n_site + i];
      if (m == -1) break;
      a.knn_ids.push_back(m);
    }
    a.knn_off[i + 1] = (int64_t)a.knn_ids.size();
  }
  return a;
}

// the tet mesh as the C ABI takes it (binary64 vertices, per-tet edge
// counts), checked against the lengths the reference indexes
struct MeshArrays {
  int32_t n_vert = 0, n_tet = 0;
  std::vector<double> verts;
};

MeshArrays mesh_arrays(const std::vector<float>& vertices,
                       const std::vector<int>& indices,
                       const std::vector<int>& v_adjs,
                       const std::vector<int>& tet_e_adjs,
                       const std::vector<int>& f_adjs,
                       const std::vector<int>& f_ids) {
  MeshArrays m;
  m.n_vert = (int32_t)(vertices.size() / 3);
  m.n_tet = (int32_t)(indices.size() / 4);
  if (m.n_tet <= 0 || m.n_vert <= 0)
    throw std::runtime_error("RPD: empty tet mesh");
  if ((int64_t)v_adjs.size() < m.n_vert ||
      (int64_t)f_adjs.size() < 4LL * m.n_tet ||
      (int64_t)f_ids.size() < 4LL * m.n_tet ||
      (int64_t)tet_e_adjs.size() < 6LL * m.n_tet)
    throw std::runtime_error("RPD: adjacency arrays too short");
  m.verts.assign(vertices.begin(), vertices.begin() + 3LL * m.n_vert);
  return m;
}

// per-tet edge counts from the triangular n_vert^2 matrix of get_edge_idx
std::vector<int> tet_edge_counts(const std::vector<int>& indices,
                                 const std::vector<int>& e_adjs,
                                 int64_t n_vert) {
  const int64_t n_tet = (int64_t)indices.size() / 4;
  std::vector<int> tet_e(6 * n_tet);
  static const int el[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
  for (int64_t t = 0; t < n_tet; ++t)
    for (int k = 0; k < 6; ++k) {
      int64_t idx = tri_edge_idx(indices[4 * t + el[k][0]],
                                 indices[4 * t + el[k][1]], n_vert);
      if (idx < 0 || idx >= (int64_t)e_adjs.size())
        throw std::runtime_error("RPD: e_adjs index out of range");
      tet_e[6 * t + k] = e_adjs[idx];
    }
  return tet_e;
}

// The last run's cells on ctx as ConvexCellHost records, planes in
// binary32 with the reference's arithmetic.  tet_pos (nullable): position
// of each tet in the clipped tet list of a partial run (the reference's
// thread index counts tets of the partial mesh); else the tet id itself.
std::vector<ConvexCellHost> cells_of_run(
    rpd_ctx* ctx, int32_t n_tet, const std::vector<float>& vertices,
    const std::vector<int>& indices, const std::vector<int>& f_adjs,
    const std::vector<int>& f_ids, const std::vector<float>& site, int n_site,
    const std::vector<float>& site_weights, std::vector<float>& site_cell_vol,
    const std::vector<int>* tet_pos) {
  rpd_stats st;
  rpd_get_stats(ctx, &st);
  std::vector<rpd_cell> cells(st.n_cells);
  std::vector<uint32_t> words(st.n_words);
  std::vector<double> geom(4 * st.n_cells);
  std::vector<double> areas(st.n_areas);
  check(ctx,
        rpd_download_cells(ctx, cells.data(), words.data(), geom.data(),
                           areas.data()),
        "rpd_download_cells");
  std::vector<int64_t> rel_off(n_tet + 1);
  std::vector<int32_t> rel_ids(std::max<int64_t>(st.n_candidates, 1));
  check(ctx, rpd_download_relation(ctx, rel_off.data(), rel_ids.data()),
        "rpd_download_relation");
  const int tet_k = st.tet_k;

  // float planes with the reference's arithmetic (tri2plane,
  // common_cuda.h:248-254; new_plane, convex_cell.cu:561-592)
  static const int fl[4][3] = {{2, 1, 3}, {0, 2, 3}, {1, 0, 3}, {0, 1, 2}};
  auto tet_plane = [&](int t, int k) {
    const float* a = &vertices[3LL * indices[4LL * t + fl[k][0]]];
    const float* b = &vertices[3LL * indices[4LL * t + fl[k][1]]];
    const float* c = &vertices[3LL * indices[4LL * t + fl[k][2]]];
    float ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
    float vx = c[0] - a[0], vy = c[1] - a[1], vz = c[2] - a[2];
    float nx = uy * vz - uz * vy, ny = uz * vx - ux * vz,
          nz = ux * vy - uy * vx;
    return cfloat4{nx, ny, nz, -(a[0] * nx + a[1] * ny + a[2] * nz)};
  };
  auto bisector = [&](int s, int m) {
    float sx = site[s], sy = site[s + n_site], sz = site[s + 2 * n_site],
          sw = site_weights[s];
    float bx = site[m], by = site[m + n_site], bz = site[m + 2 * n_site],
          bw = site_weights[m];
    float dx = sx - bx, dy = sy - by, dz = sz - bz;
    float ax = sx + bx, ay = sy + by, az = sz + bz;
    float dot = ax * dx + ay * dy + az * dz + (bw - sw);
    return cfloat4{dx, dy, dz, -dot / 2.f};
  };

  site_cell_vol.assign(n_site, 0.f);
  std::vector<double> vol_acc(n_site, 0.0);
  std::vector<ConvexCellHost> out(st.n_cells);
  for (int64_t q = 0; q < st.n_cells; ++q) {
    const rpd_cell& c = cells[q];
    ConvexCellHost& h = out[q];
    const int t = c.tet_id, s = c.seed;
    // the reference's thread index: tet * tet_k + slot of s in tet_knn
    int slot = 0;
    for (int64_t r = rel_off[t]; r < rel_off[t + 1]; ++r, ++slot)
      if (rel_ids[r] == s) break;
    h.status = success;
    h.thread_id = (tet_pos ? (*tet_pos)[t] : t) * tet_k + slot;
    h.voro_id = s;
    h.tet_id = t;
    h.weight = site_weights[s];
    h.is_active = true;
    h.nb_v = c.nb_v;
    h.nb_p = c.nb_p;
    h.nb_e = c.nb_e;
    const uint32_t* w = words.data() + c.word_off;
    for (int v = 0; v < c.nb_v; ++v)
      h.ver_data_trans[v] = cuchar4{(uchar)(w[v] & 255),
                                    (uchar)((w[v] >> 8) & 255),
                                    (uchar)((w[v] >> 16) & 255),
                                    (uchar)(w[v] >> 24)};
    const int32_t* src = reinterpret_cast<const int32_t*>(w + c.nb_v);
    for (int p = 0; p < c.nb_p; ++p) {
      cfloat4 pl;
      float hh;
      if (src[p] < 0) {
        const int k = -1 - src[p];
        pl = tet_plane(t, k);
        hh = (float)f_adjs[4LL * t + k];
        h.clip_id2_data_trans[p] = cint2{f_ids[4LL * t + k], -1};
      } else {
        pl = bisector(s, src[p]);
        hh = F_CELL_ADJ_DEFAULT;
        h.clip_id2_data_trans[p] =
            cint2{std::min(s, src[p]), std::max(s, src[p])};
      }
      h.clip_data_trans[p] = cfloat5{pl.x, pl.y, pl.z, pl.w, hh};
    }
    const uint32_t* e = w + c.nb_v + c.nb_p;
    for (int k = 0; k < c.nb_e; ++k)
      h.edge_data[k] = cuchar3{(uchar)(e[k] & 255), (uchar)((e[k] >> 8) & 255),
                               (uchar)((e[k] >> 16) & 255)};
    h.euler = UNK_FLOAT;
    h.cell_vol = (float)geom[4 * q];
    vol_acc[s] += geom[4 * q];
    h.id = (int)q;
  }
  for (int i = 0; i < n_site; ++i) site_cell_vol[i] = (float)vol_acc[i];
  return out;
}


std::vector<ConvexCellHost> run_rpd(
    const std::vector<float>& vertices, const std::vector<int>& indices,
    const std::vector<int>& v_adjs, const std::vector<int>& tet_e_adjs,
    const std::vector<int>& f_adjs, const std::vector<int>& f_ids,
    const std::vector<float>& site, int n_site,
    const std::vector<float>& site_weights, const std::vector<uint>& site_flags,
    const std::vector<int>& site_knn, int site_k,
    std::vector<float>& site_cell_vol) {
  const MeshArrays m =
      mesh_arrays(vertices, indices, v_adjs, tet_e_adjs, f_adjs, f_ids);
  check_site_arrays(site, n_site, site_weights, site_flags, site_knn, site_k);
  const SiteArrays s =
      site_arrays(site, n_site, site_weights, site_flags, site_knn, site_k);
  rpd_instance_desc d;
  d.n_vert = m.n_vert;
  d.n_tet = m.n_tet;
  d.n_site = n_site;
  d.verts = m.verts.data();
  d.tets = indices.data();
  d.v_adj = v_adjs.data();
  d.e_adj = tet_e_adjs.data();
  d.f_adj = f_adjs.data();
  d.f_id = f_ids.data();
  d.sites = s.sites.data();
  d.flags = s.flags.data();
  d.knn_off = s.knn_off.data();
  d.knn_ids = s.knn_ids.data();

  std::lock_guard<std::mutex> lock(g_mu);
  rpd_ctx* ctx = shared_ctx();
  check(ctx, rpd_upload(ctx, &d), "rpd_upload");
  const rpd_run_cfg cfg = dropin_cfg();
  check(ctx, rpd_run(ctx, &cfg), "rpd_run");
  return cells_of_run(ctx, m.n_tet, vertices, indices, f_adjs, f_ids, site,
                      n_site, site_weights, site_cell_vol, nullptr);
}

}  // namespace

// -------------------------------------------------------- entry points
std::vector<ConvexCellHost> compute_clipped_voro_diagram_GPU(
    const int, const std::vector<float>& vertices,
    const std::vector<int>& indices, const std::map<int, std::set<int>>&,
    const std::vector<int>& v_adjs, const std::vector<int>& e_adjs,
    const std::vector<int>& f_adjs, const std::vector<int>& f_ids,
    std::vector<float>& site, const int n_site,
    const std::vector<float>& site_weights, const std::vector<uint>& site_flags,
    const std::vector<int>& site_knn, const int site_k,
    std::vector<float>& site_cell_vol, const bool, int, int) {
  const std::vector<int> tet_e =
      tet_edge_counts(indices, e_adjs, (int64_t)vertices.size() / 3);
  return run_rpd(vertices, indices, v_adjs, tet_e, f_adjs, f_ids, site,
                 n_site, site_weights, site_flags, site_knn, site_k,
                 site_cell_vol);
}

std::vector<ConvexCellHost> compute_clipped_voro_diagram_GPU_tet_eadj(
    const std::vector<float>& vertices, const std::vector<int>& indices,
    const std::vector<int>& v_adjs, const std::vector<int>& tet_e_adjs,
    const std::vector<int>& f_adjs, const std::vector<int>& f_ids,
    std::vector<float>& site, const int n_site,
    const std::vector<float>& site_weights, const std::vector<uint>& site_flags,
    const std::vector<int>& site_knn, const int site_k,
    std::vector<float>& site_cell_vol) {
  return run_rpd(vertices, indices, v_adjs, tet_e_adjs, f_adjs, f_ids, site,
                 n_site, site_weights, site_flags, site_knn, site_k,
                 site_cell_vol);
}

void libmat_rpd::power_neighbors(
    const std::vector<std::array<double, 4>>& xyzr,
    const std::vector<char>& live, const double box[6],
    std::vector<char>& valid, std::vector<std::set<int>>& neighbors) {
  const int n = (int)xyzr.size();
  std::vector<int> idx;
  std::vector<double> xyzw;
  for (int i = 0; i < n; ++i) {
    if (!live[i]) continue;
    idx.push_back(i);
    xyzw.insert(xyzw.end(), {xyzr[i][0], xyzr[i][1], xyzr[i][2],
                             xyzr[i][3] * xyzr[i][3]});
  }
  rpd_nbr* h = nullptr;
  int rc = rpd_power_neighbors((int32_t)idx.size(), xyzw.data(), box, 0, &h);
  if (rc != 0) throw std::runtime_error("rpd_power_neighbors failed");
  const int m = (int)idx.size();
  std::vector<int32_t> v(m);
  std::vector<int64_t> off(m + 1);
  std::vector<int32_t> ids(std::max<int64_t>(rpd_nbr_total(h), 1));
  rpd_nbr_get(h, v.data(), off.data(), ids.data());
  rpd_nbr_free(h);
  valid.assign(n, 0);
  neighbors.assign(n, {});
  for (int a = 0; a < m; ++a) {
    valid[idx[a]] = v[a] != 0;
    for (int64_t q = off[a]; q < off[a + 1]; ++q)
      neighbors[idx[a]].insert(idx[ids[q]]);
  }
}

// ------------------------------------------------------- resident mesh
namespace libmat_rpd {

ResidentMesh::ResidentMesh(int device)
    : device_(device >= 0 ? device : rpd_current_device()) {}

ResidentMesh::~ResidentMesh() {
  if (ctx_) rpd_ctx_destroy(ctx_);
}

void ResidentMesh::upload(const std::vector<float>& vertices,
                          const std::vector<int>& indices,
                          const std::vector<int>& v_adjs,
                          const std::vector<int>& e_adjs,
                          const std::vector<int>& f_adjs,
                          const std::vector<int>& f_ids) {
  const std::vector<int> tet_e =
      tet_edge_counts(indices, e_adjs, (int64_t)vertices.size() / 3);
  const MeshArrays m =
      mesh_arrays(vertices, indices, v_adjs, tet_e, f_adjs, f_ids);
  if (!ctx_) {
    const int rc = rpd_ctx_create(device_, &ctx_);
    if (rc != 0) {
      ctx_ = nullptr;
      throw std::runtime_error("RPD: no usable GPU (rpd_ctx_create " +
                               std::to_string(rc) + ")");
    }
  }
  n_tet_ = 0;
  rpd_instance_desc d{};
  d.n_vert = m.n_vert;
  d.n_tet = m.n_tet;
  d.n_site = 0;  // mesh only: the spheres come with each run
  d.verts = m.verts.data();
  d.tets = indices.data();
  d.v_adj = v_adjs.data();
  d.e_adj = tet_e.data();
  d.f_adj = f_adjs.data();
  d.f_id = f_ids.data();
  check(ctx_, rpd_upload(ctx_, &d), "rpd_upload");
  vertices_.assign(vertices.begin(), vertices.begin() + 3LL * m.n_vert);
  indices_.assign(indices.begin(), indices.begin() + 4LL * m.n_tet);
  f_adjs_.assign(f_adjs.begin(), f_adjs.begin() + 4LL * m.n_tet);
  f_ids_.assign(f_ids.begin(), f_ids.begin() + 4LL * m.n_tet);
  n_tet_ = m.n_tet;
}

std::vector<ConvexCellHost> ResidentMesh::run(
    const std::vector<float>& site, int n_site,
    const std::vector<float>& site_weights, const std::vector<uint>& site_flags,
    const std::vector<int>& site_knn, int site_k,
    std::vector<float>& site_cell_vol, const std::vector<int>* tets) {
  if (!ctx_ || n_tet_ <= 0)
    throw std::runtime_error("RPD: ResidentMesh::run before upload");
  check_site_arrays(site, n_site, site_weights, site_flags, site_knn, site_k);
  const SiteArrays s =
      site_arrays(site, n_site, site_weights, site_flags, site_knn, site_k);
  check(ctx_,
        rpd_update_sites(ctx_, n_site, s.sites.data(), s.flags.data(),
                         s.knn_off.data(), s.knn_ids.data()),
        "rpd_update_sites");
  std::vector<int> pos;
  if (tets) {
    if (tets->empty()) {  // nothing to clip
      site_cell_vol.assign(n_site, 0.f);
      return {};
    }
    pos.assign(n_tet_, 0);
    for (size_t i = 0; i < tets->size(); ++i) {
      const int t = (*tets)[i];
      if (t < 0 || t >= n_tet_)
        throw std::runtime_error("RPD: tet id out of range");
      pos[t] = (int)i;
    }
  }
  check(ctx_,
        rpd_set_tet_subset(ctx_, tets ? (int32_t)tets->size() : 0,
                           tets ? tets->data() : nullptr),
        "rpd_set_tet_subset");
  const rpd_run_cfg cfg = dropin_cfg();
  check(ctx_, rpd_run(ctx_, &cfg), "rpd_run");
  return cells_of_run(ctx_, n_tet_, vertices_, indices_, f_adjs_, f_ids_, site,
                      n_site, site_weights, site_cell_vol,
                      tets ? &pos : nullptr);
}

}  // namespace libmat_rpd
