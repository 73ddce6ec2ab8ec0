// Host-code sanitizer driver (AddressSanitizer + UndefinedBehaviorSanitizer,
// CPU only): exercises the library's host C++ -- input preparation
// (rpd_prep.cpp: grids, spheres, tet adjacency, the power-diagram neighbour
// builder), the drop-in's run glue (rpd_compat.cpp) and its cell methods and
// adjacency (rpd_cellhost.cpp) -- and the oracle (rpd_oracle.cpp,
// adj_oracle.cpp), compiled from source with -fsanitize=address,undefined.
// The device entry points that rpd_compat.cpp calls are replaced below by
// definitions that report "no GPU"; nothing here runs device code.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

#include "libmat_rpd.h"
#include "rpd_capi.h"
#include "rpd_prep.h"

// ---- the oracle's C API (oracle/rpd_oracle.cpp, oracle/adj_oracle.cpp)
extern "C" {
typedef struct {
  int32_t n_vert, n_tet, n_site;
  const double* verts;
  const int32_t* tets;
  const int32_t* v_adj;
  const int32_t* e_adj;
  const int32_t* f_adj;
  const int32_t* f_id;
  const double* sites;
  const uint32_t* flags;
  const int64_t* knn_off;
  const int32_t* knn_ids;
} orc_input;
typedef struct orc_result orc_result;
orc_result* orc_run(const orc_input* in, const int32_t* tet_list,
                    int32_t n_list, int32_t mode, int32_t brute_force,
                    int32_t n_threads);
void orc_free(orc_result* r);
int64_t orc_n_cells(const orc_result* r);
int64_t orc_n_rel(const orc_result* r);
int64_t orc_cell_bytes(void);
void orc_get_cells(const orc_result* r, void* dst);
void orc_get_relation(const orc_result* r, int64_t* off, int32_t* ids);
int orc_adjacency(const char* cell_path, const char* es2fe_path,
                  int max_surf_fid, const char* out_path);

// ---- device entry points used by rpd_compat.cpp: no GPU in the sanitizer binary
int rpd_ctx_create(int32_t, rpd_ctx** out) {
  *out = nullptr;
  return -3;
}
void rpd_ctx_destroy(rpd_ctx*) {}
const char* rpd_last_error(const rpd_ctx*) { return "no GPU (host build)"; }
int rpd_upload(rpd_ctx*, const rpd_instance_desc*) { return -3; }
int rpd_run(rpd_ctx*, const rpd_run_cfg*) { return -3; }
int rpd_get_stats(const rpd_ctx*, rpd_stats*) { return -3; }
int rpd_download_cells(rpd_ctx*, rpd_cell*, uint32_t*, double*, double*) {
  return -3;
}
int rpd_download_relation(rpd_ctx*, int64_t*, int32_t*) { return -3; }
int rpd_update_sites(rpd_ctx*, int32_t, const double*, const uint32_t*,
                     const int64_t*, const int32_t*) {
  return -3;
}
int rpd_set_tet_subset(rpd_ctx*, int32_t, const int32_t*) { return -3; }
int32_t rpd_current_device(void) { return -1; }
}

// mirror of the oracle's CellOut
struct CellOut {
  int32_t tet, seed;
  uint8_t status, nb_v, nb_p, nb_e;
  uint32_t ver[96];
  int32_t src[64];
  uint32_t edge[152];
  double geo[4];
  double area[64];
};

#define CHECK(c)                                               \
  do {                                                         \
    if (!(c)) {                                                \
      std::printf("CHECK failed line %d: %s\n", __LINE__, #c); \
      return 1;                                                \
    }                                                          \
  } while (0)

int main() {
  CHECK(orc_cell_bytes() == (int64_t)sizeof(CellOut));
  const int n = 5, nv = (n + 1) * (n + 1) * (n + 1), nt = 6 * n * n * n;
  std::vector<double> verts(3 * nv);
  std::vector<int32_t> tets(4 * nt);
  CHECK(rpd_synth_grid_mesh(n, 0.1, 3, verts.data(), tets.data()) == 0);
  std::vector<int32_t> v_adj(nv), e_adj(6 * nt), f_adj(4 * nt), f_id(4 * nt);
  int32_t n_sf = 0;
  CHECK(rpd_tet_adjacency(nv, nt, tets.data(), v_adj.data(), e_adj.data(),
                          f_adj.data(), f_id.data(), &n_sf) == 0);
  const int ns = 90;
  std::vector<double> sph(4 * ns);
  CHECK(rpd_synth_spheres(ns, 2, 0.02, 0.08, 5, sph.data()) == 0);
  std::vector<double> xyzw(sph);
  for (int i = 0; i < ns; ++i) xyzw[4 * i + 3] *= xyzw[4 * i + 3];
  const double box[6] = {-1e-6, -1e-6, -1e-6, 1 + 1e-6, 1 + 1e-6, 1 + 1e-6};
  rpd_nbr* h = nullptr;
  CHECK(rpd_power_neighbors(ns, xyzw.data(), box, 2, &h) == 0);
  const int64_t total = rpd_nbr_total(h);
  std::vector<int32_t> valid(ns), ids(std::max<int64_t>(total, 1));
  std::vector<int64_t> off(ns + 1);
  CHECK(rpd_nbr_get(h, valid.data(), off.data(), ids.data()) == 0);
  rpd_nbr_free(h);
  // sites = valid spheres in ascending order, lists remapped
  std::vector<int> remap(ns, -1), keep;
  for (int i = 0; i < ns; ++i)
    if (valid[i]) {
      remap[i] = (int)keep.size();
      keep.push_back(i);
    }
  const int nsite = (int)keep.size();
  CHECK(nsite > 10);
  std::vector<double> sites(4 * nsite);
  std::vector<uint32_t> flags(nsite, 1);
  std::vector<int64_t> koff(nsite + 1, 0);
  std::vector<int32_t> kids;
  for (int s = 0; s < nsite; ++s) {
    const int i = keep[s];
    for (int a = 0; a < 4; ++a) sites[4 * s + a] = xyzw[4 * i + a];
    for (int64_t q = off[i]; q < off[i + 1]; ++q) kids.push_back(remap[ids[q]]);
    koff[s + 1] = (int64_t)kids.size();
  }
  orc_input in{nv,           nt,          nsite,        verts.data(),
               tets.data(),  v_adj.data(), e_adj.data(), f_adj.data(),
               f_id.data(),  sites.data(), flags.data(), koff.data(),
               kids.data()};
  orc_result* r1 = orc_run(&in, nullptr, 0, 1, 0, 2);  // reference arithmetic
  const int64_t n1 = orc_n_cells(r1);
  orc_free(r1);
  orc_result* r = orc_run(&in, nullptr, 0, 0, 0, 2);   // exact
  const int64_t nc = orc_n_cells(r);
  CHECK(nc > 0 && n1 > 0);
  std::vector<CellOut> co(nc);
  orc_get_cells(r, co.data());
  std::vector<int64_t> roff(nt + 1);
  std::vector<int32_t> rids(std::max<int64_t>(orc_n_rel(r), 1));
  orc_get_relation(r, roff.data(), rids.data());
  orc_free(r);
  double vol = 0;
  for (const auto& c : co) vol += c.geo[0];
  CHECK(std::fabs(vol - 1.0) < 1e-9);

  // ConvexCellHost records (binary32 planes, plane ids) and the host
  // methods and adjacency the consumers call
  static const int fl[4][3] = {{2, 1, 3}, {0, 2, 3}, {1, 0, 3}, {0, 1, 2}};
  std::vector<ConvexCellHost> cells(nc);
  for (int64_t q = 0; q < nc; ++q) {
    const CellOut& o = co[q];
    ConvexCellHost& c = cells[q];
    c.status = success;
    c.id = (int)q;
    c.voro_id = o.seed;
    c.tet_id = o.tet;
    c.thread_id = 0;
    c.weight = (float)sites[4 * o.seed + 3];
    c.is_active = true;
    c.nb_v = o.nb_v;
    c.nb_p = o.nb_p;
    c.nb_e = o.nb_e;
    c.cell_vol = (float)o.geo[0];
    for (int v = 0; v < o.nb_v; ++v) {
      const uint32_t w = o.ver[v];
      c.ver_data_trans[v] = cuchar4{(uchar)(w & 255), (uchar)((w >> 8) & 255),
                                    (uchar)((w >> 16) & 255), (uchar)(w >> 24)};
    }
    for (int p = 0; p < o.nb_p; ++p) {
      if (o.src[p] < 0) {
        const int k = -1 - o.src[p];
        const double* A = &verts[3 * tets[4 * o.tet + fl[k][0]]];
        const double* B = &verts[3 * tets[4 * o.tet + fl[k][1]]];
        const double* C = &verts[3 * tets[4 * o.tet + fl[k][2]]];
        float ux = (float)(B[0] - A[0]), uy = (float)(B[1] - A[1]),
              uz = (float)(B[2] - A[2]);
        float vx = (float)(C[0] - A[0]), vy = (float)(C[1] - A[1]),
              vz = (float)(C[2] - A[2]);
        float nx = uy * vz - uz * vy, ny = uz * vx - ux * vz,
              nz = ux * vy - uy * vx;
        float d = -((float)A[0] * nx + (float)A[1] * ny + (float)A[2] * nz);
        c.clip_data_trans[p] = cfloat5{nx, ny, nz, d, (float)f_adj[4 * o.tet + k]};
        c.clip_id2_data_trans[p] = cint2{f_id[4 * o.tet + k], -1};
      } else {
        const double* S = &sites[4 * o.seed];
        const double* B = &sites[4 * o.src[p]];
        float dx = (float)(S[0] - B[0]), dy = (float)(S[1] - B[1]),
              dz = (float)(S[2] - B[2]);
        float dot = (float)(S[0] + B[0]) * dx + (float)(S[1] + B[1]) * dy +
                    (float)(S[2] + B[2]) * dz + (float)(B[3] - S[3]);
        c.clip_data_trans[p] = cfloat5{dx, dy, dz, -dot / 2.f, 1.f};
        c.clip_id2_data_trans[p] =
            cint2{std::min(o.seed, o.src[p]), std::max(o.seed, o.src[p])};
      }
    }
    for (int e = 0; e < o.nb_e; ++e) {
      const uint32_t w = o.edge[e];
      c.edge_data[e] = cuchar3{(uchar)(w & 255), (uchar)((w >> 8) & 255),
                               (uchar)((w >> 16) & 255)};
    }
  }
  std::map<aint3, aint3> es2fe;
  auto adj = update_power_cells_adjacency(cells, n_sf - 1, &es2fe, 2);
  CHECK(!adj.empty());
  double euler = 0;
  for (auto& c : cells) {
    euler += c.cal_cell_euler();
    for (int p = 0; p < c.nb_p; ++p)
      if (c.active_clipping_planes[p] > 0 && c.clip_id2_const(p).y != -1) {
        const cint2 hp = c.clip_id2_const(p);
        euler += c.cal_halfplane_facet_euler(hp.x == c.voro_id ? hp.y : hp.x);
      }
    CHECK(c.is_pc_explicit && !c.is_vertex_null);
  }
  CHECK(std::isfinite(euler));
  std::vector<int> comp_in;
  std::vector<std::set<int>> groups;
  std::set<int> all;
  for (int q = 0; q < (int)nc; ++q) all.insert(q);
  std::map<int, std::set<int>> none;
  CHECK(rpd_connected_components(all, none, groups) == (int)nc);

  // the oracle's adjacency on a cell file of the same cells
  char path[] = "/tmp/rpd_san_cellsXXXXXX";
  const int fd = mkstemp(path);
  CHECK(fd >= 0);
  FILE* f = fdopen(fd, "wb");
  const int32_t hdr[2] = {1, (int32_t)nc};
  std::fwrite("RPDC", 1, 4, f);
  std::fwrite(hdr, 4, 2, f);
  for (const auto& c : cells) {
    const int32_t hh[8] = {c.id,   c.voro_id, c.tet_id, c.thread_id,
                           4,      c.nb_v,    c.nb_p,   c.nb_e};
    std::fwrite(hh, 4, 8, f);
    std::fwrite(&c.weight, 4, 1, f);
    std::fwrite(&c.cell_vol, 4, 1, f);
    for (int v = 0; v < c.nb_v; ++v) std::fwrite(&c.ver_data_trans[v].x, 1, 4, f);
    for (int p = 0; p < c.nb_p; ++p) std::fwrite(&c.clip_data_trans[p].x, 4, 5, f);
    for (int p = 0; p < c.nb_p; ++p)
      std::fwrite(&c.clip_id2_data_trans[p].x, 4, 2, f);
    for (int e = 0; e < c.nb_e; ++e) std::fwrite(&c.edge_data[e].x, 1, 3, f);
  }
  std::fclose(f);
  std::string out = std::string(path) + ".txt";
  const int rc = orc_adjacency(path, "", n_sf - 1, out.c_str());
  std::remove(path);
  std::remove(out.c_str());
  CHECK(rc == 0);
  std::printf("san ok: %d tets, %d sites, %lld cells\n", nt, nsite,
              (long long)nc);
  return 0;
}
