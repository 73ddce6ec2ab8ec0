// C++ drop-in surface of LibMAT's restricted power diagram, on top of the
// C ABI of rpd_capi.h.
//
// Re-exported reference interfaces:
//   ConvexCellHost                 src/rpd3d_base/voronoi_defs.h:19-100
//   is_convex_cell_valid           src/rpd3d_base/voronoi_defs.cxx:293-298
//   compute_clipped_voro_diagram_GPU  src/rpd3d/voronoi.h:52-61
//   RPD3D_GPU                      src/rpd3d_api/rpd_api.h:8-69
//   update_power_cells (adjacency) src/rpd3d_base/rpd_update.cxx:568-640
// The cell record keeps the reference's field names, types, order,
// capacities and alignment (sizeof 3776, align 32 with GCC on x86-64), so
// consumers of std::vector<ConvexCellHost> (topology checks, feature
// fixes, exporters) compile and read the same bytes.
#pragma once
#include <stdint.h>

#include <array>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#ifndef LIBMAT_RPD_NO_REFERENCE_TYPES
typedef unsigned char uchar;
typedef unsigned int uint;
typedef double Scalar;
typedef std::array<int, 2> aint2;
typedef std::array<int, 3> aint3;
typedef std::array<int, 4> aint4;
typedef std::array<int, 5> aint5;
#endif
// a point with a surface facet id (LibMAT's v2int = pair<Vector3, int>)
typedef std::array<double, 3> rpd_vec3;
typedef std::pair<rpd_vec3, int> rpd_v2int;
// feature edge types of TetMesh::tet_es2fe_map (input_types.h:13-17)
constexpr int RPD_EDGE_SE = 1;  // convex sharp edge
constexpr int RPD_EDGE_CE = 2;  // concave edge

#define _MAX_P_ 64
#define _MAX_T_ 96
#define _MAX_E_ 152

constexpr uchar END_OF_LIST = 255;
constexpr int UNK_INT = -1;
constexpr float UNK_FLOAT = -1.f;
constexpr float F_TET_ADJ_DEFAULT = 2.f;
constexpr float F_CELL_ADJ_DEFAULT = 1.f;

// voronoi_common.h:8-25
enum SiteFlag { no_flag = 0, is_selected = 1 };
enum Status {
  early_return = -1,
  triangle_overflow = 0,
  vertex_overflow = 1,
  inconsistent_boundary = 2,
  security_radius_not_reached = 3,
  success = 4,
  needs_exact_predicates = 5,
  no_intersection = 6,
  edge_overflow = 7,
  needs_perturb = 8
};

// common_cxx.h:22-43 (host POD vector types and their alignments)
struct __attribute__((aligned(8))) cint2 {
  int x, y;
};
struct __attribute__((aligned(2))) cuchar2 {
  unsigned char x, y;
};
struct __attribute__((aligned(4))) cuchar3 {
  unsigned char x, y, z;
};
struct __attribute__((aligned(4))) cuchar4 {
  unsigned char x, y, z, w;
};
struct __attribute__((aligned(16))) cfloat3 {
  float x, y, z;
};
struct __attribute__((aligned(16))) cfloat4 {
  float x, y, z, w;
};
struct __attribute__((aligned(32))) cfloat5 {
  float x, y, z, w, h;
};

struct ConvexCellHost {
  ConvexCellHost() {}
  // float vertex of the dual triangle t, as the reference computes it
  // (voronoi_defs.cxx:33-74)
  cfloat4 compute_vertex_coordinates(cuchar3 t, bool persp_divide = true) const;
  void print_info() const;

  inline cuchar4& ver_trans(int t) { return ver_data_trans[t]; }
  inline cfloat5& clip_trans(int p) { return clip_data_trans[p]; }
  inline cuchar3& edge(int e) { return edge_data[e]; }
  inline uchar& ith_plane(uchar t, int i) {
    return reinterpret_cast<uchar*>(&(ver_trans(t)))[i];
  }
  inline cuchar4 ver_trans_const(int t) const { return ver_data_trans[t]; }
  inline cfloat4 clip_trans_const(int f) const {
    const cfloat5& c = clip_data_trans[f];
    return cfloat4{c.x, c.y, c.z, c.w};
  }
  inline cint2 clip_id2_const(int p) const { return clip_id2_data_trans[p]; }
  inline cuchar2 edge_id2_const(int e) const {
    return cuchar2{edge_data[e].x, edge_data[e].y};
  }
  inline uchar ith_plane_const(uchar t, int i) const {
    cuchar4 v = ver_trans_const(t);
    return reinterpret_cast<uchar*>(&(v))[i];
  }

  // topology (voronoi_defs.cxx:76-267)
  bool is_active_updated = false;
  std::vector<int> active_clipping_planes;
  std::vector<int> active_edges;
  Scalar cal_cell_euler();
  Scalar cal_halfplane_facet_euler(int neigh_id);
  // declared, never defined, in the reference too (voronoi_defs.h:53)
  Scalar cal_edge_euler(aint2 neigh_min_max);
  void print_cell_detail_euler() const;
  void reload_active();

  bool is_pc_explicit = false;
  std::vector<cfloat3> pc_points;
  std::vector<std::vector<int>> pc_local_active_faces;
  std::vector<int> pc_lf2active_map;
  void reload_pc_explicit();
  bool is_vertex_null = false;

  Status status;
  int thread_id;
  int voro_id;
  int tet_id;
  float weight = -1;
  bool is_active = false;
  uchar nb_v, nb_p, nb_e;
  cuchar4 ver_data_trans[_MAX_T_];
  cfloat5 clip_data_trans[_MAX_P_];
  cint2 clip_id2_data_trans[_MAX_P_];
  cuchar3 edge_data[_MAX_E_];
  float euler = -1;
  float cell_vol = -1;
  int id = UNK_INT;
};

bool is_convex_cell_valid(const ConvexCellHost& cell);

// The reference's GPU entry point, same signature and semantics
// (voronoi.cu:455-795): returns the non-empty cells of every flagged site in
// every tet, ordered by (tet, ascending site id), ids = index.
//   vertices: AoS xyz of the tet mesh, indices: 4 per tet (positively
//   oriented), v_adjs / e_adjs (triangular n_vert^2 layout of
//   get_edge_idx, convex_cell.h:46-60) / f_adjs / f_ids as
//   load_tet_adj_info (io.cxx:238-335); site: SoA [x.., y.., z..] (copied
//   back unchanged), site_weights = r^2, site_flags, site_knn
//   (site_k+1) x n_site column-major with -1 padding.
// Unused as in the reference: num_itr_global, v2tets, site_is_transposed
// (always SoA), nb_Lloyd_iter, preferred_tet_k.  site_cell_vol receives the
// per-site sum of cell volumes (the reference leaves it zero).
// Errors throw std::runtime_error (the reference exits the process).
std::vector<ConvexCellHost> compute_clipped_voro_diagram_GPU(
    const int num_itr_global, const std::vector<float>& vertices,
    const std::vector<int>& indices, const std::map<int, std::set<int>>& v2tets,
    const std::vector<int>& v_adjs, const std::vector<int>& e_adjs,
    const std::vector<int>& f_adjs, const std::vector<int>& f_ids,
    std::vector<float>& site, const int n_site,
    const std::vector<float>& site_weights, const std::vector<uint>& site_flags,
    const std::vector<int>& site_knn, const int site_k,
    std::vector<float>& site_cell_vol, const bool site_is_transposed,
    int nb_Lloyd_iter = 1, int preferred_tet_k = 0);

// Same, with per-tet edge counts (6 per tet, local edge order
// (0,1),(0,2),(0,3),(1,2),(1,3),(2,3)) instead of the n_vert^2 matrix, for
// meshes whose triangular matrix does not fit in memory.
std::vector<ConvexCellHost> compute_clipped_voro_diagram_GPU_tet_eadj(
    const std::vector<float>& vertices, const std::vector<int>& indices,
    const std::vector<int>& v_adjs, const std::vector<int>& tet_e_adjs,
    const std::vector<int>& f_adjs, const std::vector<int>& f_ids,
    std::vector<float>& site, const int n_site,
    const std::vector<float>& site_weights, const std::vector<uint>& site_flags,
    const std::vector<int>& site_knn, const int site_k,
    std::vector<float>& site_cell_vol);

// ------------------------------------------------------------ adjacency
// The adjacency part of PowerCell (src/matbase/medial_sphere.h:145-270)
// that update_power_cells derives from the cells of one sphere: what
// get_all_voro_info fills (rpd_update.cxx:80-340) -- cell ids and tets,
// cell <-> cell through shared tet faces, cell -> tet face ids, surface face
// centroids, halfplane facets and edges of the power cell, its vertices and
// their positions keyed by [cell id, local vertex], the end vertices of its
// edges, the covered sharp / concave feature edges and sharp-line end
// points -- and the connected components of update_pc_cc_info,
// update_pc_facet_cc_info and update_pc_edge_cc_info (rpd_update.cxx:
// 409-521).
struct PowerCellAdjacency {
  int voro_id = -1;
  std::set<int> cell_ids;
  std::set<int> tet_ids;
  std::map<int, std::set<int>> cell_neighbors;
  std::map<int, std::set<int>> cell_to_tfids;
  // cell id -> [(centroid of the cell's face on the surface, surface fid)]
  std::map<int, std::vector<rpd_v2int>> cell_to_surfv2fid;
  std::map<int, std::set<int>> facet_neigh_to_cells;
  std::map<aint2, std::set<int>> e_to_cells;
  // [neigh min, neigh max] -> [[cell, lvid], [cell, lvid]] per convex edge
  std::map<aint2, std::vector<std::array<aint2, 2>>> edge_2endvertices;
  std::map<aint2, aint3> vertex_2id;
  // [cell, lvid] -> (position, surface fid or -1)
  std::map<aint2, rpd_v2int> vertex_2pos;
  // [cell, lvid1, lvid2, fe_line_id, fe_id] of covered feature edges
  std::set<aint5> se_covered_lvids, ce_covered_lvids;
  // [cell, lvid, neigh id, se_line_id] -> position
  std::map<aint4, rpd_vec3> se_line_endpos;
  // connected components (through cell_neighbors), each as a set of cells
  std::vector<std::set<int>> cc_cells;
  std::vector<std::vector<rpd_v2int>> cc_surf_v2fids;
  std::map<int, std::vector<std::set<int>>> facet_cc_cells;
  std::map<int, std::vector<std::vector<rpd_v2int>>> facet_cc_surf_v2fids;
  std::map<aint2, std::vector<std::set<int>>> edge_cc_cells;
};

// Builds the adjacency of every sphere id that owns cells (ConvexCellHost::
// voro_id), keyed by that id; reloads the cells' active elements and
// explicit faces.  max_surf_fid: largest surface facet id (tet faces with
// f_id <= it lie on the surface).  tet_es2fe_map: TetMesh::tet_es2fe_map,
// <tet, local fid min, local fid max> -> <fe_type, fe_id, fe_line_id>
// (nullptr: no feature edges).  Multithreaded over spheres.
std::map<int, PowerCellAdjacency> update_power_cells_adjacency(
    std::vector<ConvexCellHost>& cells, int max_surf_fid,
    const std::map<aint3, aint3>* tet_es2fe_map = nullptr, int n_threads = 0);

// ------------------------------------------------------------ export
// The faces of one cell appended to a polygon soup, the reference's
// get_one_convex_cell_faces_const (io_cuda.cxx:21-141): the cell's vertices
// as binary32 points, each active plane's vertex cycle as a polygon (or
// its triangle fan with is_triangle), the cell's sphere and the sphere
// across each face (-1 for tet faces); with is_boundary_only only bisector
// faces and tet faces with id < max_sf_fid.  Needs reload_active() first.
void get_one_convex_cell_faces_const(
    const ConvexCellHost& cc_trans, std::vector<cfloat3>& voro_points,
    std::vector<std::vector<unsigned>>& one_voro_cell_faces,
    std::vector<int>& voro_faces_sites, std::vector<int>& voro_faces_adj_sites,
    bool is_triangle, int max_sf_fid, bool is_boundary_only);

namespace libmat_rpd {
// mean of a cell's binary32 vertices (compute_cell_barycenter,
// io_cuda.cxx:6-17)
std::array<float, 3> cell_barycenter(const ConvexCellHost& cell);
// The cells' faces written as a version-5 binary Houdini geometry at
// `path` (save_convex_cells_houdini, io_cuda.cxx:143-186, with the file
// name given instead of derived from the run name and a time stamp):
// points with P, polygons with the int primitive attribute "PrimAttr" =
// the cell's sphere.  skip_cell(barycenter) == true leaves a cell out (the
// reference's slice plane).  Reloads the cells' active elements.
void save_convex_cells_bgeo(
    const std::string& path, std::vector<ConvexCellHost>& cells,
    int max_sf_fid, bool is_boundary_only,
    const std::function<bool(const std::array<float, 3>&)>& skip_cell = {});
}  // namespace libmat_rpd

// Connected components of the elements of `to_visit` under `neighbors`
// (edges to elements outside the set are ignored), in order of their least
// element (common_cxx.h get_CC_given_neighbors); returns their number.
int rpd_connected_components(const std::set<int>& to_visit,
                             const std::map<int, std::set<int>>& neighbors,
                             std::vector<std::set<int>>& out);

// ------------------------------------------------------------ RPD3D_GPU
// Minimal stand-ins for the fields of TetMesh (input_types.h:122-147) and
// MedialSphere (medial_sphere.h:295-...) that the RPD driver reads.  The
// driver below is a template so the reference's own types can be used.
struct RpdTetMesh {
  std::vector<float> tet_vertices;  // AoS xyz
  std::vector<int> tet_indices;     // 4 per tet
  std::map<int, std::set<int>> v2tets;
  std::vector<int> v_adjs, e_adjs, f_adjs, f_ids;
  std::map<aint3, aint3> tet_es2fe_map;  // feature edges (input_types.h:147)
};
// the PowerCell fields (medial_sphere.h:145-260) its caller fills
struct RpdPowerCell : PowerCellAdjacency {};
struct RpdSphere {
  int id = -1;
  int site_id = -1;
  int itr_cnt = -1;  // iteration that created or last moved the sphere
  bool is_deleted = false;
  bool is_rt_valid = false;
  bool is_rt_prev_valid = false;
  int rt_change_status = 0;
  std::set<int> rt_neigh_ids, rt_neigh_ids_prev;
  std::array<double, 3> center{{0, 0, 0}};
  double radius = 0;
  RpdPowerCell pcell;
};

struct rpd_ctx;  // the device context of rpd_capi.h

namespace libmat_rpd {
// power-diagram 1-rings of the live spheres (is_deleted == false) inside
// the box lo/hi; valid[i] = power cell meets the box (the RT validity of
// generate_RT_CGAL_and_mark_valid_spheres, triangulation.cxx:62-144)
void power_neighbors(const std::vector<std::array<double, 4>>& xyzr,
                     const std::vector<char>& live, const double box[6],
                     std::vector<char>& valid,
                     std::vector<std::set<int>>& neighbors);

// A tet mesh kept on the GPU between RPD calls: uploaded once, after which
// each run sends only the spheres and, for a partial RPD, the ids of the
// tets to clip (calculate_partial's partial tets, rpd_api.cxx:249-262,
// clipped in place on the full mesh instead of sent as a partial mesh).
// Owns its device context (on `device`, or the calling thread's current
// HIP device when negative); one object is not for concurrent use.
class ResidentMesh {
 public:
  explicit ResidentMesh(int device = -1);
  ~ResidentMesh();
  ResidentMesh(const ResidentMesh&) = delete;
  ResidentMesh& operator=(const ResidentMesh&) = delete;
  // the mesh arrays of compute_clipped_voro_diagram_GPU (e_adjs in the
  // triangular layout of get_edge_idx); replaces a resident mesh
  void upload(const std::vector<float>& vertices,
              const std::vector<int>& indices, const std::vector<int>& v_adjs,
              const std::vector<int>& e_adjs, const std::vector<int>& f_adjs,
              const std::vector<int>& f_ids);
  bool uploaded() const { return n_tet_ > 0; }
  // compute_clipped_voro_diagram_GPU on the resident mesh for the given
  // spheres, over the tets in `tets` (nullptr: all).  Cells keep the
  // original tet ids; ids = position; thread_id = (position of the tet in
  // `tets`) * tet_k + slot, as a run over the partial mesh of those tets
  // numbers them.
  std::vector<ConvexCellHost> run(const std::vector<float>& site, int n_site,
                                  const std::vector<float>& site_weights,
                                  const std::vector<uint>& site_flags,
                                  const std::vector<int>& site_knn, int site_k,
                                  std::vector<float>& site_cell_vol,
                                  const std::vector<int>* tets = nullptr);

 private:
  int device_;
  ::rpd_ctx* ctx_ = nullptr;
  int n_tet_ = 0;
  std::vector<float> vertices_;
  std::vector<int> indices_, f_adjs_, f_ids_;
};
}  // namespace libmat_rpd

template <class TetMeshT = RpdTetMesh, class SurfaceMeshT = void,
          class ParameterT = void, class SphereT = RpdSphere>
class RPD3D_GPU_T {
 public:
  RPD3D_GPU_T() {}
  // The reference deletes the borrowed pointers in its destructor
  // (rpd_api.cxx:8-13); here they stay owned by the caller.
  ~RPD3D_GPU_T() {}
  void init(const TetMeshT* _tet_mesh, const SurfaceMeshT* _sf_mesh,
            const ParameterT* _params) {
    tet_mesh = _tet_mesh;
    sf_mesh = _sf_mesh;
    params = _params;
    num_itr_rpd = 0;
    n_site = -1;
    site_k = -1;
    gpu_.reset();  // the new mesh is uploaded by the next RPD
  }
  void set_spheres(std::vector<SphereT>* _all_medial_spheres) {
    all_medial_spheres = _all_medial_spheres;
  }

  // Full RPD of all valid spheres (rpd_api.cxx:32-145).
  void calculate();
  // Incremental RPD of the sphere-optimisation loop (rpd_api.cxx:147-313):
  // re-computes the cells of the spheres changed in iteration
  // num_itr_global (created/moved there, or whose validity changed) and
  // of their 1-rings, over the tets of those spheres' previous cells (all
  // tets if is_given_all_tets), merges them with the other spheres'
  // previous cells and refreshes every sphere's pcell.  Increments
  // num_itr_global and resets num_sphere_added.  No-op when
  // num_sphere_added == 0.
  void calculate_partial(int& num_itr_global, int& num_sphere_added,
                         const bool is_given_all_tets = false);
  // pcell of every sphere from powercells (update_power_cells,
  // rpd_update.cxx:568-640, adjacency part)
  void update_spheres_power_cells();

  // steps shared by calculate and calculate_partial
  void mark_valid_spheres(std::set<int>& valid_sphere_ids);
  int partial_sites_and_neighbors(const std::set<int>& given,
                                  std::vector<int>& map_site2msphere,
                                  std::set<int>& spheres_and_1rings);
  void load_changed_spheres(int num_itr_global, std::set<int>& changed,
                            std::set<int>& invalid);
  void load_partial_tet_given_spheres(const std::set<int>& given_spheres,
                                      std::vector<int>& partial_tet_indices,
                                      std::vector<int>& map_tet_new2orig,
                                      std::vector<int>& partial_tet_fids,
                                      std::vector<int>& partial_tet_f_adjs);

  // site arrays of (sphere + 1ring + 2ring), flags of (sphere + 1ring)
  // (rpd_api.cxx:343-379)
  void load_partial_spheres_to_sites(const std::vector<int>& map_site2msphere,
                                     const std::set<int>& spheres_and_1rings);
  // site ids -> sphere ids, partial tet ids -> original (rpd_api.cxx:386-430)
  void update_convex_cells_voro_and_tet_ids(
      const std::vector<int>& map_site2msphere,
      const std::vector<int>& map_tet_new2orig,
      std::vector<ConvexCellHost>& cells_partials, bool is_debug);
  // new cells of (sphere + 1ring) + previous cells of the others that are
  // still valid, ids = position (rpd_api.cxx:432-479)
  std::vector<ConvexCellHost> merge_convex_cells(
      const std::set<int>& valid_sphere_ids,
      const std::set<int>& spheres_and_1rings,
      const std::vector<ConvexCellHost>& convex_cells_prev,
      const std::vector<ConvexCellHost>& convex_cells_new, bool is_debug);

 public:
  std::vector<ConvexCellHost> powercells;
  bool is_debug = false;
  const TetMeshT* tet_mesh = nullptr;
  const SurfaceMeshT* sf_mesh = nullptr;
  const ParameterT* params = nullptr;
  std::vector<SphereT>* all_medial_spheres = nullptr;
  bool site_is_transposed = true;
  std::vector<float> site;
  std::vector<float> site_weights;
  std::vector<uint> site_flags;
  std::vector<int> site_knn;
  std::vector<float> site_cell_vol;
  int num_itr_rpd = 0;
  int n_site = -1;
  int site_k = -1;
  int device = -1;  // GPU of the resident mesh (-1: the current device)

  int max_surf_fid = -1;  // largest surface facet id (-1: from f_adjs)

 private:
  // the tet mesh on the GPU, uploaded at the first RPD after init()
  std::shared_ptr<libmat_rpd::ResidentMesh> gpu_;
  libmat_rpd::ResidentMesh& gpu() {
    if (!gpu_) {
      auto g = std::make_shared<libmat_rpd::ResidentMesh>(device);
      g->upload(tet_mesh->tet_vertices, tet_mesh->tet_indices,
                tet_mesh->v_adjs, tet_mesh->e_adjs, tet_mesh->f_adjs,
                tet_mesh->f_ids);
      gpu_ = g;
    }
    return *gpu_;
  }
};

using RPD3D_GPU = RPD3D_GPU_T<>;

#include "libmat_rpd_impl.h"
