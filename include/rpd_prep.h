/* Host-side preparation of restricted-power-diagram inputs (C ABI).
 *
 * These functions produce the arrays that the reference's driver builds on
 * the host before calling the GPU RPD:
 *   - per-vertex / per-edge / per-face adjacency counts and face ids, the
 *     replacement of load_tet_adj_info (src/IO/IO_CXX/io.cxx:238-335) with
 *     per-tet edge counts instead of the triangular n_vert^2 matrix;
 *   - the power-diagram 1-ring neighbour lists of the spheres, the
 *     replacement of the CGAL regular triangulation
 *     (src/rpd3d_base/triangulation.cxx:62-144, 324-549): exact power-cell
 *     adjacency inside a domain box, no bounding-box points;
 *   - deterministic synthetic tet meshes and sphere sets (counter-based
 *     SplitMix64 RNG, identical on every host).
 *
 * All functions return 0 on success and a negative code on error; they never
 * exit the process.
 */
#ifndef LIBMAT_AMD_RPD_PREP_H
#define LIBMAT_AMD_RPD_PREP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Counter-based RNG: uniform double in [0,1) for (seed, stream, index). */
double rpd_rng_uniform(uint64_t seed, uint64_t stream, uint64_t index);

/* Unit cube split into n^3 cubes, each cut into 6 positively oriented tets
 * (Kuhn split along the main diagonal).  Interior vertices are jittered by
 * N(0, (jitter_rel*h)^2) per axis (boundary vertices stay fixed); a draw that
 * makes any incident tet non-positive is redrawn.
 * verts: 3*(n+1)^3 doubles (AoS xyz); tets: 4*6*n^3 int32. */
int rpd_synth_grid_mesh(int32_t n, double jitter_rel, uint64_t seed,
                        double* verts, int32_t* tets);

/* Spheres (x, y, z, r) AoS, 4*n doubles.
 * rule 0: centres U[0,1]^3, r ~ U[rmin, rmax].
 * rule 1: centres U[0,1]^3, r = 0.9 * dist-to-cube-boundary * U[0.5, 1].
 * rule 2: 32-component Gaussian mixture (sigma ~ U[0.02, 0.1], centres
 *         U[0.1,0.9]^3, rejection to the cube), r ~ lognormal(-3.5, 0.6)
 *         clipped to the boundary distance. */
int rpd_synth_spheres(int32_t n, int32_t rule, double rmin, double rmax,
                      uint64_t seed, double* xyzr);

/* Tet adjacency (io.cxx:238-335 semantics).
 * v_adj[n_vert]: #tets at each vertex.
 * e_adj[6*n_tet]: #tets at each tet edge, local edge order
 *   (0,1),(0,2),(0,3),(1,2),(1,3),(2,3).
 * f_adj[4*n_tet]: 1 (boundary) or 2, face i opposite vertex i.
 * f_id[4*n_tet]: boundary faces get surface ids 0..n_sf-1 in first-encounter
 *   order over (tet, face); interior faces get ids >= n_sf in first-encounter
 *   order, shared by both tets.  *n_sf_facets receives n_sf.
 * Returns -2 on a non-manifold face (more than two tets). */
int rpd_tet_adjacency(int32_t n_vert, int32_t n_tet, const int32_t* tets,
                      int32_t* v_adj, int32_t* e_adj, int32_t* f_adj,
                      int32_t* f_id, int32_t* n_sf_facets);

/* Power-diagram neighbours of spheres (x, y, z, w = r^2) inside the box
 * [box[0],box[3]] x [box[1],box[4]] x [box[2],box[5]].
 * A sphere is valid iff its power cell meets the box; neighbour j of i is a
 * sphere whose bisector carries a facet of cell(i) inside the box.  Lists are
 * symmetrised (j in N(i) <=> i in N(j)) and sorted ascending.
 * The result is kept in an opaque handle. */
typedef struct rpd_nbr rpd_nbr;
int rpd_power_neighbors(int32_t n_sphere, const double* xyzw,
                        const double* box, int32_t n_threads, rpd_nbr** out);
/* valid[n_sphere]; offsets[n_sphere+1]; ids[offsets[n_sphere]] */
int64_t rpd_nbr_total(const rpd_nbr* h);
int32_t rpd_nbr_max_degree(const rpd_nbr* h);
int rpd_nbr_get(const rpd_nbr* h, int32_t* valid, int64_t* offsets,
                int32_t* ids);
void rpd_nbr_free(rpd_nbr* h);
/* A neighbour handle from per-sphere lists (the device builder's output,
 * rpd_power_neighbors_gpu in rpd_capi.h): valid[n], counts[n], lists[n *
 * stride]; the result is the symmetric closure over valid spheres, in the
 * layout of rpd_power_neighbors. */
int rpd_nbr_from_lists(int32_t n, const int32_t* valid, const int32_t* counts,
                       const int32_t* lists, int32_t stride, rpd_nbr** out);

#ifdef __cplusplus
}
#endif
#endif
