/* C ABI of the MI355X restricted-power-diagram (RPD) engine.
 *
 * Drop-in boundary for LibMAT's GPU RPD:
 *   compute_clipped_voro_diagram_GPU        src/rpd3d/voronoi.h:52-61,
 *                                            src/rpd3d/voronoi.cu:455-795
 * whose callers are RPD3D_GPU::calculate / calculate_partial
 * (src/rpd3d_api/rpd_api.cxx:117,275).  The C++ signature of that function
 * and the RPD3D_GPU class are re-exported by include/libmat_rpd.h on top of
 * this ABI; this header is the plain-pointer layer a cgo/JNI/ctypes binding
 * would use.
 *
 * Model: a context owns one device, one HIP stream and the device buffers of
 * one resident instance (tet mesh + spheres + neighbour lists).  rpd_upload
 * copies host arrays in (H2D); rpd_run executes the device pipeline
 * (candidate gathering, convex-cell clipping, exact / capacity re-runs,
 * geometry, compaction) on the resident instance; rpd_download_* copy
 * results out (D2H).  Errors are negative return codes with a message in
 * rpd_last_error(); nothing calls exit().  A context is not thread-safe.
 */
#ifndef LIBMAT_AMD_RPD_CAPI_H
#define LIBMAT_AMD_RPD_CAPI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Cell status codes: the reference's Status enum
 * (src/rpd3d_base/voronoi_common.h:10-25). */
enum rpd_status {
  RPD_TRIANGLE_OVERFLOW = 0,
  RPD_VERTEX_OVERFLOW = 1,
  RPD_INCONSISTENT_BOUNDARY = 2,
  RPD_SECURITY_RADIUS_NOT_REACHED = 3,
  RPD_SUCCESS = 4,
  RPD_NEEDS_EXACT_PREDICATES = 5,
  RPD_NO_INTERSECTION = 6,
  RPD_EDGE_OVERFLOW = 7,
  RPD_NEEDS_PERTURB = 8
};

/* Capacities of one cell, equal to the reference's PRESET 3
 * (voronoi_common.h:56-63): planes, vertices (dual triangles), edges. */
#define RPD_MAX_P 64
#define RPD_MAX_T 96
#define RPD_MAX_E 152

/* Host arrays of one RPD instance (all little-endian, row-major).
 * Tet face i is opposite local vertex i; local edges are ordered
 * (0,1),(0,2),(0,3),(1,2),(1,3),(2,3).  Tets must be positively oriented
 * (det(v1-v0, v2-v0, v3-v0) > 0).  Neighbour lists are the power-diagram
 * 1-rings in ascending site id (the reference's site_knn columns,
 * triangulation.cxx:324-352, as CSR). */
typedef struct {
  int32_t n_vert, n_tet, n_site;
  const double* verts;      /* 3*n_vert: x, y, z                        */
  const int32_t* tets;      /* 4*n_tet                                  */
  const int32_t* v_adj;     /* n_vert: #tets at each vertex             */
  const int32_t* e_adj;     /* 6*n_tet: #tets at each tet edge          */
  const int32_t* f_adj;     /* 4*n_tet: 1 (boundary) or 2               */
  const int32_t* f_id;      /* 4*n_tet: unique face ids                 */
  const double* sites;      /* 4*n_site: x, y, z, w = r^2               */
  const uint32_t* flags;    /* n_site: 1 = compute cells of this site   */
  const int64_t* knn_off;   /* n_site+1                                 */
  const int32_t* knn_ids;   /* knn_off[n_site]                          */
} rpd_instance_desc;

typedef struct {
  int32_t cells_per_wave;   /* cells clipped concurrently per wavefront:
                               1, 2, 4, 8 or 16; 0 = the rpd_tune choice
                               (with its small_caps) or 8              */
  int32_t use_active_mask;  /* 1: skip neighbour planes proved not to cut
                               the cell's tet (results are identical)  */
  int32_t compute_geometry; /* 1: volume, centroid, facet areas (FP64) */
  int32_t force_exact;      /* 1: decide every conflict sign with exact
                               arithmetic (test hook; identical output) */
  int32_t small_tier;       /* 1: small-capacity first pass with re-run
                               of overflowing cells at full capacity   */
  int32_t count_work;       /* 1: count the arithmetic the clip kernels
                               execute into rpd_stats.work (adds atomics:
                               for roofline figures, not for timing)    */
  int32_t small_caps;       /* capacities (vertices/planes/edges) of the
                               small tier: 0 = 32/16/48, 1 = 48/24/72,
                               2 = 16/12/32, 3 = 24/14/40 (2 and 3 with
                               cells_per_wave 4, 8 or 16)               */
  int32_t learned;          /* 1: per-tet sphere sets proposed by the
                               context's model (rpd_mlp_*), certified,
                               failing tets re-run on the full relation;
                               2: the same pipeline with the first
                               proposal_k proposals by power distance
                               instead of the model; the cells are those
                               of the plain path either way            */
  float logit_threshold;    /* learned = 1: a proposal is clipped if its
                               logit exceeds this (the owner always)    */
  int32_t proposal_k;       /* learned = 2: proposals taken per tet     */
} rpd_run_cfg;

/* One emitted (non-empty) cell.  Its payload words start at word_off:
 * nb_v vertex words (p0 | p1<<8 | p2<<16 | adj<<24: the three plane
 * indices of the dual triangle and #adjacent cells), then nb_p plane-source
 * words (int32: >= 0 neighbour site id of a bisector, -1-k for tet face k),
 * then nb_e edge words (p0 | p1<<8 | adj<<16).  Facet areas of its nb_p
 * planes start at area_off in the area stream (0 for planes carrying no
 * face). */
typedef struct {
  int32_t tet_id;
  int32_t seed;
  uint8_t nb_v, nb_p, nb_e, status;
  uint32_t reserved;
  uint64_t word_off;
  uint64_t area_off;
} rpd_cell;

typedef struct {
  int64_t n_cells;            /* emitted cells                           */
  int64_t n_candidates;       /* (tet, site) pairs passing the relation  */
  int64_t n_words;            /* payload words                           */
  int64_t n_areas;            /* facet-area entries                      */
  int64_t n_dropped[9];       /* candidates dropped, per rpd_status      */
  int64_t n_rerun_large;      /* cells over the small tier's capacities,
                                 re-run (small tiers of 16 or 24 vertices:
                                 at 32/16/48 first, then at full capacity
                                 if still over)                          */
  int64_t n_rerun_exact;      /* cells re-run with exact arithmetic      */
  int64_t n_active_planes;    /* neighbour planes not proved inactive
                                 (with count_work, else 0)               */
  int64_t n_all_planes;       /* all neighbour planes of all candidates
                                 (with count_work, else 0)               */
  int64_t n_bbox_fallback;    /* spheres whose candidate box fell back to
                                 the whole domain                        */
  int64_t n_bin_registrations; /* (sphere, bin) pairs of the candidate grid */
  int32_t tet_k;              /* max candidates of one tet               */
  int32_t site_k;             /* max neighbours of one site              */
  double ms_total;            /* device time of the last rpd_run         */
  double ms_phase[8];         /* candidates, clip, re-runs, compaction   */
  double ms_kernel[8];        /* site boxes, binning, relation count,
                                 relation fill, plane masks, clip (small
                                 tier), compaction                       */
  int64_t work[3][4];         /* with count_work, per clip tier (small,
                                 capacity re-runs, exact): FP64 conflict
                                 determinants evaluated, bisector planes
                                 built, vertices of the cells the tier
                                 finished with success, exact-arithmetic
                                 determinants; otherwise 0               */
  int64_t relation_work[4];   /* with count_work: (tet, site) pairs of the
                                 candidate bin lists tested, neighbour
                                 tests of the relation, 0, related pairs */
  int64_t learned[4];         /* learned runs: first-round candidates,
                                 tets re-run (the fallback: closure over
                                 missing facet neighbours or the full
                                 relation), candidates those re-runs
                                 added, tets re-run on the full relation */
  double ms_clip_small_sum;   /* summed device time of the run's small-tier
                                 clip launches (one per clip round: the
                                 plain path has one, the learned path one
                                 per round of its certificate)            */
  int32_t n_clip_small;       /* number of those launches                */
  int32_t n_host_reads;       /* host waits on the device during the run
                                 (count read-backs that size later work,
                                 and the final statistics)               */
  double ms_features_kernel;  /* learned runs: device time of the
                                 proposal/feature kernel alone            */
  int64_t scratch_bytes;      /* rpd_run_noalloc: scratch bytes the call
                                 used (high-water mark)                   */
} rpd_stats;

typedef struct rpd_ctx rpd_ctx;

int rpd_ctx_create(int32_t device, rpd_ctx** out);
/* The HIP device current on the calling thread (hipGetDevice), -1 if none:
 * the device the drop-in's compute_clipped_voro_diagram_GPU runs on. */
int32_t rpd_current_device(void);
void rpd_ctx_destroy(rpd_ctx* ctx);
const char* rpd_last_error(const rpd_ctx* ctx);
/* hipStream_t of the context (as void*), for event timing by callers. */
void* rpd_ctx_stream(rpd_ctx* ctx);

/* Copy an instance to the device (replaces the resident one); the host
 * staging of voronoi.cu:455-521 and copy_tet_data (voronoi.cu:324-362). */
int rpd_upload(rpd_ctx* ctx, const rpd_instance_desc* desc);

/* Spheres, flags and neighbour lists of one instance. */
typedef struct {
  int32_t n_site;
  const double* sites;      /* 4*n_site: x, y, z, w = r^2               */
  const uint32_t* flags;    /* n_site                                   */
  const int64_t* knn_off;   /* n_site+1                                 */
  const int32_t* knn_ids;   /* knn_off[n_site], instance-local ids      */
} rpd_sites_desc;

/* A batch of independent instances made resident together (replaces the
 * resident batch); one rpd_run then computes all of them at once.  No
 * reference counterpart: the reference computes one instance per call.
 * Ids in the results (rpd_cell.tet_id / seed, the relation CSR) are
 * batch-wide: instance b owns tets [tet_off[b], tet_off[b+1]) and sites
 * [site_off[b], site_off[b+1]) (rpd_batch_offsets). */
int rpd_upload_batch(rpd_ctx* ctx, int32_t n_inst,
                     const rpd_instance_desc* descs);
int rpd_update_sites_batch(rpd_ctx* ctx, int32_t n_inst,
                           const rpd_sites_desc* sites);
int rpd_batch_offsets(const rpd_ctx* ctx, int64_t* tet_off /* n_inst+1 */,
                      int64_t* site_off /* n_inst+1 */);
int32_t rpd_batch_size(const rpd_ctx* ctx);
/* Replace only the spheres / flags / neighbour lists (same n_site or not),
 * keeping the resident tet mesh: the per-iteration update of sphere
 * optimisation (load_partial_spheres_to_sites, rpd_api.cxx:343-379, and the
 * site H2D of voronoi.cu:469-521).  Single-instance batches only. */
int rpd_update_sites(rpd_ctx* ctx, int32_t n_site, const double* sites,
                     const uint32_t* flags, const int64_t* knn_off,
                     const int32_t* knn_ids);

/* Run the device pipeline on the context stream.  rpd_run returns when the
 * results are complete (it reads sizes back between phases); rpd_sync waits
 * for everything enqueued on the context stream. */
int rpd_run(rpd_ctx* ctx, const rpd_run_cfg* cfg);

/* Autotuner (no reference counterpart): runs the pipeline on the resident
 * batch for every clip shape -- 4, 8 or 16 cells per wavefront and the four
 * small-tier capacities of rpd_run_cfg.small_caps -- times the clip phases
 * (clip plus re-runs; best of two after a warm-up) and keeps the fastest
 * for runs with cells_per_wave = 0.  The other fields of base (or the
 * defaults) are used as given.  ms_out[3 * caps + shape] (12 entries)
 * receives the times. */
int rpd_tune(rpd_ctx* ctx, const rpd_run_cfg* base, int32_t* best_G,
             int32_t* best_caps, double* ms_out);
int rpd_sync(rpd_ctx* ctx);

int rpd_get_stats(const rpd_ctx* ctx, rpd_stats* out);

/* Regular-triangulation neighbours of n spheres (x, y, z, w = r^2) in the
 * box [x0, y0, z0, x1, y1, z1], computed on the device: replaces the
 * reference's CGAL regular triangulation and neighbour extraction
 * (triangulation.cxx:62-144, 396-438) as the host builder
 * rpd_power_neighbors (rpd_prep.h) does, with the same exact predicates and
 * the same result (valid flags, symmetric ascending CSR lists; read with
 * rpd_nbr_get, free with rpd_nbr_free).  ms_out (nullable) receives the
 * device time of the cell clipping.  Fails (non-zero, rpd_last_error)
 * when a cell needs more than 186 bisector planes at once, 256 vertices or
 * 160 neighbours; planes that no longer carry a vertex are dropped as the
 * cell is clipped, so only the final cell's size counts. */
typedef struct rpd_nbr rpd_nbr;
int rpd_power_neighbors_gpu(rpd_ctx* ctx, int32_t n, const double* xyzw,
                            const double* box, double* ms_out,
                            rpd_nbr** out);

/* Measurement aid: one plain device copy of `bytes` (multiple of 16) with a
 * streaming kernel, whose HBM traffic (bytes read, bytes written) is known;
 * used to calibrate hardware traffic counters. */
int rpd_calib_copy(rpd_ctx* ctx, int64_t bytes);

/* ---- learned proposals (no reference counterpart; the reference gathers
 * every related sphere).  Model: 48 -> 128 -> 128 -> 32, ReLU, BF16 MFMA
 * with FP32 accumulation; RPD_MLP_NPARAM fp32 parameters in the order
 * W1[48][128], b1[128], W2[128][128], b2[128], W3[128][32], b3[32]
 * (row-major [in][out]). */
#define RPD_MLP_IN 48
#define RPD_MLP_HID 128
#define RPD_MLP_OUT 32
#define RPD_MLP_NPARAM 26912

typedef struct {
  int32_t epochs;
  int32_t batch;      /* rows per step, multiple of 16 */
  float lr;           /* Adam step size (betas 0.9, 0.999, eps 1e-8) */
  float pos_weight;   /* weight of positive labels in the loss */
  uint64_t seed;      /* sample order */
} rpd_train_cfg;

/* Xavier-uniform initial weights (project RNG, seed), zero biases. */
int rpd_mlp_init(rpd_ctx* ctx, uint64_t seed);
int rpd_mlp_set(rpd_ctx* ctx, const float* params);
int rpd_mlp_get(rpd_ctx* ctx, float* params);
/* Proposals and features of every resident tet (device-resident for
 * training and scoring); optional copies: knn[n_tet*32] (site ids, -1 =
 * none; slot 0 the owner of the centroid), X[n_tet*48] (bf16 bits). */
int rpd_features(rpd_ctx* ctx, int32_t* knn, uint16_t* X);
/* Labels of the proposals from the last plain rpd_run of the resident
 * batch: Y = 1 if the sphere has a cell in the tet, M = 1 for valid
 * slots; optional copies Y, M [n_tet*32]. */
int rpd_labels(rpd_ctx* ctx, float* Y, float* M);
/* Trains the context's model on its resident features and labels;
 * loss[epochs] receives the mean loss per epoch. */
int rpd_mlp_train(rpd_ctx* ctx, const rpd_train_cfg* cfg, float* loss);
/* Test hook: loads params as the model, then the gradient of the loss
 * over n given rows (X bf16 bits [n*48], Y, M [n*32]) into grad, and the
 * loss. */
int rpd_mlp_grad(rpd_ctx* ctx, const float* params, int32_t n,
                 const uint16_t* X, const float* Y, const float* M,
                 float pos_weight, float* grad, float* loss);
/* Logits of the model for the resident features: logits[n_tet*32]. */
int rpd_mlp_scores(rpd_ctx* ctx, float* logits);

/* D2H copies of the last run's results; NULL pointers are skipped.
 * cells[n_cells], words[n_words], geom[4*n_cells] (volume, centroid x,y,z),
 * areas[n_areas]. */
int rpd_download_cells(rpd_ctx* ctx, rpd_cell* cells, uint32_t* words,
                       double* geom, double* areas);
/* Candidate relation of the last run: per-tet CSR of related site ids in
 * ascending order (the reference's tet_knn, voronoi.cu:284-317).
 * offsets[n_tet+1], ids[n_candidates]. */
int rpd_download_relation(rpd_ctx* ctx, int64_t* offsets, int32_t* ids);

/* Per-site centroids of the last run's cells: the Lloyd update of
 * compute_new_site (voronoi.cu:40-69) over the per-seed barycentre and
 * volume sums of the clip kernel (convex_cell.cu:1009-1054).  For each
 * (batch-wide) site s: out[4s..4s+2] = sum(vol * centroid) / sum(vol) over
 * its cells, out[4s+3] = sum(vol); a site without volume keeps its centre
 * (compute_new_site skips cell_vol == 0).  FP64 in a fixed summation
 * order, so the result does not depend on scheduling.  Needs a last run
 * with compute_geometry on the current sites.  out: host, 4*n_site
 * doubles (nullable); *dev_out (nullable) receives the device copy, owned
 * by the context and valid until the next call. */
int rpd_site_centroids(rpd_ctx* ctx, double* out, const double** dev_out);

/* The plain pipeline of the resident batch captured once into a hipGraph
 * and replayed: the batched, repeated RPD of sphere optimisation as one
 * graph launch per iteration (no reference counterpart: the reference
 * computes one instance per call and sizes every phase on the host).
 * rpd_graph_build runs cfg eagerly (its results are the current ones),
 * sizes every launch and buffer at the counts it saw times
 * (1 + headroom), and captures the stages with every count kept on the
 * device: the kernels read the counts the previous kernels left and stop
 * at the capacity.  rpd_graph_run replays the graph on the current data
 * (rpd_update_sites_batch may replace spheres and neighbour lists between
 * runs) and reads the counts back once; if one went past its capacity, or
 * the batch's per-instance tet or site counts, the tet subset switch or a
 * device buffer changed, it captures again through rpd_graph_build (whose
 * eager run gives the results) and sets *replayed = 0, else 1.  Results
 * are read as after rpd_run.  Plain and learned paths (cfg->learned 0, 1
 * or 2); count_work = 0, small_tier = 1.  A learned capture holds the
 * certificate, one closure round more than its eager run needed (a replay
 * needing more is captured again), the fallback and the output order, so
 * a replay's cells and per-tet site sets are those of an eager learned
 * run; its stats carry learned[0..3] and no re-run tier counts. */
int rpd_graph_build(rpd_ctx* ctx, const rpd_run_cfg* cfg, double headroom);
int rpd_graph_run(rpd_ctx* ctx, int32_t* replayed);
int rpd_graph_destroy(rpd_ctx* ctx);

/* Per-sphere power-cell adjacency of the last run's cells, computed on the
 * device: what get_all_voro_info and update_pc_cc_info /
 * update_pc_facet_cc_info / update_pc_edge_cc_info derive per sphere
 * (rpd_update.cxx:80-340, 409-521) for the topology checks of
 * fix_topo.cxx:87-270, plus the Euler sums of cal_cell_euler /
 * cal_halfplane_facet_euler (voronoi_defs.cxx:190-267).  Two cells of one
 * sphere are neighbours when both carry the tet face between their tets
 * as an active plane (cell_neighbors); a facet is (site, neighbour site),
 * the cells with an active bisector toward that neighbour
 * (facet_neigh_to_cells); an edge is (site, neighbour, neighbour), the
 * cells with an active edge between those two bisectors (e_to_cells).
 * Components are taken over cell neighbours (within the facet's or edge's
 * cells for those); a component is named by its smallest cell index.
 * Cell indices are positions in the last run's cell stream; site ids are
 * batch-wide.  Sizes go to *sizes (nullable); results stay on the device
 * until rpd_adjacency_download. */
typedef struct {
  int64_t n_cells;       /* cells of the last run                      */
  int64_t n_sites;       /* resident sites                             */
  int64_t n_cell_pairs;  /* neighbouring cell pairs (shared tet faces) */
  int64_t n_facets;      /* (site, neighbour) facets                   */
  int64_t n_facet_cells; /* (facet, cell) records                      */
  int64_t n_edges;       /* (site, neighbour pair) edges               */
  int64_t n_edge_cells;  /* (edge, cell) records                       */
} rpd_adj_sizes;
int rpd_adjacency(rpd_ctx* ctx, rpd_adj_sizes* sizes);

/* D2H of rpd_adjacency's results; NULL pointers are skipped.
 *   cell_euler[n_cells]   cal_cell_euler of each cell (binary32)
 *   cell_cc[n_cells]      component of each cell (its smallest cell)
 *   cell_nbr[4*n_cells]   neighbour cell across local tet face k, or -1
 *   site_cells, site_cc [n_sites]  number of cells, of components
 *   site_euler[n_sites]   sum of cell_euler (binary64, ascending cells)
 *                         minus the number of cells (is_to_fix_voro_euler)
 *   facets[5*n_facets]    site, neighbour, first record, #cells,
 *                         #components; sorted by (site, neighbour)
 *   facet_euler[n_facets] binary64 sum of the records' facet Euler values
 *                         in cell order (is_to_fix_facet_euler)
 *   facet_cells[2*n_facet_cells]  per record: cell, its component
 *   facet_cell_euler[n_facet_cells]  cal_halfplane_facet_euler (binary32)
 *   edges[6*n_edges]      site, neighbour min, neighbour max, first
 *                         record, #cells, #components; sorted
 *   edge_cells[2*n_edge_cells]    per record: cell, its component */
typedef struct {
  float* cell_euler;
  int32_t* cell_cc;
  int32_t* cell_nbr;
  int32_t* site_cells;
  int32_t* site_cc;
  double* site_euler;
  int32_t* facets;
  double* facet_euler;
  int32_t* facet_cells;
  float* facet_cell_euler;
  int32_t* edges;
  int32_t* edge_cells;
} rpd_adj_out;
int rpd_adjacency_download(rpd_ctx* ctx, const rpd_adj_out* out);

/* Restrict the following rpd_run calls to n tets of the resident mesh (ids
 * in [0, n_tet)); other tets take no candidates.  The partial RPD of
 * calculate_partial (rpd_api.cxx:249-262) clips only the tets of the
 * changed spheres' previous cells; a tet's cells depend only on the tet
 * and the spheres, so a subset run gives the cells of a full run on those
 * tets, with their original tet ids, without sending a partial mesh.  n = 0
 * clears the subset; rpd_upload / rpd_upload_batch clear it too.  Plain
 * runs only (a learned run fails while a subset is set). */
int rpd_set_tet_subset(rpd_ctx* ctx, int32_t n, const int32_t* tet_ids);

/* ---- allocation-free RPD.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  One call = rpd_upload +
 * the plain rpd_run + the cell stream copied into the caller's device
 * buffers, with every intermediate buffer carved from the library's static
 * device scratch of RPD_NOALLOC_SCRATCH_MIB MiB per GPU (held from the
 * library's load until the process ends); it calls no hipMalloc, hipFree or
 * allocating library routine.  Inputs are host arrays as for rpd_upload;
 * outputs are device buffers of the given capacities (entries), laid out
 * as rpd_download_cells writes them (geom: 4 doubles per cell).  Plain path
 * only (cfg->learned = 0, count_work = 0).  Errors: -4 when the instance
 * needs more scratch than the maximum, -5 when an output buffer is too
 * small (stats then hold nothing), with the sizes in rpd_last_error.  The
 * context's resident instance and last results are left as they were
 * (rpd_download_* still return the last rpd_run's); *st_out receives this
 * call's statistics.  Every call on one GPU shares that GPU's scratch and
 * leaves its contents behind: calls that overlap on one GPU (from several
 * contexts or threads) race. */
#define RPD_NOALLOC_SCRATCH_MIB 256
typedef struct {
  rpd_cell* cells;   /* device, cells_cap records                     */
  int64_t cells_cap;
  uint32_t* words;   /* device, words_cap payload words               */
  int64_t words_cap;
  double* geom;      /* device, 4 * cells_cap: volume, centroid       */
  double* areas;     /* device, areas_cap facet areas                 */
  int64_t areas_cap;
} rpd_noalloc_out;
int rpd_run_noalloc(rpd_ctx* ctx, const rpd_instance_desc* desc,
                    const rpd_run_cfg* cfg, const rpd_noalloc_out* out,
                    rpd_stats* st_out);

#ifdef __cplusplus
}
#endif
#endif
