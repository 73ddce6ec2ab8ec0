"""ctypes layer over the C-ABI of include/rpd_capi.h.

:class:`Context` mirrors the reference's one-shot entry point
``compute_clipped_voro_diagram_GPU`` (src/rpd3d/voronoi.h:52-61) as an
upload / run / download sequence on a resident instance.  There is no CPU
fallback: constructing a Context without the native library or without a
GPU raises.
"""
import ctypes

import numpy as np

from ._native import lib

CELL_DTYPE = np.dtype([
    ("tet", np.int32), ("seed", np.int32), ("nb_v", np.uint8),
    ("nb_p", np.uint8), ("nb_e", np.uint8), ("status", np.uint8),
    ("reserved", np.uint32), ("word_off", np.uint64), ("area_off", np.uint64),
])
assert CELL_DTYPE.itemsize == 32

SUCCESS = 4
STATUS_NAMES = ["triangle_overflow", "vertex_overflow",
                "inconsistent_boundary", "security_radius_not_reached",
                "success", "needs_exact_predicates", "no_intersection",
                "edge_overflow", "needs_perturb"]


class InstanceDesc(ctypes.Structure):
    _fields_ = [("n_vert", ctypes.c_int32), ("n_tet", ctypes.c_int32),
                ("n_site", ctypes.c_int32)] + [
        (n, ctypes.c_void_p) for n in
        ("verts", "tets", "v_adj", "e_adj", "f_adj", "f_id", "sites",
         "flags", "knn_off", "knn_ids")]


class SitesDesc(ctypes.Structure):
    _fields_ = [("n_site", ctypes.c_int32), ("sites", ctypes.c_void_p),
                ("flags", ctypes.c_void_p), ("knn_off", ctypes.c_void_p),
                ("knn_ids", ctypes.c_void_p)]


class TrainCfg(ctypes.Structure):
    _fields_ = [("epochs", ctypes.c_int32), ("batch", ctypes.c_int32),
                ("lr", ctypes.c_float), ("pos_weight", ctypes.c_float),
                ("seed", ctypes.c_uint64)]


MLP_NPARAM = 26912


class RunCfg(ctypes.Structure):
    _fields_ = [("cells_per_wave", ctypes.c_int32),
                ("use_active_mask", ctypes.c_int32),
                ("compute_geometry", ctypes.c_int32),
                ("force_exact", ctypes.c_int32),
                ("small_tier", ctypes.c_int32),
                ("count_work", ctypes.c_int32),
                ("small_caps", ctypes.c_int32),
                ("learned", ctypes.c_int32),
                ("logit_threshold", ctypes.c_float),
                ("proposal_k", ctypes.c_int32)]


class NoAllocOut(ctypes.Structure):
    _fields_ = [("cells", ctypes.c_void_p), ("cells_cap", ctypes.c_int64),
                ("words", ctypes.c_void_p), ("words_cap", ctypes.c_int64),
                ("geom", ctypes.c_void_p), ("areas", ctypes.c_void_p),
                ("areas_cap", ctypes.c_int64)]


class Stats(ctypes.Structure):
    _fields_ = [("n_cells", ctypes.c_int64), ("n_candidates", ctypes.c_int64),
                ("n_words", ctypes.c_int64), ("n_areas", ctypes.c_int64),
                ("n_dropped", ctypes.c_int64 * 9),
                ("n_rerun_large", ctypes.c_int64),
                ("n_rerun_exact", ctypes.c_int64),
                ("n_active_planes", ctypes.c_int64),
                ("n_all_planes", ctypes.c_int64),
                ("n_bbox_fallback", ctypes.c_int64),
                ("n_bin_registrations", ctypes.c_int64),
                ("tet_k", ctypes.c_int32), ("site_k", ctypes.c_int32),
                ("ms_total", ctypes.c_double),
                ("ms_phase", ctypes.c_double * 8),
                ("ms_kernel", ctypes.c_double * 8),
                ("work", (ctypes.c_int64 * 4) * 3),
                ("relation_work", ctypes.c_int64 * 4),
                ("learned", ctypes.c_int64 * 4),
                ("ms_clip_small_sum", ctypes.c_double),
                ("n_clip_small", ctypes.c_int32),
                ("n_host_reads", ctypes.c_int32),
                ("ms_features_kernel", ctypes.c_double),
                ("scratch_bytes", ctypes.c_int64)]

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if hasattr(v, "__len__"):
                v = [list(x) if hasattr(x, "__len__") else x for x in v]
            d[name] = v
        return d


def _declare_capi(L):
    vp, i32 = ctypes.c_void_p, ctypes.c_int32
    L.rpd_ctx_create.argtypes = [i32, ctypes.POINTER(vp)]
    L.rpd_ctx_create.restype = ctypes.c_int
    L.rpd_ctx_destroy.argtypes = [vp]
    L.rpd_ctx_destroy.restype = None
    L.rpd_last_error.argtypes = [vp]
    L.rpd_last_error.restype = ctypes.c_char_p
    L.rpd_ctx_stream.argtypes = [vp]
    L.rpd_ctx_stream.restype = vp
    L.rpd_upload.argtypes = [vp, ctypes.POINTER(InstanceDesc)]
    L.rpd_upload.restype = ctypes.c_int
    L.rpd_update_sites.argtypes = [vp, i32, vp, vp, vp, vp]
    L.rpd_update_sites.restype = ctypes.c_int
    L.rpd_upload_batch.argtypes = [vp, i32, ctypes.POINTER(InstanceDesc)]
    L.rpd_upload_batch.restype = ctypes.c_int
    L.rpd_update_sites_batch.argtypes = [vp, i32, ctypes.POINTER(SitesDesc)]
    L.rpd_update_sites_batch.restype = ctypes.c_int
    L.rpd_batch_offsets.argtypes = [vp, vp, vp]
    L.rpd_batch_offsets.restype = ctypes.c_int
    L.rpd_batch_size.argtypes = [vp]
    L.rpd_batch_size.restype = i32
    L.rpd_tune.argtypes = [vp, vp, vp, vp, vp]
    L.rpd_tune.restype = ctypes.c_int
    for n, a in (("rpd_mlp_init", [vp, ctypes.c_uint64]),
                 ("rpd_mlp_set", [vp, vp]), ("rpd_mlp_get", [vp, vp]),
                 ("rpd_features", [vp, vp, vp]),
                 ("rpd_labels", [vp, vp, vp]),
                 ("rpd_mlp_train", [vp, ctypes.POINTER(TrainCfg), vp]),
                 ("rpd_mlp_grad", [vp, vp, i32, vp, vp, vp, ctypes.c_float,
                                   vp, vp]),
                 ("rpd_mlp_scores", [vp, vp])):
        getattr(L, n).argtypes = a
        getattr(L, n).restype = ctypes.c_int
    L.rpd_run.argtypes = [vp, ctypes.POINTER(RunCfg)]
    L.rpd_run.restype = ctypes.c_int
    L.rpd_sync.argtypes = [vp]
    L.rpd_sync.restype = ctypes.c_int
    L.rpd_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    L.rpd_get_stats.restype = ctypes.c_int
    L.rpd_download_cells.argtypes = [vp, vp, vp, vp, vp]
    L.rpd_download_cells.restype = ctypes.c_int
    L.rpd_download_relation.argtypes = [vp, vp, vp]
    L.rpd_download_relation.restype = ctypes.c_int
    L.rpd_calib_copy.argtypes = [vp, ctypes.c_int64]
    L.rpd_calib_copy.restype = ctypes.c_int
    L.rpd_power_neighbors_gpu.argtypes = [vp, ctypes.c_int32, vp, vp, vp,
                                          ctypes.POINTER(vp)]
    L.rpd_power_neighbors_gpu.restype = ctypes.c_int
    L.rpd_run_noalloc.argtypes = [vp, ctypes.POINTER(InstanceDesc),
                                  ctypes.POINTER(RunCfg),
                                  ctypes.POINTER(NoAllocOut),
                                  ctypes.POINTER(Stats)]
    L.rpd_run_noalloc.restype = ctypes.c_int
    L.rpd_site_centroids.argtypes = [vp, vp, ctypes.POINTER(vp)]
    L.rpd_site_centroids.restype = ctypes.c_int
    L.rpd_set_tet_subset.argtypes = [vp, i32, vp]
    L.rpd_set_tet_subset.restype = ctypes.c_int
    L.rpd_graph_build.argtypes = [vp, ctypes.POINTER(RunCfg), ctypes.c_double]
    L.rpd_graph_build.restype = ctypes.c_int
    L.rpd_graph_run.argtypes = [vp, vp]
    L.rpd_graph_run.restype = ctypes.c_int
    L.rpd_graph_destroy.argtypes = [vp]
    L.rpd_graph_destroy.restype = ctypes.c_int
    L.rpd_adjacency.argtypes = [vp, ctypes.POINTER(AdjSizes)]
    L.rpd_adjacency.restype = ctypes.c_int
    L.rpd_adjacency_download.argtypes = [vp, ctypes.POINTER(AdjOut)]
    L.rpd_adjacency_download.restype = ctypes.c_int


class AdjSizes(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in
                ("n_cells", "n_sites", "n_cell_pairs", "n_facets",
                 "n_facet_cells", "n_edges", "n_edge_cells")]


class AdjOut(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("cell_euler", "cell_cc", "cell_nbr", "site_cells", "site_cc",
                 "site_euler", "facets", "facet_euler", "facet_cells",
                 "facet_cell_euler", "edges", "edge_cells")]


CAPI_SYMBOLS = ("rpd_ctx_create", "rpd_ctx_destroy", "rpd_last_error",
                "rpd_ctx_stream", "rpd_upload", "rpd_update_sites",
                "rpd_upload_batch", "rpd_update_sites_batch",
                "rpd_batch_offsets", "rpd_batch_size", "rpd_tune", "rpd_run",
                "rpd_mlp_init", "rpd_mlp_set", "rpd_mlp_get", "rpd_features",
                "rpd_labels", "rpd_mlp_train", "rpd_mlp_grad",
                "rpd_mlp_scores",
                "rpd_sync", "rpd_get_stats", "rpd_download_cells",
                "rpd_download_relation", "rpd_calib_copy",
                "rpd_power_neighbors_gpu", "rpd_run_noalloc",
                "rpd_site_centroids", "rpd_set_tet_subset", "rpd_adjacency",
                "rpd_adjacency_download", "rpd_graph_build", "rpd_graph_run",
                "rpd_graph_destroy", "rpd_current_device")


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class RPDResult:
    """Emitted cells of one run, in (tet, ascending site) order."""

    def __init__(self, cells, words, geom, areas):
        self.cells = cells
        self.words = words
        self.geom = geom.reshape(-1, 4) if geom is not None else None
        self.areas = areas

    def __len__(self):
        return self.cells.size

    def cell(self, q):
        c = self.cells[q]
        o = int(c["word_off"])
        nv, np_, ne = int(c["nb_v"]), int(c["nb_p"]), int(c["nb_e"])
        ver = self.words[o:o + nv]
        src = self.words[o + nv:o + nv + np_].view(np.int32)
        edge = self.words[o + nv + np_:o + nv + np_ + ne]
        a = int(c["area_off"])
        area = self.areas[a:a + np_] if self.areas is not None else None
        return ver, src, edge, area


class Context:
    def __init__(self, device=0):
        L = lib()
        self._L = L
        h = ctypes.c_void_p()
        rc = L.rpd_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"rpd_ctx_create failed ({rc}): no usable GPU")
        self._h = h
        self._keep = []

    def close(self):
        if getattr(self, "_h", None):
            self._L.rpd_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = self._L.rpd_last_error(self._h).decode()
            raise RuntimeError(f"{what} failed ({rc}): {msg}")

    @property
    def stream(self):
        return self._L.rpd_ctx_stream(self._h)

    @staticmethod
    def _arrays(inst):
        return dict(
            verts=np.ascontiguousarray(inst.verts, np.float64),
            tets=np.ascontiguousarray(inst.tets, np.int32),
            v_adj=np.ascontiguousarray(inst.v_adj, np.int32),
            e_adj=np.ascontiguousarray(inst.e_adj, np.int32),
            f_adj=np.ascontiguousarray(inst.f_adj, np.int32),
            f_id=np.ascontiguousarray(inst.f_id, np.int32),
            sites=np.ascontiguousarray(inst.sites, np.float64),
            flags=np.ascontiguousarray(inst.flags, np.uint32),
            knn_off=np.ascontiguousarray(inst.knn_off, np.int64),
            knn_ids=np.ascontiguousarray(inst.knn_ids, np.int32))

    @staticmethod
    def _desc(inst, a):
        return InstanceDesc(inst.n_vert, inst.n_tet, inst.n_site,
                            *[_p(a[n]) for n in
                              ("verts", "tets", "v_adj", "e_adj", "f_adj",
                               "f_id", "sites", "flags", "knn_off",
                               "knn_ids")])

    def upload(self, inst):
        a = self._arrays(inst)
        d = self._desc(inst, a)
        self._check(self._L.rpd_upload(self._h, ctypes.byref(d)),
                    "rpd_upload")
        self.n_tet = inst.n_tet

    def upload_batch(self, insts):
        """Make a list of instances resident together; returns the
        batch-wide tet and site offsets (n_inst + 1 each)."""
        arrs = [self._arrays(i) for i in insts]
        ds = (InstanceDesc * len(insts))(
            *[self._desc(i, a) for i, a in zip(insts, arrs)])
        self._check(self._L.rpd_upload_batch(self._h, len(insts), ds),
                    "rpd_upload_batch")
        self.n_tet = sum(i.n_tet for i in insts)
        return self.batch_offsets()

    def batch_offsets(self):
        n = self._L.rpd_batch_size(self._h)
        to = np.empty(n + 1, np.int64)
        so = np.empty(n + 1, np.int64)
        self._check(self._L.rpd_batch_offsets(self._h, _p(to), _p(so)),
                    "rpd_batch_offsets")
        return to, so

    def update_sites(self, sites, flags, knn_off, knn_ids):
        sites = np.ascontiguousarray(sites, np.float64)
        flags = np.ascontiguousarray(flags, np.uint32)
        knn_off = np.ascontiguousarray(knn_off, np.int64)
        knn_ids = np.ascontiguousarray(knn_ids, np.int32)
        self._check(self._L.rpd_update_sites(
            self._h, sites.shape[0], _p(sites), _p(flags), _p(knn_off),
            _p(knn_ids)), "rpd_update_sites")

    def update_sites_batch(self, site_sets):
        """site_sets: one (sites, flags, knn_off, knn_ids) per instance."""
        keep = [tuple(np.ascontiguousarray(x, t) for x, t in
                      zip(ss, (np.float64, np.uint32, np.int64, np.int32)))
                for ss in site_sets]
        ds = (SitesDesc * len(keep))(
            *[SitesDesc(k[0].shape[0], _p(k[0]), _p(k[1]), _p(k[2]),
                        _p(k[3])) for k in keep])
        self._check(self._L.rpd_update_sites_batch(self._h, len(keep), ds),
                    "rpd_update_sites_batch")

    def tune(self):
        """Pick the fastest clip shape for the resident batch (used by
        runs with cells_per_wave=0); returns (G, caps, {(caps, G): ms})."""
        g, c = ctypes.c_int32(), ctypes.c_int32()
        ms = np.zeros(12)
        self._check(self._L.rpd_tune(self._h, None, ctypes.byref(g),
                                     ctypes.byref(c), _p(ms)), "rpd_tune")
        return g.value, c.value, {(cc, G): ms[3 * cc + q]
                                  for cc in range(4)
                                  for q, G in enumerate((4, 8, 16))}

    # ---- learned proposals
    def mlp_init(self, seed=0):
        self._check(self._L.rpd_mlp_init(self._h, seed), "rpd_mlp_init")

    def mlp_set(self, params):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == MLP_NPARAM
        self._check(self._L.rpd_mlp_set(self._h, _p(p)), "rpd_mlp_set")

    def mlp_get(self):
        p = np.empty(MLP_NPARAM, np.float32)
        self._check(self._L.rpd_mlp_get(self._h, _p(p)), "rpd_mlp_get")
        return p

    def features(self, download=True):
        """Proposals (n_tet, 32) and bf16 feature bits (n_tet, 48)."""
        if not download:
            self._check(self._L.rpd_features(self._h, None, None),
                        "rpd_features")
            return None
        knn = np.empty((self.n_tet, 32), np.int32)
        X = np.empty((self.n_tet, 48), np.uint16)
        self._check(self._L.rpd_features(self._h, _p(knn), _p(X)),
                    "rpd_features")
        return knn, X

    def labels(self, download=True):
        if not download:
            self._check(self._L.rpd_labels(self._h, None, None), "rpd_labels")
            return None
        Y = np.empty((self.n_tet, 32), np.float32)
        M = np.empty((self.n_tet, 32), np.float32)
        self._check(self._L.rpd_labels(self._h, _p(Y), _p(M)), "rpd_labels")
        return Y, M

    def train(self, epochs=4, batch=1024, lr=2e-3, pos_weight=4.0, seed=0):
        loss = np.empty(epochs, np.float32)
        cfg = TrainCfg(epochs, batch, lr, pos_weight, seed)
        self._check(self._L.rpd_mlp_train(self._h, ctypes.byref(cfg),
                                          _p(loss)), "rpd_mlp_train")
        return loss

    def mlp_grad(self, params, X, Y, M, pos_weight):
        params = np.ascontiguousarray(params, np.float32)
        X = np.ascontiguousarray(X, np.uint16)
        Y = np.ascontiguousarray(Y, np.float32)
        M = np.ascontiguousarray(M, np.float32)
        g = np.empty(MLP_NPARAM, np.float32)
        loss = np.zeros(1, np.float32)
        self._check(self._L.rpd_mlp_grad(
            self._h, _p(params), X.shape[0], _p(X), _p(Y), _p(M),
            float(pos_weight), _p(g), _p(loss)), "rpd_mlp_grad")
        return g, float(loss[0])

    def scores(self):
        z = np.empty((self.n_tet, 32), np.float32)
        self._check(self._L.rpd_mlp_scores(self._h, _p(z)), "rpd_mlp_scores")
        return z

    def run(self, cells_per_wave=8, use_active_mask=True, geometry=True,
            force_exact=False, small_tier=True, count_work=False,
            small_caps=0, learned=0, logit_threshold=0.0, proposal_k=8):
        cfg = RunCfg(cells_per_wave, int(use_active_mask), int(geometry),
                     int(force_exact), int(small_tier), int(count_work),
                     int(small_caps), int(learned), float(logit_threshold),
                     int(proposal_k))
        self._check(self._L.rpd_run(self._h, ctypes.byref(cfg)), "rpd_run")
        return self.stats()

    def graph_build(self, headroom=0.1, cells_per_wave=8, geometry=True,
                    small_caps=0, learned=0, logit_threshold=0.0,
                    proposal_k=8):
        """Capture the pipeline of the resident batch (plain, or learned
        with learned = 1 / 2) into a hipGraph (rpd_graph_build); the eager
        run it makes gives the results."""
        cfg = RunCfg(cells_per_wave, 1, int(geometry), 0, 1, 0,
                     int(small_caps), int(learned), float(logit_threshold),
                     int(proposal_k))
        self._check(self._L.rpd_graph_build(self._h, ctypes.byref(cfg),
                                            float(headroom)),
                    "rpd_graph_build")
        return self.stats()

    def graph_run(self):
        """Replay the captured pipeline on the current data; returns (stats,
        replayed) -- replayed False when it had to capture again."""
        r = ctypes.c_int32(0)
        self._check(self._L.rpd_graph_run(self._h, ctypes.byref(r)),
                    "rpd_graph_run")
        return self.stats(), bool(r.value)

    def graph_destroy(self):
        self._check(self._L.rpd_graph_destroy(self._h), "rpd_graph_destroy")

    def run_noalloc(self, inst, out, cells_per_wave=8, small_caps=0):
        """rpd_run_noalloc of ``inst`` (host arrays) into the caller's device
        buffers ``out`` = (cells, cells_cap, words, words_cap, geom, areas,
        areas_cap) (device addresses and capacities); returns (rc, stats
        dict, last error)."""
        a = self._arrays(inst)
        d = self._desc(inst, a)
        cfg = RunCfg(cells_per_wave, 1, 1, 0, 1, 0, int(small_caps), 0, 0.0,
                     8)
        o = NoAllocOut(*out)
        st = Stats()
        rc = self._L.rpd_run_noalloc(self._h, ctypes.byref(d),
                                     ctypes.byref(cfg), ctypes.byref(o),
                                     ctypes.byref(st))
        err = self._L.rpd_last_error(self._h).decode() if rc else ""
        return rc, st.as_dict(), err

    def stats(self):
        st = Stats()
        self._check(self._L.rpd_get_stats(self._h, ctypes.byref(st)),
                    "rpd_get_stats")
        return st.as_dict()

    def download(self, geometry=True):
        st = self.stats()
        cells = np.empty(st["n_cells"], CELL_DTYPE)
        words = np.empty(st["n_words"], np.uint32)
        geom = np.empty(4 * st["n_cells"], np.float64) if geometry else None
        areas = np.empty(st["n_areas"], np.float64) if geometry else None
        nul = ctypes.c_void_p()
        self._check(self._L.rpd_download_cells(
            self._h, _p(cells), _p(words),
            _p(geom) if geometry else nul, _p(areas) if geometry else nul),
            "rpd_download_cells")
        return RPDResult(cells, words, geom, areas)

    def power_neighbors(self, xyzw, box):
        """Regular-triangulation neighbours computed on the device:
        (valid[n], offsets[n+1], ids, device ms), the layout of
        prep.power_neighbors."""
        xyzw = np.ascontiguousarray(xyzw, np.float64)
        box = np.ascontiguousarray(box, np.float64)
        n = xyzw.shape[0]
        h = ctypes.c_void_p()
        ms = ctypes.c_double(0.0)
        self._check(self._L.rpd_power_neighbors_gpu(
            self._h, n, _p(xyzw), _p(box), ctypes.byref(ms),
            ctypes.byref(h)), "rpd_power_neighbors_gpu")
        try:
            total = self._L.rpd_nbr_total(h)
            valid = np.empty(n, np.int32)
            off = np.empty(n + 1, np.int64)
            ids = np.empty(max(total, 1), np.int32)
            self._check(self._L.rpd_nbr_get(h, _p(valid), _p(off), _p(ids)),
                        "rpd_nbr_get")
        finally:
            self._L.rpd_nbr_free(h)
        return valid, off, ids[:total], ms.value

    def calib_copy(self, nbytes):
        self._check(self._L.rpd_calib_copy(self._h, int(nbytes)),
                    "rpd_calib_copy")

    def site_centroids(self):
        """(n_site, 4): per site the volume-weighted centroid of its cells
        in the last run and their total volume (a site without volume
        keeps its centre)."""
        n = int(self._site_count())
        out = np.empty((max(n, 1), 4), np.float64)
        self._check(self._L.rpd_site_centroids(self._h, _p(out), None),
                    "rpd_site_centroids")
        return out[:n]

    def _site_count(self):
        to, so = self.batch_offsets()
        return so[-1]

    def adjacency(self):
        """Power-cell adjacency of the last run's cells, computed on the
        device (rpd_adjacency); returns a dict of numpy arrays named as the
        fields of rpd_adj_out plus the sizes."""
        z = AdjSizes()
        self._check(self._L.rpd_adjacency(self._h, ctypes.byref(z)),
                    "rpd_adjacency")
        nc, ns, nf, nfc, ne, nec = (z.n_cells, z.n_sites, z.n_facets,
                                    z.n_facet_cells, z.n_edges,
                                    z.n_edge_cells)
        a = dict(
            cell_euler=np.empty(nc, np.float32),
            cell_cc=np.empty(nc, np.int32),
            cell_nbr=np.empty((nc, 4), np.int32),
            site_cells=np.empty(ns, np.int32),
            site_cc=np.empty(ns, np.int32),
            site_euler=np.empty(ns, np.float64),
            facets=np.empty((nf, 5), np.int32),
            facet_euler=np.empty(nf, np.float64),
            facet_cells=np.empty((nfc, 2), np.int32),
            facet_cell_euler=np.empty(nfc, np.float32),
            edges=np.empty((ne, 6), np.int32),
            edge_cells=np.empty((nec, 2), np.int32))
        o = AdjOut(*[a[f].ctypes.data if a[f].size else None
                     for f, _ in AdjOut._fields_])
        self._check(self._L.rpd_adjacency_download(self._h, ctypes.byref(o)),
                    "rpd_adjacency_download")
        a["sizes"] = {f: getattr(z, f) for f, _ in AdjSizes._fields_}
        return a

    def set_tet_subset(self, tet_ids=None):
        """Clip only these tets of the resident mesh in the following
        (plain) runs; None or empty clears the subset."""
        if tet_ids is None or len(tet_ids) == 0:
            self._check(self._L.rpd_set_tet_subset(self._h, 0, None),
                        "rpd_set_tet_subset")
            return
        ids = np.ascontiguousarray(tet_ids, np.int32)
        self._check(self._L.rpd_set_tet_subset(self._h, ids.size, _p(ids)),
                    "rpd_set_tet_subset")

    def relation(self):
        st = self.stats()
        off = np.empty(self.n_tet + 1, np.int64)
        ids = np.empty(max(st["n_candidates"], 1), np.int32)
        self._check(self._L.rpd_download_relation(self._h, _p(off), _p(ids)),
                    "rpd_download_relation")
        return off, ids[:st["n_candidates"]]
