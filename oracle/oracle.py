"""ORACLE (test infrastructure only): ctypes binding of oracle/liboracle.so.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, always as the checker of the HIP path, never as the path.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
MAX_P, MAX_T, MAX_E = 64, 96, 152

# mirrors struct CellOut in rpd_oracle.cpp
CELL_DTYPE = np.dtype([
    ("tet", np.int32), ("seed", np.int32),
    ("status", np.uint8), ("nb_v", np.uint8), ("nb_p", np.uint8),
    ("nb_e", np.uint8),
    ("ver", np.uint32, MAX_T), ("src", np.int32, MAX_P),
    ("edge", np.uint32, MAX_E), ("geo", np.float64, 4),
    ("area", np.float64, MAX_P),
], align=True)


class _Input(ctypes.Structure):
    _fields_ = [("n_vert", ctypes.c_int32), ("n_tet", ctypes.c_int32),
                ("n_site", ctypes.c_int32)] + [
        (n, ctypes.c_void_p) for n in
        ("verts", "tets", "v_adj", "e_adj", "f_adj", "f_id", "sites",
         "flags", "knn_off", "knn_ids")]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        L.orc_run.argtypes = [ctypes.POINTER(_Input), vp, ctypes.c_int32,
                              ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
        L.orc_run.restype = vp
        L.orc_clip.argtypes = [ctypes.POINTER(_Input), vp, ctypes.c_int32,
                               vp, vp, ctypes.c_int32, ctypes.c_int32]
        L.orc_clip.restype = vp
        L.orc_free.argtypes = [vp]
        L.orc_n_cells.argtypes = [vp]
        L.orc_n_cells.restype = ctypes.c_int64
        L.orc_n_rel.argtypes = [vp]
        L.orc_n_rel.restype = ctypes.c_int64
        L.orc_cell_bytes.restype = ctypes.c_int64
        L.orc_get_relation.argtypes = [vp, vp, vp]
        L.orc_get_cells.argtypes = [vp, vp]
        L.orc_get_drops.argtypes = [vp, vp]
        L.orc_get_counters.argtypes = [vp, vp]
        L.orc_adjacency.argtypes = [ctypes.c_char_p, ctypes.c_char_p,
                                    ctypes.c_int32, ctypes.c_char_p]
        L.orc_adjacency.restype = ctypes.c_int
        L.orc_exact_det4_sign.argtypes = [vp]
        L.orc_exact_det4_sign.restype = ctypes.c_int
        assert L.orc_cell_bytes() == CELL_DTYPE.itemsize, \
            (L.orc_cell_bytes(), CELL_DTYPE.itemsize)
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleResult:
    def __init__(self, rel_off, rel_ids, cells, drops, counters, tets):
        self.rel_off = rel_off      # per requested tet
        self.rel_ids = rel_ids
        self.cells = cells          # structured array, CELL_DTYPE
        self.drops = drops          # dropped cells per status code 0..8
        self.counters = counters    # n_det, n_bisector, n_exact, n_vout
        self.tets = tets


def _input(inst, keep):
    def arr(a, dt):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return _p(a)

    return _Input(inst.n_vert, inst.n_tet, inst.n_site,
                  arr(inst.verts, np.float64), arr(inst.tets, np.int32),
                  arr(inst.v_adj, np.int32), arr(inst.e_adj, np.int32),
                  arr(inst.f_adj, np.int32), arr(inst.f_id, np.int32),
                  arr(inst.sites, np.float64), arr(inst.flags, np.uint32),
                  arr(inst.knn_off, np.int64), arr(inst.knn_ids, np.int32))


def _collect(L, h, tet_arr):
    try:
        nc, nr = L.orc_n_cells(h), L.orc_n_rel(h)
        off = np.empty(tet_arr.size + 1, np.int64)
        ids = np.empty(max(nr, 1), np.int32)
        L.orc_get_relation(h, _p(off), _p(ids))
        cells = np.empty(nc, CELL_DTYPE)
        if nc:
            L.orc_get_cells(h, _p(cells))
        drops = np.empty(9, np.int32)
        L.orc_get_drops(h, _p(drops))
        cnt = np.empty(4, np.int64)
        L.orc_get_counters(h, _p(cnt))
    finally:
        L.orc_free(h)
    return OracleResult(off, ids[:nr], cells, drops, cnt, tet_arr)


def clip(inst, tets, cand_off, cand_ids, mode=0, n_threads=8):
    """Clip only, for the given per-tet candidate lists (CSR over tets)."""
    L = lib()
    keep = []
    I = _input(inst, keep)
    tet_arr = np.ascontiguousarray(tets, np.int32)
    off = np.ascontiguousarray(cand_off, np.int64)
    ids = np.ascontiguousarray(cand_ids, np.int32)
    h = L.orc_clip(ctypes.byref(I), _p(tet_arr), tet_arr.size, _p(off),
                   _p(ids), mode, n_threads)
    return _collect(L, h, tet_arr)


def run(inst, tets=None, mode=0, brute_force=False, n_threads=8):
    """Relation + clip for `tets` (default: all) of a prep.Instance."""
    L = lib()
    keep = []
    I = _input(inst, keep)
    if tets is None:
        tl, n_list = None, 0
        tet_arr = np.arange(inst.n_tet, dtype=np.int32)
    else:
        tet_arr = np.ascontiguousarray(tets, np.int32)
        tl, n_list = _p(tet_arr), tet_arr.size
    h = L.orc_run(ctypes.byref(I), tl, n_list, mode, int(brute_force),
                  n_threads)
    return _collect(L, h, tet_arr)


def exact_det4_sign(m):
    m = np.ascontiguousarray(m, np.float64).reshape(16)
    return lib().orc_exact_det4_sign(_p(m))


def adjacency(cell_path, es2fe_path, max_surf_fid, out_path):
    """Per-sphere power-cell adjacency and Euler sums of the cells in a cell
    file (format in adj_oracle.cpp), written as the canonical text dump."""
    rc = lib().orc_adjacency(str(cell_path).encode(),
                             str(es2fe_path).encode() if es2fe_path else b"",
                             int(max_surf_fid), str(out_path).encode())
    if rc != 0:
        raise RuntimeError(f"orc_adjacency failed ({rc})")
