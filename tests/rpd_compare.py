"""Comparison of GPU results with the oracle (test helper)."""
import numpy as np


def gpu_cells_to_arrays(res, max_p=64, max_t=96, max_e=152):
    """Expand the compact GPU stream into fixed-width arrays (oracle layout)."""
    n = len(res)
    ver = np.zeros((n, max_t), np.uint32)
    src = np.zeros((n, max_p), np.int32)
    edge = np.zeros((n, max_e), np.uint32)
    area = np.zeros((n, max_p), np.float64)
    for q in range(n):
        v, s, e, a = res.cell(q)
        ver[q, :v.size] = v
        src[q, :s.size] = s
        edge[q, :e.size] = e
        if a is not None:
            area[q, :a.size] = a
    return ver, src, edge, area


def compare_with_oracle(res, orc, geom_rtol=1e-6, check_geometry=True):
    """Bit-exact combinatorics and geometry within `geom_rtol` (relative to
    the cell's own size).  Returns a dict of measured maxima; raises
    AssertionError with the first mismatch."""
    oc = orc.cells
    gc = res.cells
    assert gc.size == oc.size, f"cell count {gc.size} != oracle {oc.size}"
    assert np.array_equal(gc["tet"], oc["tet"]), "tet ids differ"
    assert np.array_equal(gc["seed"], oc["seed"]), "seed ids differ"
    for f in ("nb_v", "nb_p", "nb_e"):
        bad = np.nonzero(gc[f] != oc[f])[0]
        assert bad.size == 0, f"{f} differs at cell {bad[:5]}"
    ver, src, edge, area = gpu_cells_to_arrays(res)
    for name, g, o, cnt in (("ver", ver, oc["ver"], oc["nb_v"]),
                            ("src", src, oc["src"], oc["nb_p"]),
                            ("edge", edge, oc["edge"], oc["nb_e"])):
        valid = np.arange(g.shape[1])[None, :] < cnt[:, None].astype(np.int64)
        bad = np.nonzero(((g != o) & valid).any(axis=1))[0]
        assert bad.size == 0, (f"{name} differs at cells {bad[:5]} "
                               f"(tet {oc['tet'][bad[:1]]}, seed "
                               f"{oc['seed'][bad[:1]]})")
    out = {"n_cells": int(gc.size)}
    if check_geometry:
        gg, og = res.geom, oc["geo"]
        scale = np.cbrt(np.maximum(np.abs(og[:, 0]), 1e-300))
        vol_err = np.abs(gg[:, 0] - og[:, 0]) / np.maximum(
            np.abs(og[:, 0]), 1e-300)
        cen_err = np.abs(gg[:, 1:] - og[:, 1:]).max(axis=1) / scale
        area_scale = np.maximum(og[:, 0:1] ** (2.0 / 3.0), 1e-300)
        pvalid = np.arange(area.shape[1])[None, :] < \
            oc["nb_p"][:, None].astype(np.int64)
        ar_err = (np.abs(area - oc["area"]) * pvalid / area_scale).max(axis=1)
        out.update(max_vol_rel_err=float(vol_err.max(initial=0)),
                   max_centroid_rel_err=float(cen_err.max(initial=0)),
                   max_area_rel_err=float(ar_err.max(initial=0)))
        assert out["max_vol_rel_err"] <= geom_rtol, out
        assert out["max_centroid_rel_err"] <= geom_rtol, out
        assert out["max_area_rel_err"] <= geom_rtol, out
    return out


def compare_relation(gpu_off, gpu_ids, orc):
    assert np.array_equal(gpu_off, orc.rel_off), "relation offsets differ"
    assert np.array_equal(gpu_ids, orc.rel_ids), "relation ids differ"


def digest_oracle(cells):
    """SHA-256 over the combinatorics of oracle cells, in order."""
    import hashlib
    h = hashlib.sha256()
    for c in cells:
        nv, np_, ne = int(c["nb_v"]), int(c["nb_p"]), int(c["nb_e"])
        h.update(np.array([c["tet"], c["seed"], nv, np_, ne],
                          np.int64).tobytes())
        h.update(c["ver"][:nv].astype(np.uint32).tobytes())
        h.update(c["src"][:np_].astype(np.int32).tobytes())
        h.update(c["edge"][:ne].astype(np.uint32).tobytes())
    return h.hexdigest()


def digest_gpu(res):
    """Same digest over the compact GPU stream."""
    import hashlib
    h = hashlib.sha256()
    for q in range(len(res)):
        c = res.cells[q]
        nv, np_, ne = int(c["nb_v"]), int(c["nb_p"]), int(c["nb_e"])
        h.update(np.array([c["tet"], c["seed"], nv, np_, ne],
                          np.int64).tobytes())
        v, s, e, _ = res.cell(q)
        h.update(v.astype(np.uint32).tobytes())
        h.update(s.astype(np.int32).tobytes())
        h.update(e.astype(np.uint32).tobytes())
    return h.hexdigest()


def tet_volumes(inst):
    V = inst.verts[inst.tets]
    return np.abs(np.einsum("ij,ij->i", V[:, 1] - V[:, 0],
                            np.cross(V[:, 2] - V[:, 0],
                                     V[:, 3] - V[:, 0]))) / 6.0
