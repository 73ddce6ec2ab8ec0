"""Host-side preparation of RPD inputs (thin ctypes layer over rpd_prep.h).

An :class:`Instance` holds exactly the arrays the reference's driver hands to
``compute_clipped_voro_diagram_GPU`` (src/rpd3d/voronoi.h:52-61), in this
package's layout: binary64 vertices and spheres, int32 tets, per-tet
adjacency counts and face ids (io.cxx:238-335 semantics) and CSR neighbour
lists (the RT 1-ring of triangulation.cxx:324-549, ascending site ids).
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from ._native import lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with code {rc}")


def grid_mesh(n, jitter_rel=0.0, seed=0):
    """Unit cube, n^3 cubes x 6 Kuhn tets; returns (verts[nv,3], tets[nt,4])."""
    nv = (n + 1) ** 3
    nt = 6 * n ** 3
    verts = np.empty((nv, 3), np.float64)
    tets = np.empty((nt, 4), np.int32)
    _check(lib().rpd_synth_grid_mesh(n, float(jitter_rel), seed, _p(verts),
                                     _p(tets)), "rpd_synth_grid_mesh")
    return verts, tets


def spheres(n, rule, seed, rmin=0.02, rmax=0.08):
    """(x, y, z, r) per sphere; rule 0 uniform radii, 1 boundary rule, 2 GMM."""
    out = np.empty((n, 4), np.float64)
    _check(lib().rpd_synth_spheres(n, rule, rmin, rmax, seed, _p(out)),
           "rpd_synth_spheres")
    return out


def tet_adjacency(n_vert, tets):
    tets = np.ascontiguousarray(tets, np.int32)
    nt = tets.shape[0]
    v_adj = np.empty(n_vert, np.int32)
    e_adj = np.empty((nt, 6), np.int32)
    f_adj = np.empty((nt, 4), np.int32)
    f_id = np.empty((nt, 4), np.int32)
    n_sf = ctypes.c_int32(0)
    _check(lib().rpd_tet_adjacency(n_vert, nt, _p(tets), _p(v_adj), _p(e_adj),
                                   _p(f_adj), _p(f_id),
                                   ctypes.cast(ctypes.byref(n_sf),
                                               ctypes.c_void_p)),
           "rpd_tet_adjacency")
    return v_adj, e_adj, f_adj, f_id, n_sf.value


def power_neighbors(xyzw, box, n_threads=0):
    """Returns (valid[n], offsets[n+1], ids) of the power-diagram 1-rings."""
    xyzw = np.ascontiguousarray(xyzw, np.float64)
    box = np.ascontiguousarray(box, np.float64)
    n = xyzw.shape[0]
    h = ctypes.c_void_p()
    _check(lib().rpd_power_neighbors(n, _p(xyzw), _p(box), n_threads,
                                     ctypes.byref(h)), "rpd_power_neighbors")
    try:
        total = lib().rpd_nbr_total(h)
        valid = np.empty(n, np.int32)
        off = np.empty(n + 1, np.int64)
        ids = np.empty(max(total, 1), np.int32)
        _check(lib().rpd_nbr_get(h, _p(valid), _p(off), _p(ids)),
               "rpd_nbr_get")
    finally:
        lib().rpd_nbr_free(h)
    return valid, off, ids[:total]


@dataclass
class Instance:
    verts: np.ndarray          # (n_vert, 3) f64
    tets: np.ndarray           # (n_tet, 4) i32
    v_adj: np.ndarray          # (n_vert,) i32
    e_adj: np.ndarray          # (n_tet, 6) i32
    f_adj: np.ndarray          # (n_tet, 4) i32
    f_id: np.ndarray           # (n_tet, 4) i32
    n_sf: int
    sites: np.ndarray          # (n_site, 4) f64: x, y, z, w = r^2
    flags: np.ndarray          # (n_site,) u32
    knn_off: np.ndarray        # (n_site + 1,) i64
    knn_ids: np.ndarray        # i32, ascending per site
    site2sphere: np.ndarray = field(default=None)  # (n_site,) i32

    @property
    def n_vert(self):
        return self.verts.shape[0]

    @property
    def n_tet(self):
        return self.tets.shape[0]

    @property
    def n_site(self):
        return self.sites.shape[0]

    @property
    def site_k(self):
        d = np.diff(self.knn_off)
        return int(d.max()) if d.size else 0


UNIT_BOX = np.array([-1e-6, -1e-6, -1e-6, 1 + 1e-6, 1 + 1e-6, 1 + 1e-6])


def build_instance(verts, tets, xyzr, box=UNIT_BOX, n_threads=0):
    """Adjacency, power neighbours and sites of the non-hidden spheres.

    Sites are the spheres whose power cell meets ``box``, in ascending sphere
    order; every site is flagged (a full RPD, as RPD3D_GPU::calculate)."""
    verts = np.ascontiguousarray(verts, np.float64)
    tets = np.ascontiguousarray(tets, np.int32)
    v_adj, e_adj, f_adj, f_id, n_sf = tet_adjacency(verts.shape[0], tets)
    xyzw = np.ascontiguousarray(xyzr, np.float64).copy()
    xyzw[:, 3] = xyzw[:, 3] * xyzw[:, 3]
    valid, off, ids = power_neighbors(xyzw, box, n_threads)
    keep = np.nonzero(valid)[0].astype(np.int32)
    remap = -np.ones(xyzw.shape[0], np.int64)
    remap[keep] = np.arange(keep.size)
    deg = np.diff(off)[keep]
    knn_off = np.zeros(keep.size + 1, np.int64)
    np.cumsum(deg, out=knn_off[1:])
    sel = np.concatenate([np.arange(off[i], off[i + 1]) for i in keep]) \
        if keep.size else np.zeros(0, np.int64)
    knn_ids = remap[ids[sel]].astype(np.int32) if sel.size else \
        np.zeros(0, np.int32)
    assert (knn_ids >= 0).all()
    return Instance(verts, tets, v_adj, e_adj, f_adj, f_id, n_sf,
                    np.ascontiguousarray(xyzw[keep]),
                    np.ones(keep.size, np.uint32), knn_off, knn_ids, keep)


def sites_of(xyzr, box=UNIT_BOX, n_threads=0, ctx=None):
    """Sites (non-hidden spheres in ascending order, w = r^2), flags and
    CSR neighbour lists of a sphere set: the per-iteration step of sphere
    optimisation, with the neighbours from the host builder, or from the
    device builder of ``ctx`` (an rpd.Context; same result).  Returns
    (sites, flags, knn_off, knn_ids, keep)."""
    xyzw = np.ascontiguousarray(xyzr, np.float64).copy()
    xyzw[:, 3] = xyzw[:, 3] * xyzw[:, 3]
    if ctx is not None:
        valid, off, ids, _ = ctx.power_neighbors(xyzw, box)
    else:
        valid, off, ids = power_neighbors(xyzw, box, n_threads)
    keep = np.nonzero(valid)[0].astype(np.int32)
    remap = -np.ones(xyzw.shape[0], np.int64)
    remap[keep] = np.arange(keep.size)
    deg = np.diff(off)[keep]
    knn_off = np.zeros(keep.size + 1, np.int64)
    np.cumsum(deg, out=knn_off[1:])
    sel = np.concatenate([np.arange(off[i], off[i + 1]) for i in keep]) \
        if keep.size else np.zeros(0, np.int64)
    knn_ids = remap[ids[sel]].astype(np.int32) if sel.size else \
        np.zeros(0, np.int32)
    assert (knn_ids >= 0).all()
    return (np.ascontiguousarray(xyzw[keep]), np.ones(keep.size, np.uint32),
            knn_off, knn_ids, keep)


def sites_both(xyzr, box=UNIT_BOX, ctx=None):
    """The sites of ``sites_of`` (non-hidden spheres, renumbered) and the
    same neighbour lists over every sphere in its own numbering (flag 1 for
    the non-hidden ones, 0 with an empty list for the hidden ones, which no
    list holds): the second form keeps each instance's site count fixed
    across sphere updates.  One neighbour computation for both."""
    xyzw = np.ascontiguousarray(xyzr, np.float64).copy()
    xyzw[:, 3] = xyzw[:, 3] * xyzw[:, 3]
    if ctx is not None:
        valid, off, ids, _ = ctx.power_neighbors(xyzw, box)
    else:
        valid, off, ids = power_neighbors(xyzw, box, 0)
    valid = np.asarray(valid).astype(bool)
    keep = np.nonzero(valid)[0].astype(np.int32)
    remap = -np.ones(xyzw.shape[0], np.int64)
    remap[keep] = np.arange(keep.size)
    deg = np.diff(off)[keep]
    knn_off = np.zeros(keep.size + 1, np.int64)
    np.cumsum(deg, out=knn_off[1:])
    sel = np.concatenate([np.arange(off[i], off[i + 1]) for i in keep]) \
        if keep.size else np.zeros(0, np.int64)
    knn_ids = remap[ids[sel]].astype(np.int32) if sel.size else \
        np.zeros(0, np.int32)
    compact = (np.ascontiguousarray(xyzw[keep]), np.ones(keep.size, np.uint32),
               knn_off, knn_ids)
    deg_all = np.where(valid, np.diff(off), 0)
    off_all = np.zeros(xyzw.shape[0] + 1, np.int64)
    np.cumsum(deg_all, out=off_all[1:])
    ids_all = ids[sel].astype(np.int32) if sel.size else np.zeros(0, np.int32)
    every = (xyzw, valid.astype(np.uint32), off_all, ids_all)
    return compact, every


@dataclass
class BatchWorkload:
    """Independent RPD instances of uneven size and their sphere sets."""
    instances: list
    spheres: list     # (n_sphere, 4) x, y, z, r per instance
    grid_n: list      # realised grid resolution per instance


def batch_workload(n_inst=256, seed=0, tets=(2000, 200000),
                   n_spheres=(100, 20000), n_threads=0, progress=None):
    """Tets log-uniform in ``tets`` (a jittered n^3 Kuhn grid with
    n = round((T/6)^(1/3)), so 6 n^3 tets), spheres log-uniform in
    ``n_spheres`` with centres U(0,1)^3 and radii U[0.02, 0.08]; sizes drawn
    by numpy's PCG64 from ``seed``, geometry by the project generator with
    seed ``seed * 100003 + instance``."""
    rng = np.random.default_rng(seed)
    insts, sph, ns = [], [], []
    for b in range(n_inst):
        t = np.exp(rng.uniform(np.log(tets[0]), np.log(tets[1])))
        m = int(round(np.exp(rng.uniform(np.log(n_spheres[0]),
                                         np.log(n_spheres[1])))))
        n = max(1, int(round((t / 6.0) ** (1.0 / 3.0))))
        s = seed * 100003 + b
        verts, tt = grid_mesh(n, 0.1, s)
        xyzr = spheres(m, 0, s, 0.02, 0.08)
        insts.append(build_instance(verts, tt, xyzr, n_threads=n_threads))
        sph.append(xyzr)
        ns.append(n)
        if progress is not None:
            progress(b)
    return BatchWorkload(insts, sph, ns)


def gmm_workload(n_inst=128, seed=0, tets=(50000, 500000),
                 n_spheres=(5000, 40000), n_threads=0, progress=None):
    """Clustered-sphere stress batch: tets log-uniform in ``tets`` (jittered
    n^3 Kuhn grids as in :func:`batch_workload`), spheres log-uniform in
    ``n_spheres`` from the 32-component Gaussian mixture of
    :func:`spheres` rule 2 (sigma ~ U[0.02, 0.1], r ~ lognormal(-3.5, 0.6)
    clipped to the boundary distance); sizes by numpy's PCG64 from
    ``seed``, geometry by the project generator with seed
    ``seed * 100003 + instance + 50000``.  ``progress(i)`` is called after
    each instance."""
    rng = np.random.default_rng(seed + 7777)
    insts, sph, ns = [], [], []
    for b in range(n_inst):
        t = np.exp(rng.uniform(np.log(tets[0]), np.log(tets[1])))
        m = int(round(np.exp(rng.uniform(np.log(n_spheres[0]),
                                         np.log(n_spheres[1])))))
        n = max(1, int(round((t / 6.0) ** (1.0 / 3.0))))
        s = seed * 100003 + b + 50000
        verts, tt = grid_mesh(n, 0.1, s)
        xyzr = spheres(m, 2, s)
        insts.append(build_instance(verts, tt, xyzr, n_threads=n_threads))
        sph.append(xyzr)
        ns.append(n)
        if progress is not None:
            progress(b)
    return BatchWorkload(insts, sph, ns)


def jitter_spheres(xyzr, sigma, seed):
    """Sphere centres moved by N(0, sigma^2) per coordinate (numpy PCG64
    from ``seed``), kept inside the open unit cube; radii unchanged."""
    rng = np.random.default_rng(seed)
    out = np.array(xyzr, np.float64, copy=True)
    out[:, :3] = np.clip(out[:, :3] + rng.normal(0.0, sigma, (out.shape[0], 3)),
                         1e-6, 1 - 1e-6)
    return out


def cube10_instance(seed=0, n_threads=0):
    """10^3 Kuhn grid (6,000 tets), 500 spheres r ~ U[0.02, 0.08]."""
    verts, tets = grid_mesh(10, 0.0, seed)
    return build_instance(verts, tets, spheres(500, 0, seed, 0.02, 0.08),
                          n_threads=n_threads)


def cube64_instance(seed=0, n_threads=0, n=64, n_spheres=50000):
    """64^3 jittered grid (1,572,864 tets), 50,000 spheres, boundary rule."""
    verts, tets = grid_mesh(n, 0.1, seed)
    return build_instance(verts, tets, spheres(n_spheres, 1, seed),
                          n_threads=n_threads)
