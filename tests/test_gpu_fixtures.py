"""GPU parity on hand-built fixtures: the reference's plane capacity with
masked neighbours, and degenerate sphere sets (bisectors on tet faces and
through tet vertices) on the plain and the learned pipelines."""
import numpy as np
import pytest

import oracle
from libmat_amd import prep
from rpd_compare import compare_with_oracle

pytestmark = pytest.mark.gpu


def _same_result(a, b):
    for f in ("cells", "words", "geom", "areas"):
        x, y = getattr(a, f), getattr(b, f)
        if f == "cells":
            for k in ("tet", "seed", "nb_v", "nb_p", "nb_e", "status",
                      "word_off", "area_off"):
                assert np.array_equal(x[k], y[k]), (f, k)
        else:
            assert np.array_equal(x, y), f


def cap_instance(n_cut, n_far):
    """One tet (corner simplex of edge 4); site 0 at (0.8, 0.8, 0.8) is the
    only flagged site.  Its list holds n_cut spheres on the +x axis whose
    bisectors x = 1.7 - 0.01 k (k = 1..n_cut) each cut the cell (so every
    one adds a plane and the cell keeps 6 vertices), then n_far spheres far
    on the -x side whose bisectors leave the whole tet on the seed's side
    (the plane masks skip them)."""
    verts = np.array([[0, 0, 0], [4, 0, 0], [0, 4, 0], [0, 0, 4]], np.float64)
    tets = np.array([[0, 1, 2, 3]], np.int32)
    v_adj, e_adj, f_adj, f_id, n_sf = prep.tet_adjacency(4, tets)
    c = np.array([0.8, 0.8, 0.8])
    sites = [[*c, 0.0]]
    for k in range(1, n_cut + 1):
        d = 2 * (0.9 - 0.01 * k)
        sites.append([c[0] + d, c[1], c[2], 0.0])
    for k in range(n_far):
        sites.append([c[0] - 20 - k, c[1], c[2], 0.0])
    sites = np.array(sites, np.float64)
    n = sites.shape[0]
    flags = np.zeros(n, np.uint32)
    flags[0] = 1
    knn_off = np.zeros(n + 1, np.int64)
    knn_off[1:] = n - 1
    knn_ids = np.arange(1, n, dtype=np.int32)
    return prep.Instance(verts, tets, v_adj, e_adj, f_adj, f_id, n_sf, sites,
                         flags, knn_off, knn_ids, np.arange(n, dtype=np.int32))


@pytest.mark.parametrize("n_cut,n_far", [(47, 5), (48, 0), (48, 5), (49, 0),
                                         (49, 5)])
@pytest.mark.parametrize("mask", [True, False])
def test_capacity_with_masked_neighbours(ctx, n_cut, n_far, mask):
    # every kept plane appends at least 3 edges, so the edge capacity (152)
    # is reached before the plane capacity (64): 48 cuts leave 150 edges and
    # 52 planes, the 49th cut overflows the edge list (status 7); the
    # neighbours the masks skip change nothing
    inst = cap_instance(n_cut, n_far)
    orc = oracle.run(inst)
    print(f"[cap] cut={n_cut} far={n_far} mask={mask}: oracle cells "
          f"{orc.cells.size} drops {orc.drops.tolist()}", flush=True)
    ctx.upload(inst)
    st = ctx.run(use_active_mask=mask)
    res = ctx.download()
    compare_with_oracle(res, orc)
    assert st["n_dropped"] == [int(x) for x in orc.drops]
    if n_cut >= 49:
        assert len(res) == 0 and st["n_dropped"][7] == 1
    else:
        assert len(res) == 1 and int(res.cells["nb_p"][0]) == 4 + n_cut
        assert int(res.cells["nb_e"][0]) == 6 + 3 * n_cut


def octant_instance():
    # bisectors on grid planes: vertex-on-plane determinants exactly 0
    verts, tets = prep.grid_mesh(4, 0.0, 0)
    sph = [[x, y, z, 0.1] for x in (0.25, 0.75) for y in (0.25, 0.75)
           for z in (0.25, 0.75)]
    return prep.build_instance(verts, tets, np.array(sph))


def apex_instance():
    """One tet with a corner at the origin and the rest in x < 0; five
    equal spheres on the unit sphere: four around the x = 0 plane and
    sphere 4 at (1, 0, 0), whose power cell is the cone x >= |y|, x >= |z|
    with its apex at the origin.  Every bisector of sphere 4 has a tet
    corner strictly on its side, so the reference relation keeps it, and
    its cell in the tet is the single point (0, 0, 0): a zero-volume cell
    the plain path emits.  The origin's owner is sphere 0 (ties go to the
    smaller id), so sphere 4 is no owner."""
    v = np.array([[0, 0, 0], [-1, 3, 0], [-1, -3, -2], [-1, -3, 2]],
                 np.float64)  # positively oriented
    tets = np.array([[0, 1, 2, 3]], np.int32)
    v_adj, e_adj, f_adj, f_id, n_sf = prep.tet_adjacency(4, tets)
    S = np.array([[0, 1, 0, 0.01], [0, -1, 0, 0.01], [0, 0, 1, 0.01],
                  [0, 0, -1, 0.01], [1, 0, 0, 0.01]], np.float64)
    # power-diagram neighbours (0-1 and 2-3 meet only at the origin)
    lists = [[2, 3, 4], [2, 3, 4], [0, 1, 4], [0, 1, 4], [0, 1, 2, 3]]
    off = np.zeros(6, np.int64)
    off[1:] = np.cumsum([len(x) for x in lists])
    ids = np.array(sum(lists, []), np.int32)
    return prep.Instance(v, tets, v_adj, e_adj, f_adj, f_id, n_sf, S,
                         np.ones(5, np.uint32), off, ids,
                         np.arange(5, dtype=np.int32))


def lattice_instance():
    # 27 equal spheres on the lattice {0.25, 0.5, 0.75}^3 over an 8^3 grid:
    # axis bisectors on grid planes, diagonal ones through grid vertices,
    # cells of neighbours that touch a tet only in a face, edge or point
    verts, tets = prep.grid_mesh(8, 0.0, 0)
    sph = [[x, y, z, 0.05] for x in (0.25, 0.5, 0.75)
           for y in (0.25, 0.5, 0.75) for z in (0.25, 0.5, 0.75)]
    return prep.build_instance(verts, tets, np.array(sph))


@pytest.mark.parametrize("which", ["octant", "lattice", "apex"])
def test_degenerate_spheres_learned_matches_plain(ctx, which):
    inst = {"octant": octant_instance, "lattice": lattice_instance,
            "apex": apex_instance}[which]()
    orc = oracle.run(inst)
    ctx.upload(inst)
    st = ctx.run()
    plain = ctx.download()
    m = compare_with_oracle(plain, orc)
    zero_vol = int((np.abs(plain.geom[:, 0]) == 0).sum())
    if which == "apex":
        assert zero_vol == 1 and int(plain.cells["seed"][4]) == 4
    print(f"[degenerate] {which}: {m}, exact re-runs {st['n_rerun_exact']}, "
          f"zero-volume cells {zero_vol}", flush=True)
    ctx.mlp_init(11)
    for kw in (dict(learned=1, logit_threshold=0.0),
               dict(learned=1, logit_threshold=2.0),
               dict(learned=2, proposal_k=1), dict(learned=2, proposal_k=4)):
        st = ctx.run(**kw)
        res = ctx.download()
        print(f"[degenerate] {which} {kw}: first-round {st['learned'][0]}, "
              f"re-run tets {st['learned'][1]}, full relation "
              f"{st['learned'][3]}", flush=True)
        _same_result(res, plain)
        compare_with_oracle(res, orc)
        # the same step captured (exact tier, full-relation fallback and
        # closure rounds on the device): the eager run's cells and sets
        off_e, ids_e = ctx.relation()
        ctx.graph_build(headroom=0.1, cells_per_wave=8, **kw)
        st_g, rep = ctx.graph_run()
        assert rep and st_g["learned"][3] == st["learned"][3]
        _same_result(ctx.download(), plain)
        off_g, ids_g = ctx.relation()
        assert np.array_equal(off_g, off_e) and np.array_equal(ids_g, ids_e)
        ctx.graph_destroy()
