"""GPU: per-site centroids of the cells (the Lloyd update of
compute_new_site, voronoi.cu:40-69, over the barycentre sums of
convex_cell.cu:1009-1054) and the tet-subset run of the partial RPD
(calculate_partial, rpd_api.cxx:249-262), through the C-ABI.

Centroids are checked against FP64 sums over the oracle's cells (tolerance
1e-9 relative: the cells' own geometry agrees with the oracle to about
1e-11) and against closed forms: one sphere owns the whole cube (centroid
(1/2, 1/2, 1/2), volume 1); two equal spheres split it at x = 1/2
(centroids at x = 1/4 and 3/4, volume 1/2 each).  A subset run must give
bit for bit the cells of the full run on those tets.
"""
import numpy as np
import pytest

import oracle
from libmat_amd import prep
from rpd_compare import compare_with_oracle

pytestmark = pytest.mark.gpu

TOL = 1e-9


def oracle_centroids(inst, orc):
    """Per site: sum vol * centroid / sum vol, sum vol over the oracle's
    cells; sites without volume keep their centre."""
    n = inst.n_site
    vol = np.zeros(n)
    mom = np.zeros((n, 3))
    seed = orc.cells["seed"].astype(np.int64)
    geo = orc.cells["geo"]
    np.add.at(vol, seed, geo[:, 0])
    np.add.at(mom, seed, geo[:, 0:1] * geo[:, 1:4])
    out = np.empty((n, 4))
    out[:, 3] = vol
    has = vol != 0
    out[has, :3] = mom[has] / vol[has, None]
    out[~has, :3] = inst.sites[~has, :3]
    return out


@pytest.mark.parametrize("seed", [0, 2])
def test_site_centroids_match_oracle(ctx, seed):
    inst = prep.cube10_instance(seed)
    orc = oracle.run(inst)
    print(f"[centroids] cube10 seed={seed} n_site={inst.n_site}", flush=True)
    ctx.upload(inst)
    ctx.run()
    got = ctx.site_centroids()
    exp = oracle_centroids(inst, orc)
    assert got.shape == exp.shape
    err_v = np.max(np.abs(got[:, 3] - exp[:, 3]))
    err_c = np.max(np.abs(got[:, :3] - exp[:, :3]))
    print(f"[centroids] max |dvol| {err_v:.2e} max |dc| {err_c:.2e}",
          flush=True)
    assert err_v <= TOL * max(1.0, np.max(exp[:, 3]))
    assert err_c <= TOL
    # sites without cells keep their centre exactly
    empty = exp[:, 3] == 0
    assert np.array_equal(got[empty, :3], inst.sites[empty, :3])
    # the cells tile the unit cube
    assert abs(got[:, 3].sum() - 1.0) < 1e-12


@pytest.mark.parametrize("n_grid", [4, 5])
def test_site_centroids_closed_forms(ctx, n_grid):
    verts, tets = prep.grid_mesh(n_grid, 0.0, 0)
    one = prep.build_instance(verts, tets, np.array([[0.4, 0.45, 0.55, 0.1]]))
    ctx.upload(one)
    ctx.run()
    c = ctx.site_centroids()
    print(f"[centroids] one sphere n={n_grid}: {c[0]}", flush=True)
    assert np.allclose(c[0], [0.5, 0.5, 0.5, 1.0], rtol=0, atol=1e-13)
    two = prep.build_instance(verts, tets, np.array(
        [[0.25, 0.5, 0.5, 0.1], [0.75, 0.5, 0.5, 0.1]]))
    ctx.upload(two)
    ctx.run()
    c = ctx.site_centroids()
    print(f"[centroids] two spheres n={n_grid}: {c}", flush=True)
    assert np.allclose(c, [[0.25, 0.5, 0.5, 0.5], [0.75, 0.5, 0.5, 0.5]],
                       rtol=0, atol=1e-13)
    # an unflagged site (the 2-ring of a partial RPD) still clips its
    # neighbour but has no cells: it keeps its centre, with volume 0
    two.flags = np.array([1, 0], np.uint32)
    ctx.upload(two)
    ctx.run()
    c = ctx.site_centroids()
    assert np.allclose(c[0], [0.25, 0.5, 0.5, 0.5], rtol=0, atol=1e-13)
    assert np.array_equal(c[1], [0.75, 0.5, 0.5, 0.0])


def test_site_centroids_learned_equals_plain(ctx):
    inst = prep.cube10_instance(1)
    ctx.upload(inst)
    ctx.run()
    a = ctx.site_centroids()
    ctx.mlp_init(0)
    for kw in (dict(learned=1), dict(learned=2, proposal_k=4)):
        ctx.run(**kw)
        b = ctx.site_centroids()
        print(f"[centroids] learned {kw}: identical "
              f"{np.array_equal(a, b)}", flush=True)
        assert np.array_equal(a, b)


def test_site_centroids_batch_equals_single(ctx):
    insts = [prep.cube10_instance(s) for s in (0, 3)]
    single = []
    for inst in insts:
        ctx.upload(inst)
        ctx.run()
        single.append(ctx.site_centroids())
    _, so = ctx.upload_batch(insts)
    ctx.run()
    both = ctx.site_centroids()
    for b in range(2):
        assert np.array_equal(both[so[b]:so[b + 1]], single[b])


def test_site_centroids_need_geometry(ctx):
    inst = prep.cube10_instance(0)
    ctx.upload(inst)
    ctx.run(geometry=False)
    with pytest.raises(RuntimeError, match="geometry"):
        ctx.site_centroids()


def _cells_of(res, keep_tet):
    """(tet, seed) -> (words, geom, areas) of the cells whose tet is kept."""
    out = {}
    for q in range(len(res)):
        c = res.cells[q]
        t = int(c["tet"])
        if not keep_tet[t]:
            continue
        o, a = int(c["word_off"]), int(c["area_off"])
        n = int(c["nb_v"]) + int(c["nb_p"]) + int(c["nb_e"])
        out[(t, int(c["seed"]))] = (
            tuple(int(c[f]) for f in ("nb_v", "nb_p", "nb_e", "status")),
            res.words[o:o + n].tobytes(), res.geom[q].tobytes(),
            res.areas[a:a + int(c["nb_p"])].tobytes())
    return out


@pytest.mark.parametrize("frac", [0.0005, 0.3])
def test_tet_subset_equals_full_run_on_those_tets(ctx, frac):
    inst = prep.cube10_instance(2)
    ctx.upload(inst)
    ctx.run()
    full = ctx.download()
    off_f, ids_f = ctx.relation()
    rng = np.random.default_rng(5)
    n = max(1, int(frac * inst.n_tet))
    sub = rng.choice(inst.n_tet, n, replace=False).astype(np.int32)
    keep = np.zeros(inst.n_tet, bool)
    keep[sub] = True
    print(f"[subset] {n} of {inst.n_tet} tets", flush=True)
    ctx.set_tet_subset(sub)
    ctx.run()
    part = ctx.download()
    off_p, ids_p = ctx.relation()
    # relation rows: the full run's on subset tets, empty elsewhere
    cnt_p = np.diff(off_p)
    cnt_f = np.diff(off_f)
    assert np.array_equal(cnt_p[keep], cnt_f[keep])
    assert not cnt_p[~keep].any()
    a, b = _cells_of(full, keep), _cells_of(part, keep)
    assert len(b) == len(part) and a == b
    # the learned path does not take a subset
    ctx.mlp_init(0)
    with pytest.raises(RuntimeError, match="subset"):
        ctx.run(learned=1)
    # cleared: the full run again
    ctx.set_tet_subset(None)
    ctx.run()
    again = ctx.download()
    assert np.array_equal(again.words, full.words)
    assert np.array_equal(again.geom, full.geom)


def test_tet_subset_rejects_bad_ids(ctx):
    inst = prep.cube10_instance(0)
    ctx.upload(inst)
    with pytest.raises(RuntimeError, match="out of range"):
        ctx.set_tet_subset(np.array([0, inst.n_tet], np.int32))
    # a new upload clears a subset
    ctx.set_tet_subset(np.array([1, 2, 3], np.int32))
    ctx.upload(inst)
    ctx.run()
    orc = oracle.run(inst)
    compare_with_oracle(ctx.download(), orc)
