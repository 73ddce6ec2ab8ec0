"""GPU: the pipeline of a resident batch captured into a hipGraph
(rpd_graph_build / rpd_graph_run, through the C-ABI), plain and learned
(certificate, closure rounds, fallback and output order on the device).
Every replay must give bit for bit the cells (and the relation CSR) of an
eager rpd_run on the same data: after the capture, after sphere updates
that keep each instance's site count (a replay), and after updates that
change it or outgrow the capacities (a new capture, whose eager run gives
the results).
"""
import numpy as np
import pytest

from libmat_amd import prep

pytestmark = pytest.mark.gpu


def _same(a, b):
    for f in ("cells", "words", "geom", "areas"):
        x, y = getattr(a, f), getattr(b, f)
        assert x.shape == y.shape and np.array_equal(x, y), f


def _eager(ctx, sets, **kw):
    ctx.update_sites_batch(sets)
    st = ctx.run(**kw)
    return st, ctx.download()


def _sets(w):
    return [(i.sites, i.flags, i.knn_off, i.knn_ids) for i in w.instances]


def _jittered(w, it, keep_counts):
    """Sites of every instance after a centre jitter; with keep_counts the
    instances whose valid-sphere count changed keep their old sites."""
    out = []
    for inst, x, b in zip(w.instances, w.spheres, range(len(w.spheres))):
        y = prep.jitter_spheres(x, 0.002, 7919 * it + b)
        s = prep.sites_of(y)[:4]
        if keep_counts and s[0].shape[0] != inst.n_site:
            s = (inst.sites, inst.flags, inst.knn_off, inst.knn_ids)
        out.append(s)
    return out


@pytest.fixture(scope="module")
def workload():
    return prep.batch_workload(6, 3, tets=(2000, 20000), n_spheres=(100, 1500))


@pytest.mark.parametrize("small_caps,G", [(0, 8), (2, 16), (3, 8)])
def test_graph_replays_equal_eager_runs(ctx, workload, small_caps, G):
    w = workload
    kw = dict(cells_per_wave=G, small_caps=small_caps)
    ctx.upload_batch(w.instances)
    base = _sets(w)
    st0, ref0 = _eager(ctx, base, **kw)
    print(f"[graph] caps={small_caps} G={G} cells={st0['n_cells']} "
          f"reruns={st0['n_rerun_large']}/{st0['n_rerun_exact']}", flush=True)
    ctx.graph_build(headroom=0.2, **kw)
    _same(ctx.download(), ref0)
    st, rep = ctx.graph_run()
    assert rep
    assert st["n_cells"] == st0["n_cells"]
    assert st["n_candidates"] == st0["n_candidates"]
    assert st["tet_k"] == st0["tet_k"]
    _same(ctx.download(), ref0)
    # sphere updates that keep the per-instance site counts: replays
    replays = 0
    for it in (1, 2):
        sets = _jittered(w, it, keep_counts=True)
        st_e, ref = _eager(ctx, sets, **kw)
        ctx.update_sites_batch(sets)
        st, rep = ctx.graph_run()
        replays += rep
        print(f"[graph] iteration {it}: replayed {rep} cells "
              f"{st['n_cells']} ms {st['ms_total']:.3f}", flush=True)
        assert st["n_cells"] == st_e["n_cells"]
        _same(ctx.download(), ref)
    assert replays == 2
    ctx.graph_destroy()


def test_graph_recaptures_when_counts_change(ctx, workload):
    w = workload
    ctx.upload_batch(w.instances)
    base = _sets(w)
    ctx.update_sites_batch(base)
    ctx.graph_build(headroom=0.0)
    # every instance with new sites (counts may change): a new capture or a
    # replay, the cells are the eager run's either way
    sets = _jittered(w, 5, keep_counts=False)
    st_e, ref = _eager(ctx, sets)
    ctx.update_sites_batch(sets)
    st, rep = ctx.graph_run()
    print(f"[graph] changed counts: replayed {rep}", flush=True)
    _same(ctx.download(), ref)
    # larger spheres: more candidates and cells than the zero-headroom
    # capture holds -- the overflow is caught and the batch captured again
    big = []
    for x in w.spheres:
        y = x.copy()
        y[:, 3] *= 1.5
        big.append(prep.sites_of(y)[:4])
    st_e, ref = _eager(ctx, big)
    ctx.update_sites_batch(big)
    st, rep = ctx.graph_run()
    print(f"[graph] larger spheres: replayed {rep} cells {st['n_cells']}",
          flush=True)
    assert st["n_cells"] == st_e["n_cells"]
    _same(ctx.download(), ref)
    # and the recaptured graph replays
    st, rep = ctx.graph_run()
    assert rep
    _same(ctx.download(), ref)
    ctx.graph_destroy()


def test_graph_with_a_tet_subset(ctx):
    inst = prep.cube10_instance(1)
    ctx.upload(inst)
    sub = np.arange(0, inst.n_tet, 3, dtype=np.int32)
    ctx.set_tet_subset(sub)
    ctx.run()
    ref = ctx.download()
    ctx.graph_build(headroom=0.1)
    st, rep = ctx.graph_run()
    assert rep
    _same(ctx.download(), ref)
    ctx.set_tet_subset(None)  # the switch changed: captured again
    ctx.run()
    full = ctx.download()
    st, rep = ctx.graph_run()
    assert not rep
    _same(ctx.download(), full)
    ctx.graph_destroy()


def _same_relation(ctx, ref_rel):
    off, ids = ctx.relation()
    assert np.array_equal(off, ref_rel[0]) and np.array_equal(ids, ref_rel[1])


def _every(w, it):
    """Every sphere a site (hidden ones unflagged, with empty lists), after
    jitter iteration it (0: the original spheres): site counts stay fixed,
    the neighbour lists' total length changes."""
    out = []
    for b, x in enumerate(w.spheres):
        y = x if it == 0 else prep.jitter_spheres(x, 0.002, 7919 * it + b)
        out.append(prep.sites_both(y)[1])
    return out


@pytest.mark.parametrize("mode,sites", [("topk", "visible"),
                                        ("model", "visible"),
                                        ("model", "every")])
def test_learned_graph_replays_equal_eager_runs(ctx, workload, mode, sites):
    w = workload
    ctx.upload_batch(w.instances)
    if mode == "model":
        ctx.mlp_init(3)
        kw = dict(learned=1, logit_threshold=0.0)
    else:
        kw = dict(learned=2, proposal_k=4)
    base = _sets(w) if sites == "visible" else _every(w, 0)
    st0, ref0 = _eager(ctx, base, **kw)
    rel0 = ctx.relation()
    # the plain path: the same cells
    _, plain = _eager(ctx, base)
    _same(plain, ref0)
    print(f"[graph] learned {mode}: cells={st0['n_cells']} "
          f"learned={st0['learned']}", flush=True)
    ctx.graph_build(headroom=0.2, **kw)
    _same(ctx.download(), ref0)
    _same_relation(ctx, rel0)
    st, rep = ctx.graph_run()
    assert rep
    assert st["n_cells"] == st0["n_cells"]
    assert st["n_host_reads"] == 1
    assert st["learned"][0] == st0["learned"][0]
    _same(ctx.download(), ref0)
    _same_relation(ctx, rel0)
    # sphere updates with eager learned runs in between (they swap the
    # offset buffers the graph writes): replays
    replays = 0
    for it in (1, 2):
        sets = (_jittered(w, it, keep_counts=True) if sites == "visible"
                else _every(w, it))
        st_e, ref = _eager(ctx, sets, **kw)
        rel = ctx.relation()
        ctx.update_sites_batch(sets)
        st, rep = ctx.graph_run()
        replays += rep
        why = "" if rep else ctx._L.rpd_last_error(ctx._h).decode()
        print(f"[graph] learned {mode} iteration {it}: replayed {rep} cells "
              f"{st['n_cells']} learned {st['learned']} ms "
              f"{st['ms_total']:.3f} {why}", flush=True)
        assert st["n_cells"] == st_e["n_cells"]
        _same(ctx.download(), ref)
        _same_relation(ctx, rel)
    assert replays == 2
    ctx.graph_destroy()


def test_learned_graph_recaptures_when_counts_grow(ctx, workload):
    w = workload
    ctx.upload_batch(w.instances)
    kw = dict(learned=2, proposal_k=2)  # few proposals: closure rounds
    ctx.update_sites_batch(_sets(w))
    ctx.graph_build(headroom=0.0, **kw)
    big = []
    for x in w.spheres:
        y = x.copy()
        y[:, 3] *= 1.5
        big.append(prep.sites_of(y)[:4])
    st_e, ref = _eager(ctx, big, **kw)
    rel = ctx.relation()
    ctx.update_sites_batch(big)
    st, rep = ctx.graph_run()
    print(f"[graph] learned, larger spheres: replayed {rep}", flush=True)
    assert st["n_cells"] == st_e["n_cells"]
    _same(ctx.download(), ref)
    _same_relation(ctx, rel)
    st, rep = ctx.graph_run()  # the new capture replays
    assert rep
    _same(ctx.download(), ref)
    ctx.graph_destroy()


def test_graph_rejects_other_paths(ctx):
    inst = prep.cube10_instance(0)
    ctx.upload(inst)
    ctx.graph_destroy()
    with pytest.raises(RuntimeError, match="no graph"):
        ctx.graph_run()
    from libmat_amd.rpd import RunCfg
    import ctypes
    cfg = RunCfg(8, 1, 1, 0, 1, 1, 0, 0, 0.0, 8)  # work counters
    rc = ctx._L.rpd_graph_build(ctx._h, ctypes.byref(cfg), 0.1)
    assert rc != 0
    assert "work counters" in ctx._L.rpd_last_error(ctx._h).decode()
