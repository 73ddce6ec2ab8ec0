"""GPU parity: the HIP RPD against the oracle, through the C-ABI.

Combinatorics (per-tet candidate lists, per-cell vertex / plane / edge
arrays in the reference's own order) must be bit-exact; volumes, centroids
and facet areas must agree within 1e-6 relative (measured maxima printed).
"""
import numpy as np
import pytest

import oracle
from libmat_amd import prep
from rpd_compare import compare_relation, compare_with_oracle

pytestmark = pytest.mark.gpu

_cache = {}


def c1(seed):
    if seed not in _cache:
        inst = prep.cube10_instance(seed)
        _cache[seed] = (inst, oracle.run(inst, mode=0))
    return _cache[seed]


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_cube10_parity(ctx, seed):
    inst, orc = c1(seed)
    print(f"[parity] cube10 seed={seed} n_tet={inst.n_tet} "
          f"n_site={inst.n_site}", flush=True)
    ctx.upload(inst)
    st = ctx.run()
    off, ids = ctx.relation()
    compare_relation(off, ids, orc)
    res = ctx.download()
    m = compare_with_oracle(res, orc)
    print(f"[parity] seed={seed} {m} stats: cells={st['n_cells']} "
          f"cand={st['n_candidates']} active/all planes="
          f"{st['n_active_planes']}/{st['n_all_planes']} rerun="
          f"{st['n_rerun_large']}/{st['n_rerun_exact']} "
          f"ms={st['ms_total']:.3f}", flush=True)
    assert st["n_dropped"][6] == orc.drops[6]


def test_clustered_spheres_parity(ctx):
    # spheres from the 32-component Gaussian mixture (radii lognormal,
    # clipped to the boundary): dense clusters, many hidden spheres
    verts, tets = prep.grid_mesh(12, 0.1, 7)
    inst = prep.build_instance(verts, tets, prep.spheres(3000, 2, 7))
    orc = oracle.run(inst, mode=0)
    print(f"[parity] clustered n_tet={inst.n_tet} n_site={inst.n_site} "
          f"oracle cells={orc.cells.size}", flush=True)
    ctx.upload(inst)
    st = ctx.run()
    off, ids = ctx.relation()
    compare_relation(off, ids, orc)
    m = compare_with_oracle(ctx.download(), orc)
    print(f"[parity] clustered {m} cells={st['n_cells']} "
          f"rerun={st['n_rerun_large']}/{st['n_rerun_exact']}", flush=True)
    assert st["n_cells"] == orc.cells.size


@pytest.mark.parametrize("cfg", [
    dict(cells_per_wave=1), dict(cells_per_wave=2), dict(cells_per_wave=4),
    dict(cells_per_wave=8), dict(cells_per_wave=16),
    dict(use_active_mask=False), dict(small_tier=False),
    dict(force_exact=True), dict(cells_per_wave=4, small_caps=2),
    dict(cells_per_wave=8, small_caps=3), dict(cells_per_wave=16,
                                               small_caps=2)])
def test_cube10_variants_identical(ctx, cfg):
    inst, orc = c1(0)
    print(f"[variant] {cfg}", flush=True)
    ctx.upload(inst)
    st = ctx.run(**cfg)
    res = ctx.download()
    compare_with_oracle(res, orc)
    if cfg.get("force_exact"):
        assert st["n_rerun_exact"] > 0


def test_determinism(ctx):
    inst, _ = c1(1)
    ctx.upload(inst)
    ctx.run()
    a = ctx.download()
    ctx.run()
    b = ctx.download()
    for f in ("cells", "words", "geom", "areas"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_cube10_matches_golden_fixture(ctx, seed):
    import os
    from rpd_compare import digest_gpu
    g = np.load(os.path.join(os.path.dirname(__file__), "golden",
                             f"cube10_seed{seed}.npz"))
    inst = prep.cube10_instance(seed)
    ctx.upload(inst)
    ctx.run()
    off, ids = ctx.relation()
    assert np.array_equal(off, g["rel_off"]) and np.array_equal(ids,
                                                                g["rel_ids"])
    res = ctx.download()
    assert len(res) == int(g["n_cells"])
    assert digest_gpu(res) == str(g["cells_digest"])


def test_octant_spheres_exact_path_on_gpu(ctx):
    # bisectors on tet faces: vertex-on-plane determinants are exactly 0 and
    # the filtered predicate must hand them to the exact re-run
    verts, tets = prep.grid_mesh(4, 0.0, 0)
    sph = [[x, y, z, 0.1] for x in (0.25, 0.75) for y in (0.25, 0.75)
           for z in (0.25, 0.75)]
    inst = prep.build_instance(verts, tets, np.array(sph))
    orc = oracle.run(inst)
    ctx.upload(inst)
    st = ctx.run()
    print(f"[exact] reruns large/exact = {st['n_rerun_large']}/"
          f"{st['n_rerun_exact']}", flush=True)
    assert st["n_rerun_exact"] > 0
    compare_with_oracle(ctx.download(), orc)


def test_cube64_sample_and_partition(ctx):
    from libmat_amd.rpd import RPDResult
    from rpd_compare import tet_volumes
    inst = prep.cube64_instance(0, n_threads=16)
    print(f"[cube64] n_tet={inst.n_tet} n_site={inst.n_site}", flush=True)
    ctx.upload(inst)
    st = ctx.run()
    res = ctx.download()
    off, ids = ctx.relation()
    vol = np.zeros(inst.n_tet)
    np.add.at(vol, res.cells["tet"], res.geom[:, 0])
    tv = tet_volumes(inst)
    err = np.max(np.abs(vol - tv) / tv)
    print(f"[cube64] cells={st['n_cells']} partition max rel err {err:.2e}",
          flush=True)
    assert err < 1e-10
    sample = np.sort(np.random.default_rng(7).choice(
        inst.n_tet, 400, replace=False)).astype(np.int32)
    orc = oracle.run(inst, tets=sample, n_threads=16)
    for q, t in enumerate(sample):
        assert np.array_equal(ids[off[t]:off[t + 1]],
                              orc.rel_ids[orc.rel_off[q]:orc.rel_off[q + 1]])
    idx = np.nonzero(np.isin(res.cells["tet"], sample))[0]
    sub = RPDResult(res.cells[idx], res.words, res.geom[idx].reshape(-1),
                    res.areas)
    m = compare_with_oracle(sub, orc)
    print(f"[cube64] sample {m}", flush=True)


def compare_dropin_with_oracle(cells, inst, site2sphere, orc):
    """ConvexCellHost records of RPD3D_GPU::calculate against the oracle's
    cells of the same instance: order, ids, thread_id = tet * tet_k + slot,
    dual vertices, edges and plane ids bit-exact, planes by what they must
    satisfy in binary32, cell volumes to binary32 precision."""
    oc = orc.cells
    assert len(cells) == oc.size
    rel_n = np.diff(orc.rel_off)
    tet_k = int(rel_n.max())
    rng = np.random.default_rng(0)
    max_vol_err = 0.0
    for q, (c, o) in enumerate(zip(cells, oc)):
        t, s = int(o["tet"]), int(o["seed"])
        nv, np_, ne = int(o["nb_v"]), int(o["nb_p"]), int(o["nb_e"])
        assert c["id"] == q and c["tet_id"] == t, q
        assert c["voro_id"] == site2sphere[s], q
        assert c["status"] == 4, q
        slot = int(np.nonzero(orc.rel_ids[orc.rel_off[t]:orc.rel_off[t + 1]]
                              == s)[0][0])
        assert c["thread_id"] == t * tet_k + slot, q
        assert (c["nb_v"], c["nb_p"], c["nb_e"]) == (nv, np_, ne), q
        assert np.array_equal(c["ver"], o["ver"][:nv]), q
        assert np.array_equal(c["edge"], o["edge"][:ne] & 0xFFFFFF), q
        w = np.float32(inst.sites[s, 3])
        assert np.float32(c["weight"]) == w, q
        for p in range(np_):
            src = int(o["src"][p])
            a, b, cc, d, h = (float(x) for x in c["clip"][p])
            if src < 0:
                k = -1 - src
                assert tuple(c["id2"][p]) == (int(inst.f_id[t, k]), -1), q
                assert h == float(inst.f_adj[t, k])
                face = [inst.tets[t, i] for i in ((2, 1, 3), (0, 2, 3),
                                                  (1, 0, 3), (0, 1, 2))[k]]
                n = np.sqrt(a * a + b * b + cc * cc)
                for v in face:
                    x = inst.verts[v]
                    assert abs(a * x[0] + b * x[1] + cc * x[2] + d) <= \
                        1e-6 * n, (q, p)
                x = inst.verts[inst.tets[t, k]]
                assert a * x[0] + b * x[1] + cc * x[2] + d > 0, (q, p)
            else:
                lo, hi = min(s, src), max(s, src)
                assert tuple(c["id2"][p]) == (site2sphere[lo],
                                              site2sphere[hi]), q
                assert h == 1.0
                S, B = inst.sites[s], inst.sites[src]
                for x in rng.random((2, 3)):
                    pd_s = ((x - S[:3]) ** 2).sum() - S[3]
                    pd_b = ((x - B[:3]) ** 2).sum() - B[3]
                    val = a * x[0] + b * x[1] + cc * x[2] + d
                    assert abs(val - 0.5 * (pd_b - pd_s)) <= 1e-6, (q, p)
        vo = float(o["geo"][0])
        err = abs(c["cell_vol"] - vo) / max(abs(vo), 1e-30)
        max_vol_err = max(max_vol_err, err)
    assert max_vol_err < 1e-6
    return {"cells": len(cells), "tet_k": tet_k,
            "max_cell_vol_rel_err": max_vol_err}


def test_cpp_dropin_on_gpu(tmp_path):
    import os
    import subprocess
    from cellfile import dropin_instance, read_cells, write_cube_es2fe
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lib = os.path.join(root, "libmat_amd", "_lib")
    exe = tmp_path / "compat_test"
    subprocess.run(["g++", "-O2", "-std=c++17",
                    "-I" + os.path.join(root, "include"),
                    os.path.join(root, "tests/cpp/compat_test.cpp"), "-o",
                    str(exe), "-L" + lib, "-lrpd_amd", "-Wl,-rpath," + lib],
                   check=True)
    dump = tmp_path / "dropin_cells.bin"
    out = subprocess.run([str(exe), "--gpu", "--dump", str(dump)],
                         capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0 and "gpu ok" in out.stdout
    # the drop-in's cells against the oracle on the instance it builds
    inst, site2sphere = dropin_instance(0)
    orc = oracle.run(inst)
    cells = read_cells(dump)
    m = compare_dropin_with_oracle(cells, inst, site2sphere, orc)
    print(f"[dropin] {m}", flush=True)
    # and the adjacency the consumers read, on these very cells, against
    # the oracle's restatement of get_all_voro_info
    adj = tmp_path / "adj_dump"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-I" + os.path.join(root, "include"),
                    os.path.join(root, "tests/cpp/adj_dump.cpp"), "-o",
                    str(adj), "-L" + lib, "-lrpd_amd", "-Wl,-rpath," + lib],
                   check=True)
    ef = tmp_path / "es2fe.bin"
    write_cube_es2fe(ef, inst)
    pd_, od = tmp_path / "product.txt", tmp_path / "oracle.txt"
    subprocess.run([str(adj), str(dump), str(ef), str(inst.n_sf - 1),
                    str(pd_)], check=True, timeout=120)
    oracle.adjacency(dump, ef, inst.n_sf - 1, od)
    a, b = open(pd_).read(), open(od).read()
    print(f"[dropin] adjacency dump {len(a.splitlines())} lines, identical "
          f"{a == b}", flush=True)
    assert a == b


def test_work_counters_match_the_oracle(ctx):
    # without plane masks and with every cell in the full-capacity tier the
    # kernel evaluates exactly the oracle's conflict determinants, bisectors
    # and output vertices; masks only remove determinants
    inst = prep.cube10_instance(0)
    orc = oracle.run(inst)
    n_det, n_bis, _, n_vout = (int(x) for x in orc.counters)
    ctx.upload(inst)
    st = ctx.run(use_active_mask=False, small_tier=False, count_work=True)
    w = st["work"]
    print(f"[work] gpu {w} oracle det={n_det} bis={n_bis} vout={n_vout}",
          flush=True)
    assert w[0] == [0, 0, 0, 0] and st["n_rerun_exact"] == 0
    assert w[1][0] == n_det and w[1][1] == n_bis and w[1][2] == n_vout
    st2 = ctx.run(count_work=True)
    w2 = st2["work"]
    assert sum(x[2] for x in w2) == n_vout
    assert w2[0][0] + w2[1][0] < n_det
    assert 0 < st2["n_active_planes"] < st2["n_all_planes"]
    compare_with_oracle(ctx.download(), orc)


def _instance_slice(res, off, ids, to, so, b):
    """Cells and relation of instance b of a batched run, in local ids."""
    t0, t1 = int(to[b]), int(to[b + 1])
    roff = off[t0:t1 + 1] - off[t0]
    rids = ids[off[t0]:off[t1]] - so[b]
    sel = np.nonzero((res.cells["tet"] >= t0) & (res.cells["tet"] < t1))[0]
    cells = res.cells[sel]
    words, geom, areas = [], res.geom[sel], []
    for q in sel:
        c = res.cells[q]
        o, nv, np_, ne = (int(c["word_off"]), int(c["nb_v"]), int(c["nb_p"]),
                          int(c["nb_e"]))
        w = res.words[o:o + nv + np_ + ne].copy()
        src = w[nv:nv + np_].view(np.int32)
        src[src >= 0] -= so[b]
        words.append(w)
        a = int(c["area_off"])
        areas.append(res.areas[a:a + np_])
    return (roff, rids, cells["tet"] - t0, cells["seed"] - so[b],
            _counts(cells), words, geom, areas)


def _counts(cells):
    return np.stack([cells[f].astype(np.int64)
                     for f in ("nb_v", "nb_p", "nb_e", "status")])


def _single_arrays(res, off, ids):
    words, areas = [], []
    for q in range(len(res)):
        c = res.cells[q]
        o, n = int(c["word_off"]), int(c["nb_v"]) + int(c["nb_p"]) + int(c["nb_e"])
        words.append(res.words[o:o + n])
        a = int(c["area_off"])
        areas.append(res.areas[a:a + int(c["nb_p"])])
    return (off, ids, res.cells["tet"], res.cells["seed"],
            _counts(res.cells), words, res.geom, areas)


def _same(a, b):
    for x, y in zip(a, b):
        if isinstance(x, list):
            assert len(x) == len(y)
            for u, v in zip(x, y):
                assert np.array_equal(u, v)
        else:
            assert np.array_equal(x, y)


def test_batch_matches_single_instance_runs(ctx):
    from dataclasses import replace
    from libmat_amd.rpd import Context
    w = prep.batch_workload(10, 3, (500, 6000), (50, 800))
    to, so = ctx.upload_batch(w.instances)
    print(f"[batch] tets {np.diff(to).tolist()} sites {np.diff(so).tolist()}",
          flush=True)
    single = Context(0)
    insts = list(w.instances)
    for it in range(2):
        st = ctx.run()
        res = ctx.download()
        off, ids = ctx.relation()
        assert st["n_cells"] == len(res)
        for b, inst in enumerate(insts):
            single.upload(inst)
            single.run()
            ra = _single_arrays(single.download(), *single.relation())
            _same(_instance_slice(res, off, ids, to, so, b), ra)
        # two instances against the oracle as well
        from test_gpu_configs import cells_of, subset
        for b in (0, len(insts) - 1):
            t0, t1, s0 = int(to[b]), int(to[b + 1]), int(so[b])
            orc = oracle.run(insts[b])
            assert np.array_equal(off[t0:t1 + 1] - off[t0], orc.rel_off)
            assert np.array_equal(ids[off[t0]:off[t1]] - s0, orc.rel_ids)
            sel = cells_of(res, np.arange(t0, t1))
            compare_with_oracle(subset(res, sel, t0, s0), orc)
        print(f"[batch] iteration {it}: {len(res)} cells identical to the "
              f"single-instance runs", flush=True)
        # next sphere-update iteration: centres jittered, host neighbours
        sets, new = [], []
        for b, inst in enumerate(insts):
            xyzr = prep.jitter_spheres(w.spheres[b], 0.002, 1000 + b)
            sites, flags, koff, kids, keep = prep.sites_of(xyzr)
            sets.append((sites, flags, koff, kids))
            new.append(replace(inst, sites=sites, flags=flags, knn_off=koff,
                               knn_ids=kids, site2sphere=keep))
        ctx.update_sites_batch(sets)
        to, so = ctx.batch_offsets()
        insts = new
    single.close()


def test_tune_picks_a_shape_with_identical_results(ctx):
    inst, orc = c1(2)
    ctx.upload(inst)
    G, caps, ms = ctx.tune()
    print(f"[tune] G={G} caps={caps} ms={ms}", flush=True)
    assert G in (4, 8, 16) and caps in (0, 1, 2, 3)
    assert all(v > 0 for v in ms.values()) and len(ms) == 12
    ctx.run(cells_per_wave=0)
    compare_with_oracle(ctx.download(), orc)
    for c in range(4):
        st = ctx.run(cells_per_wave=16, small_caps=c)
        print(f"[tune] caps={c}: capacity re-runs {st['n_rerun_large']}",
              flush=True)
        compare_with_oracle(ctx.download(), orc)


@pytest.mark.parametrize("case", ["uniform500", "boundary20k", "gmm5k",
                                  "boundary50k"])
def test_device_power_neighbours_match_host(ctx, case):
    # the device builder (rpd_power_neighbors_gpu) against the host builder:
    # same exact predicates, so the same validity and the same lists
    n, rule, seed = {"uniform500": (500, 0, 0), "boundary20k": (20000, 1, 3),
                     "gmm5k": (5000, 2, 4),
                     "boundary50k": (50000, 1, 0)}[case]
    xyzr = prep.spheres(n, rule, seed)
    xyzw = xyzr.copy()
    xyzw[:, 3] **= 2
    v_h, off_h, ids_h = prep.power_neighbors(xyzw, prep.UNIT_BOX)
    v_d, off_d, ids_d, ms = ctx.power_neighbors(xyzw, prep.UNIT_BOX)
    print(f"[neighbours] {case}: valid {int(v_d.sum())} of {n}, pairs "
          f"{ids_d.size}, device clip {ms:.2f} ms", flush=True)
    assert np.array_equal(v_h > 0, v_d > 0)
    assert np.array_equal(off_h, off_d)
    assert np.array_equal(ids_h, ids_d)
