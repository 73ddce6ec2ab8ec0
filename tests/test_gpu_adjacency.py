"""GPU power-cell adjacency (rpd_adjacency, through the C-ABI) against the
oracle's restatement of get_all_voro_info and the component updates
(rpd_update.cxx:80-340, 409-521) and of cal_cell_euler /
cal_halfplane_facet_euler (voronoi_defs.cxx:190-267).

The oracle reads a cell file of its own cells (tests/cellfile.py); the
device cells of the same instance are the same cells in the same order, so
cell ids agree.  Compared exactly: cell neighbours, per-sphere cells and
components, facet and edge cell sets and components, and every Euler value
(binary32 per cell and per facet record, binary64 sums in ascending cell
order recomputed here from the oracle's values).
"""
import numpy as np
import pytest

import oracle
from cellfile import write_from_oracle
from libmat_amd import prep

pytestmark = pytest.mark.gpu


def _set(txt):
    txt = txt.strip()
    assert txt[0] == "{" and txt[-1] == "}", txt
    return [int(x) for x in txt[1:-1].split()]


def _groups(txt):
    """'{a b} {c}' -> [[a, b], [c]]"""
    out, cur = [], None
    for tok in txt.replace("{", " { ").replace("}", " } ").split():
        if tok == "{":
            cur = []
        elif tok == "}":
            out.append(cur)
            cur = None
        else:
            cur.append(int(tok))
    return out


def parse_full_dump(path):
    ce, fe, S = {}, {}, {}
    cur = None
    for line in open(path):
        w = line.split(None, 1)
        tag, rest = w[0], (w[1] if len(w) > 1 else "")
        if tag == "ce":
            c, v = rest.split()
            ce[int(c)] = float.fromhex(v)
        elif tag == "fe":
            c, j, v = rest.split()
            fe[(int(c), int(j))] = float.fromhex(v)
        elif tag == "S":
            cur = S.setdefault(int(rest), dict(nbr={}, facet={}, edge={},
                                               fcc={}, ecc={}))
        elif tag == "cells":
            cur["cells"] = _set(rest)
        elif tag == "nbr":
            c, st = rest.split(None, 1)
            cur["nbr"][int(c)] = _set(st)
        elif tag == "facet":
            j, st = rest.split(None, 1)
            cur["facet"][int(j)] = _set(st)
        elif tag == "edge":
            j, k, st = rest.split(None, 2)
            cur["edge"][(int(j), int(k))] = _set(st)
        elif tag == "cc":
            cur["cc"] = _groups(rest)
        elif tag == "fcc":
            j, st = rest.split(None, 1)
            cur["fcc"][int(j)] = _groups(st)
        elif tag == "ecc":
            j, k, st = rest.split(None, 2)
            cur["ecc"][(int(j), int(k))] = _groups(st)
    return ce, fe, S


def oracle_dump(tmp_path, inst, orc):
    cf, out = tmp_path / "cells.bin", tmp_path / "oracle_adj.txt"
    write_from_oracle(cf, inst, orc)
    oracle.adjacency(cf, None, inst.n_sf - 1, out)
    return parse_full_dump(out)


def label_of(groups):
    return {c: min(g) for g in groups for c in g}


def check_against_oracle(a, dump, n_site, cell0=0, site0=0):
    """a: rpd_adjacency output; cells / sites of the oracle's instance sit
    at cell0 / site0 of the device's stream (batch-wide ids)."""
    ce, fe, S = dump
    nc = len(ce)
    # per cell: Euler value, neighbours, component
    got_e = a["cell_euler"][cell0:cell0 + nc]
    assert np.array_equal(got_e, np.array([ce[c] for c in range(nc)],
                                          np.float32))
    nbr = a["cell_nbr"][cell0:cell0 + nc]
    for s, d in S.items():
        cells = d["cells"]
        for c in cells:
            g = sorted(int(x) - cell0 for x in nbr[c] if x >= 0)
            assert g == d["nbr"].get(c, []), (s, c)
        lab = label_of(d["cc"])
        for c in cells:
            assert a["cell_cc"][cell0 + c] == cell0 + lab[c], (s, c)
        ss = site0 + s
        assert a["site_cells"][ss] == len(cells)
        assert a["site_cc"][ss] == len(d["cc"])
        es = 0.0
        for c in sorted(cells):
            es += float(np.float32(ce[c]))
        assert a["site_euler"][ss] == es - len(cells), s
    # facets and edges of this instance's sites
    fac = {}
    for r in a["facets"]:
        s, j, b, n, ncc = (int(x) for x in r)
        if site0 <= s < site0 + n_site:
            fac[(s - site0, j - site0)] = (b, n, ncc)
    exp_fac = {(s, j): cs for s, d in S.items() for j, cs in d["facet"].items()}
    assert set(fac) == set(exp_fac)
    frow = {(int(a["facets"][q][0]), int(a["facets"][q][1])): q
            for q in range(len(a["facets"]))}
    for (s, j), cs in exp_fac.items():
        b, n, ncc = fac[(s, j)]
        rec = a["facet_cells"][b:b + n]
        assert [int(x) - cell0 for x in rec[:, 0]] == sorted(cs)
        groups = S[s]["fcc"][j]
        assert ncc == len(groups)
        lab = label_of(groups)
        assert [int(x) - cell0 for x in rec[:, 1]] == [lab[c] for c in
                                                       sorted(cs)]
        fev = a["facet_cell_euler"][b:b + n]
        assert np.array_equal(fev, np.array([fe[(c, j)] for c in sorted(cs)],
                                            np.float32)), (s, j)
        tot = 0.0
        for c in sorted(cs):
            tot += float(np.float32(fe[(c, j)]))
        assert a["facet_euler"][frow[(s + site0, j + site0)]] == tot
    edg = {}
    for r in a["edges"]:
        s, j, k, b, n, ncc = (int(x) for x in r)
        if site0 <= s < site0 + n_site:
            edg[(s - site0, j - site0, k - site0)] = (b, n, ncc)
    exp_edg = {(s, j, k): cs for s, d in S.items()
               for (j, k), cs in d["edge"].items()}
    assert set(edg) == set(exp_edg)
    for (s, j, k), cs in exp_edg.items():
        b, n, ncc = edg[(s, j, k)]
        rec = a["edge_cells"][b:b + n]
        assert [int(x) - cell0 for x in rec[:, 0]] == sorted(cs)
        groups = S[s]["ecc"][(j, k)]
        assert ncc == len(groups)
        lab = label_of(groups)
        assert [int(x) - cell0 for x in rec[:, 1]] == [lab[c] for c in
                                                       sorted(cs)]
    return len(exp_fac), len(exp_edg)


def _instances():
    yield "cube10 seed 0", prep.cube10_instance(0)
    yield "cube10 seed 3", prep.cube10_instance(3)
    verts, tets = prep.grid_mesh(8, 0.1, 7)
    yield "clustered", prep.build_instance(verts, tets,
                                           prep.spheres(1200, 2, 7))
    verts, tets = prep.grid_mesh(4, 0.0, 0)
    sph = [[x, y, z, 0.1] for x in (0.25, 0.75) for y in (0.25, 0.75)
           for z in (0.25, 0.75)]
    yield "octant", prep.build_instance(verts, tets, np.array(sph))


@pytest.mark.parametrize("case", [0, 1, 2, 3])
def test_adjacency_matches_oracle(ctx, tmp_path, case):
    name, inst = list(_instances())[case]
    orc = oracle.run(inst)
    print(f"[adjacency] {name} n_tet={inst.n_tet} n_site={inst.n_site} "
          f"cells={orc.cells.size}", flush=True)
    ctx.upload(inst)
    ctx.run()
    a = ctx.adjacency()
    assert a["sizes"]["n_cells"] == orc.cells.size
    dump = oracle_dump(tmp_path, inst, orc)
    nf, ne = check_against_oracle(a, dump, inst.n_site)
    print(f"[adjacency] {name}: facets {nf} edges {ne} pairs "
          f"{a['sizes']['n_cell_pairs']} identical", flush=True)
    if name.startswith("cube10"):
        # generic position: each restricted power cell is convex, so one
        # component and Euler number 1 (up to binary32 rounding)
        has = a["site_cells"] > 0
        assert (a["site_cc"][has] == 1).all()
        assert np.abs(a["site_euler"][has] - 1.0).max() < 1e-4


def test_adjacency_batch_equals_single(ctx, tmp_path):
    insts = [prep.cube10_instance(s) for s in (1, 2)]
    orcs = [oracle.run(i) for i in insts]
    _, so = ctx.upload_batch(insts)
    ctx.run()
    a = ctx.adjacency()
    cell0 = 0
    for b, (inst, orc) in enumerate(zip(insts, orcs)):
        d = oracle_dump(tmp_path, inst, orc)
        check_against_oracle(a, d, inst.n_site, cell0=cell0,
                             site0=int(so[b]))
        cell0 += orc.cells.size
    assert cell0 == a["sizes"]["n_cells"]


def test_adjacency_learned_equals_plain(ctx):
    inst = prep.cube10_instance(1)
    ctx.upload(inst)
    ctx.run()
    a = ctx.adjacency()
    ctx.mlp_init(0)
    ctx.run(learned=2, proposal_k=4)
    b = ctx.adjacency()
    for k in a:
        if k != "sizes":
            assert np.array_equal(a[k], b[k]), k


def test_adjacency_needs_a_run(ctx):
    inst = prep.cube10_instance(0)
    ctx.upload(inst)
    with pytest.raises(RuntimeError, match="no run"):
        ctx.adjacency()
