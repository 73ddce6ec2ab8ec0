"""Power-cell adjacency and Euler sums (CPU only): the product's
update_power_cells_adjacency / cal_cell_euler / cal_halfplane_facet_euler
(host C++ in the library) against the oracle's restatement of
get_all_voro_info and the connected-component updates
(rpd_update.cxx:80-340, 409-521; voronoi_defs.cxx:190-267), on the oracle's
cells of the 10^3 grid, and the Euler sums against what topology fixes."""
import os
import subprocess
from collections import defaultdict

import numpy as np
import pytest

import oracle
from cellfile import parse_dump, write_cube_es2fe, write_from_oracle
from libmat_amd import prep

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def adj_dump(tmp_path_factory):
    exe = tmp_path_factory.mktemp("adj") / "adj_dump"
    lib = os.path.join(ROOT, "libmat_amd", "_lib")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests/cpp/adj_dump.cpp"), "-o",
                    str(exe), "-L" + lib, "-lrpd_amd", "-Wl,-rpath," + lib],
                   check=True)
    return str(exe)


def run_both(adj_dump, tmp_path, cell_file, es2fe, max_surf_fid):
    prod, orc = tmp_path / "product.txt", tmp_path / "oracle.txt"
    r = subprocess.run([adj_dump, str(cell_file), str(es2fe) if es2fe else
                        "-", str(max_surf_fid), str(prod)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    oracle.adjacency(cell_file, es2fe, max_surf_fid, orc)
    a, b = open(prod).read().splitlines(), open(orc).read().splitlines()
    for q, (x, y) in enumerate(zip(a, b)):
        assert x == y, f"line {q}: product {x[:200]!r} oracle {y[:200]!r}"
    assert len(a) == len(b)
    return prod, a


def euler_invariants(dump, cells_by_q):
    """Per sphere: sum of cell Euler numbers minus the number of cells; per
    (sphere, neighbour) facet: sum of facet Euler numbers.  Both are 1 for a
    convex power cell restricted to the cube in general position (every
    vertex, edge and face of the cell complex is counted once through the
    1 / #adjacent weights).  A facet whose neighbour id equals the face id
    of an active tet face of one of its cells is left out: the facet lookup
    (voronoi_defs.cxx:234-244) matches ids without the -1 test and finds
    the tet face there."""
    ce, fe, by_sphere = parse_dump(dump)
    cell_sum = {s: sum(ce[c] for c in cs) - len(cs)
                for s, cs in by_sphere.items()}
    facet = defaultdict(float)
    skip = set()
    for (c, nb), v in fe.items():
        s = int(cells_by_q["seed"][c])
        facet[(s, nb)] += v
        if nb in cells_by_q["active_faces"][c]:
            skip.add((s, nb))
    return cell_sum, {k: v for k, v in facet.items() if k not in skip}, skip


def active_face_ids(inst, orc):
    out = []
    for c in orc.cells:
        nv, np_ = int(c["nb_v"]), int(c["nb_p"])
        used = set()
        for w in c["ver"][:nv]:
            used |= {int(w) & 255, (int(w) >> 8) & 255, (int(w) >> 16) & 255}
        ids = set()
        for p in range(np_):
            src = int(c["src"][p])
            if src < 0 and p in used:
                ids.add(int(inst.f_id[int(c["tet"]), -1 - src]))
        out.append(ids)
    return out


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_adjacency_matches_oracle(adj_dump, tmp_path, seed):
    inst = prep.cube10_instance(seed)
    orc = oracle.run(inst)
    cf, ef = tmp_path / "cells.bin", tmp_path / "es2fe.bin"
    write_from_oracle(cf, inst, orc)
    n_fe = write_cube_es2fe(ef, inst)
    prod, lines = run_both(adj_dump, tmp_path, cf, ef, inst.n_sf - 1)
    kinds = defaultdict(int)
    for x in lines:
        kinds[x.split()[0]] += 1
    print(f"[adjacency] seed {seed}: {len(lines)} identical lines, "
          f"{n_fe} feature-edge keys, " +
          ", ".join(f"{k} {v}" for k, v in sorted(kinds.items())))
    # every kind of entry is exercised
    for k in ("ce", "fe", "S", "nbr", "tfid", "surf", "facet", "edge",
              "e2v", "v2id", "v2pos", "se", "endpos", "cc", "fcc", "ecc"):
        assert kinds[k] > 0, k
    q = {"seed": orc.cells["seed"], "active_faces": active_face_ids(inst, orc)}
    cell_sum, facet_sum, skipped = euler_invariants(prod, q)
    ce_err = max(abs(v - 1.0) for v in cell_sum.values())
    fe_err = max(abs(v - 1.0) for v in facet_sum.values())
    print(f"[euler] {len(cell_sum)} spheres: max |sum - cells - 1| "
          f"{ce_err:.2e}; {len(facet_sum)} facets: max |sum - 1| "
          f"{fe_err:.2e} ({len(skipped)} facets with an id collision left "
          f"out)")
    assert ce_err < 1e-4 and fe_err < 1e-4


def test_adjacency_without_feature_edges(adj_dump, tmp_path):
    # no feature-edge map: no sharp/concave pieces, the rest unchanged
    inst = prep.cube10_instance(0)
    orc = oracle.run(inst)
    cf = tmp_path / "cells.bin"
    write_from_oracle(cf, inst, orc)
    _, lines = run_both(adj_dump, tmp_path, cf, None, inst.n_sf - 1)
    assert not any(x.startswith(("se ", "ce5 ", "endpos ")) for x in lines)


def slab_keep(inst, orc):
    # cells outside the slab 0.42 < x < 0.58 of tet centroids: spheres that
    # straddle it get cells in two (or more) connected components
    cx = inst.verts[inst.tets].mean(axis=1)[:, 0]
    x = cx[orc.cells["tet"]]
    return (x < 0.42) | (x > 0.58)


def test_adjacency_with_several_components(adj_dump, tmp_path):
    inst = prep.cube10_instance(0)
    orc = oracle.run(inst)
    cf, ef = tmp_path / "cells.bin", tmp_path / "es2fe.bin"
    write_from_oracle(cf, inst, orc, slab_keep(inst, orc))
    write_cube_es2fe(ef, inst)
    _, lines = run_both(adj_dump, tmp_path, cf, ef, inst.n_sf - 1)
    multi = [x for x in lines if x.startswith("cc ") and x.count("{") > 1]
    fmulti = [x for x in lines if x.startswith("fcc ") and x.count("{") > 1]
    print(f"[adjacency] slab removed: {len(multi)} spheres and "
          f"{len(fmulti)} facets with several components")
    assert multi and fmulti


@pytest.mark.parametrize("plant", ["euler_double", "face_walk",
                                   "cc_largest"])
def test_comparison_catches_a_mistake_planted_in_the_oracle(
        adj_dump, tmp_path, plant, monkeypatch):
    # binary64 vertex terms in the Euler sums (the reference divides in
    # binary32), face cycles walked from another start (the surface
    # centroids' binary32 sums change), components grown from the largest
    # element (their order changes): each must make the dumps differ
    inst = prep.cube10_instance(2)
    orc = oracle.run(inst)
    cf, ef = tmp_path / "cells.bin", tmp_path / "es2fe.bin"
    write_from_oracle(cf, inst, orc, slab_keep(inst, orc))
    write_cube_es2fe(ef, inst)
    prod, orcd = tmp_path / "p.txt", tmp_path / "o.txt"
    subprocess.run([adj_dump, str(cf), str(ef), str(inst.n_sf - 1),
                    str(prod)], check=True)
    monkeypatch.setenv("RPD_ORACLE_PLANT", plant)
    oracle.adjacency(cf, ef, inst.n_sf - 1, orcd)
    a, b = open(prod).read().splitlines(), open(orcd).read().splitlines()
    n_diff = sum(x != y for x, y in zip(a, b)) + abs(len(a) - len(b))
    print(f"[planted] {plant}: {n_diff} differing lines")
    assert n_diff > 0
