"""Host-side preparation and the C-ABI surface (CPU only)."""
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from libmat_amd import _native, prep
from libmat_amd.rpd import CAPI_SYMBOLS
from rpd_compare import tet_volumes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"^\s*(?:[\w\*]+\s+)+\**(rpd_\w+)\s*\(", txt,
                          flags=re.M))


def test_library_exports_every_declared_symbol():
    L = _native.lib()
    names = declared_functions("rpd_capi.h") | declared_functions("rpd_prep.h")
    assert len(names) >= 20
    for n in sorted(names):
        assert hasattr(L, n), n
    for n in CAPI_SYMBOLS:
        assert n in names


def test_cpp_dropin_layout(tmp_path):
    exe = tmp_path / "compat_test"
    lib = os.path.join(ROOT, "libmat_amd", "_lib")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests/cpp/compat_test.cpp"), "-o",
                    str(exe), "-L" + lib, "-lrpd_amd", "-Wl,-rpath," + lib],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True,
                         check=True).stdout
    assert "layout ok" in out


def test_core_library_leaves_cell_methods_to_the_program(tmp_path):
    """librpd_amd_core.so exports the RPD entry points but defines no
    ConvexCellHost method, so a program with its own voronoi_defs.cxx
    links it with one definition of each."""
    lib = os.path.join(ROOT, "libmat_amd", "_lib")
    core = os.path.join(lib, "librpd_amd_core.so")
    syms = subprocess.run(["nm", "-DC", "--defined-only", core],
                          capture_output=True, text=True, check=True).stdout
    assert "ConvexCellHost::" not in syms
    assert "is_convex_cell_valid" not in syms
    assert "compute_clipped_voro_diagram_GPU(" in syms
    assert "libmat_rpd::ResidentMesh::run(" in syms
    full = subprocess.run(["nm", "-DC", "--defined-only",
                           os.path.join(lib, "librpd_amd.so")],
                          capture_output=True, text=True, check=True).stdout
    assert "ConvexCellHost::reload_active()" in full
    exe = tmp_path / "core_link"
    subprocess.run(["g++", "-O1", "-std=c++17",
                    "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests/cpp/core_link.cpp"), "-o",
                    str(exe), "-L" + lib, "-lrpd_amd_core",
                    "-Wl,-rpath," + lib], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True,
                         check=True).stdout
    assert "core link ok 1 1" in out


def test_grid_mesh_is_conforming_and_positive():
    for n, jit in ((3, 0.0), (5, 0.1)):
        verts, tets = prep.grid_mesh(n, jit, 3)
        assert verts.shape == ((n + 1) ** 3, 3) and tets.shape == (6 * n ** 3, 4)
        V = verts[tets]
        det = np.einsum("ij,ij->i", V[:, 1] - V[:, 0],
                        np.cross(V[:, 2] - V[:, 0], V[:, 3] - V[:, 0]))
        assert (det > 0).all()
        np.testing.assert_allclose(tet_volumes(prep.Instance(
            verts, tets, *([None] * 5), None, None, None, None)).sum(), 1.0,
            rtol=1e-12)
        # boundary vertices stay on the cube
        b = (verts <= 0) | (verts >= 1)
        on = b.any(axis=1)
        g = np.round(verts[on] * n) / n
        assert np.abs(verts[on] - g).max() < 1e-15


def test_tet_adjacency_counts():
    n = 3
    verts, tets = prep.grid_mesh(n, 0.0, 0)
    v_adj, e_adj, f_adj, f_id, n_sf = prep.tet_adjacency(verts.shape[0], tets)
    assert v_adj.sum() == 4 * tets.shape[0]
    assert n_sf == 12 * n * n == (f_adj == 1).sum()
    # brute-force edge counts
    from collections import Counter
    el = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    cnt = Counter()
    for t in tets:
        for a, b in el:
            cnt[tuple(sorted((t[a], t[b])))] += 1
    for ti, t in enumerate(tets):
        for k, (a, b) in enumerate(el):
            assert e_adj[ti, k] == cnt[tuple(sorted((t[a], t[b])))]
    # face ids: boundary < n_sf, interior shared by exactly two tets
    ids, counts = np.unique(f_id, return_counts=True)
    assert (counts[ids < n_sf] == 1).all() and (counts[ids >= n_sf] == 2).all()
    assert ((f_id < n_sf) == (f_adj == 1)).all()


def test_power_neighbours_symmetric_sorted_and_hidden():
    sph = prep.spheres(300, 1, 4)
    xyzw = sph.copy()
    xyzw[:, 3] **= 2
    valid, off, ids = prep.power_neighbors(xyzw, prep.UNIT_BOX)
    assert valid.sum() < 300  # the boundary radius rule hides spheres
    pairs = set()
    for i in range(300):
        L = ids[off[i]:off[i + 1]]
        assert (np.diff(L) > 0).all()
        assert valid[i] or L.size == 0
        pairs |= {(i, int(j)) for j in L}
    assert all((j, i) in pairs for (i, j) in pairs)


@pytest.mark.timeout(120)
def test_power_neighbours_with_a_lone_weight_class():
    """A weight class of one sphere (here the one tiny sphere) has a zero
    extent; its search grid must still end.  Checked against the regular
    triangulation of the reference's own predicates when that builds, else
    against a dense clip of every sphere by all others."""
    rng = np.random.default_rng(5)
    xyzr = np.c_[rng.uniform(0.1, 0.9, (200, 3)), np.full(200, 0.05)]
    xyzr = np.r_[xyzr, [[0.5, 0.5, 0.5, 1e-5]]]
    xyzw = xyzr.copy()
    xyzw[:, 3] **= 2
    big = np.array([-100, -100, -100, 101, 101, 101.])
    valid, off, ids = prep.power_neighbors(xyzw, big)
    dense = np.array(xyzw)
    # dense: the sphere set with every sphere in one class (equal weights
    # up to 1e-12) gives the same neighbours when the weights are equal, so
    # compare the 200 equal spheres' lists with the lone sphere removed
    v2, off2, ids2 = prep.power_neighbors(dense[:200], big)
    for i in range(200):
        a = set(ids[off[i]:off[i + 1]]) - {200}
        b = set(ids2[off2[i]:off2[i + 1]])
        if 200 not in set(ids[off[i]:off[i + 1]]):
            assert a == b
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ref
    if ref.available():
        roff, rids = ref.rt_neighbors(xyzw)
        assert np.array_equal(np.diff(roff) > 0, valid > 0)
        for i in range(201):
            assert set(ids[off[i]:off[i + 1]]) <= \
                set(rids[roff[i]:roff[i + 1]])


def test_rng_is_deterministic():
    a = prep.spheres(10, 2, 9)
    b = prep.spheres(10, 2, 9)
    assert np.array_equal(a, b)
    assert (a[:, :3] > 0).all() and (a[:, :3] < 1).all()


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_gmm_workload_sizes_and_radius_rule():
    w = prep.gmm_workload(2, 3, tets=(6000, 20000), n_spheres=(500, 1500))
    assert len(w.instances) == 2
    for inst, xyzr, n in zip(w.instances, w.spheres, w.grid_n):
        assert inst.n_tet == 6 * n ** 3 and 500 <= len(xyzr) <= 1500
        c, r = xyzr[:, :3], xyzr[:, 3]
        bd = np.minimum(c, 1 - c).min(axis=1)
        assert (r > 0).all() and (r <= bd + 1e-15).all()
        assert ((c > 0) & (c < 1)).all()
