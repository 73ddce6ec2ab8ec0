"""Face export (CPU only): the product's get_one_convex_cell_faces_const and
save_convex_cells_bgeo (io_cuda.cxx:6-186) on the oracle's cells of the
10^3 grid, read back with a reader of the version-5 binary geometry format
written here from the format (big-endian header, points x y z w,
primitive attribute definitions, polygons of type 1 closed with '<', 16-bit
indices up to 65536 points, trailing 0x00 0xff).

What pins the export is independent of its code: each cell's polygons must
close into an oriented surface (every edge once in each direction, V - E +
F = 2), the divergence-theorem volume of that surface must match the
oracle's cell volume, each polygon's Newell area the oracle's facet area
(binary32 vertices: within 1e-3 of the tet's volume and face scale), the
triangle fans must tile the polygons, and the boundary file must hold exactly the bisector faces and
the surface tet faces.  Byte identity with the reference's own writer is
parity unpinned: it needs CUDA's vector types to build.
"""
import os
import struct
import subprocess
from collections import Counter

import numpy as np
import pytest

import oracle
from cellfile import write_from_oracle
from libmat_amd import prep

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def read_bgeo(path):
    b = open(path, "rb").read()
    o = 0

    def take(fmt):
        nonlocal o
        v = struct.unpack_from(">" + fmt, b, o)
        o += struct.calcsize(">" + fmt)
        return v
    magic, = take("I")
    assert magic == 0x4267656F
    ver, version = take("cI")
    assert ver == b"V" and version == 5
    n_pts, n_prim, npg, nprg, npa, nva, nra, nda = take("8I")
    assert (npg, nprg, npa, nva, nda) == (0, 0, 0, 0, 0) and nra == 1
    P = np.frombuffer(b, ">f4", 4 * n_pts, o).reshape(n_pts, 4).astype(
        np.float64)
    o += 16 * n_pts
    assert np.all(P[:, 3] == 1.0)
    ln, = take("H")
    name = b[o:o + ln].decode()
    o += ln
    assert name == "PrimAttr"
    size, typ, default = take("HiI")
    assert (size, typ, default) == (1, 1, 0)
    wide = n_pts > 65536
    polys, attr = [], []
    for _ in range(n_prim):
        kind, n = take("II")
        close, = take("c")
        assert kind == 1 and close == b"<"
        idx = take(("I" if wide else "H") * n)
        polys.append(idx)
        attr.append(take("i")[0])
    assert b[o:] == b"\x00\xff"
    return P[:, :3], polys, attr


@pytest.fixture(scope="module")
def exported(tmp_path_factory):
    d = tmp_path_factory.mktemp("geo")
    lib = os.path.join(ROOT, "libmat_amd", "_lib")
    exe = d / "geo_dump"
    subprocess.run(["g++", "-O2", "-std=c++17",
                    "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests/cpp/geo_dump.cpp"), "-o",
                    str(exe), "-L" + lib, "-lrpd_amd", "-Wl,-rpath," + lib],
                   check=True)
    inst = prep.cube10_instance(0)
    orc = oracle.run(inst)
    cf = d / "cells.bin"
    write_from_oracle(cf, inst, orc)
    max_sf = inst.n_sf - 1
    out = [d / "all.bgeo", d / "boundary.bgeo", d / "tri.txt"]
    subprocess.run([str(exe), str(cf), str(max_sf)] + [str(x) for x in out],
                   check=True, timeout=300)
    return inst, orc, max_sf, out


def _cells_faces(orc, polys, attr):
    """Split the polygon soup into cells (a cell's polygons are its active
    planes in plane order; its points follow the previous cells')."""
    cells = []
    q = 0
    for c in orc.cells:
        nv, np_ = int(c["nb_v"]), int(c["nb_p"])
        used = np.zeros(np_, bool)
        for v in c["ver"][:nv]:
            for k in range(3):
                used[(int(v) >> (8 * k)) & 255] = True
        k = int(used.sum())
        cells.append((np.nonzero(used)[0], polys[q:q + k], attr[q:q + k]))
        q += k
    assert q == len(polys)
    return cells


def test_bgeo_faces_close_and_match_the_oracle(exported):
    inst, orc, _, (all_path, _, _) = exported
    P, polys, attr = read_bgeo(all_path)
    assert P.shape[0] == int(orc.cells["nb_v"].sum())
    # binary32 vertices from binary32 planes lose digits on small cells, so
    # errors are measured against the tet's size
    T = inst.verts[inst.tets]
    tvol = np.abs(np.einsum("ij,ij->i", T[:, 1] - T[:, 0], np.cross(
        T[:, 2] - T[:, 0], T[:, 3] - T[:, 0]))) / 6
    row = 0
    vol_err, area_err, signs = 0.0, 0.0, set()
    for c, (planes, cp, ca) in zip(orc.cells, _cells_faces(orc, polys,
                                                           attr)):
        nv = int(c["nb_v"])
        assert all(a == int(c["seed"]) for a in ca)
        # an oriented closed surface over the cell's points
        edges = Counter()
        for f in cp:
            assert len(f) >= 3 and all(row <= i < row + nv for i in f)
            for a, b in zip(f, f[1:] + f[:1]):
                edges[(a, b)] += 1
        for (a, b), n in edges.items():
            assert n == 1 and edges[(b, a)] == 1
        assert nv - len(edges) // 2 + len(cp) == 2
        # volume (divergence theorem) and facet areas (Newell)
        V = P[row:row + nv]
        o = V[0]
        vol = 0.0
        for p, f in zip(planes, cp):
            X = P[list(f)]
            n = np.zeros(3)
            for a, b in zip(X, np.roll(X, -1, axis=0)):
                n += np.cross(a, b)
            area = 0.5 * np.linalg.norm(n)
            area_err = max(area_err, abs(area - c["area"][p]) /
                           tvol[int(c["tet"])] ** (2 / 3))
            for k in range(1, len(X) - 1):
                vol += np.dot(X[0] - o, np.cross(X[k] - o, X[k + 1] - o)) / 6
        signs.add(np.sign(vol))
        vol_err = max(vol_err, abs(abs(vol) - c["geo"][0]) /
                      tvol[int(c["tet"])])
        row += nv
    print(f"[export] cells {orc.cells.size} max volume error / tet volume "
          f"{vol_err:.2e}, max area error / tet volume^(2/3) {area_err:.2e}")
    assert len(signs) == 1  # every cell oriented the same way
    # (measured 7e-5 and 2e-4: binary32 vertices; a wrong face cycle or a
    # lost face moves these by a sizeable fraction of the tet)
    assert vol_err < 1e-3 and area_err < 1e-3


def test_triangle_fans_tile_the_polygons(exported):
    _, orc, _, (all_path, _, tri_path) = exported
    P, polys, attr = read_bgeo(all_path)
    tris = np.loadtxt(tri_path, dtype=np.int64).reshape(-1, 5)
    assert len(tris) == sum(len(f) - 2 for f in polys)
    q = 0
    for f, a in zip(polys, attr):
        for k in range(1, len(f) - 1):
            s, adj, i, j, l = tris[q]
            assert s == a and (i, j, l) == (f[0], f[k], f[k + 1])
            q += 1


def test_boundary_only_keeps_bisectors_and_surface_faces(exported):
    inst, orc, max_sf, (_, bnd_path, _) = exported
    _, polys, attr = read_bgeo(bnd_path)
    expect = 0
    for c in orc.cells:
        nv, np_ = int(c["nb_v"]), int(c["nb_p"])
        used = np.zeros(np_, bool)
        for v in c["ver"][:nv]:
            for k in range(3):
                used[(int(v) >> (8 * k)) & 255] = True
        for p in np.nonzero(used)[0]:
            src = int(c["src"][p])
            if src >= 0 or inst.f_id[int(c["tet"]), -1 - src] < max_sf:
                expect += 1
    assert len(polys) == expect
