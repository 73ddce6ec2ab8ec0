"""Test helper: cell files (the binary format of oracle/adj_oracle.cpp) and
the cube's feature-edge map, for the adjacency comparisons.

write_from_oracle() turns oracle cells into the records a LibMAT consumer
reads from ConvexCellHost: tet-face planes and power bisectors in binary32
from binary32 inputs (tri2plane, common_cuda.h:248-254; new_plane,
convex_cell.cu:561-592), h = the face's #adjacent tets (1 for bisectors),
plane ids (f_id, -1) or (min, max) of the two site ids.  Both sides of a
comparison read the same file, so these values only have to be
consistent with the cells.
"""
import struct

import numpy as np

F = np.float32
FACES = ((2, 1, 3), (0, 2, 3), (1, 0, 3), (0, 1, 2))
EDGES = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))


def _tet_plane(V, t, k, tets):
    a, b, c = (V[tets[t, i]] for i in FACES[k])
    u, v = b - a, c - a
    n = np.array([u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2],
                  u[0] * v[1] - u[1] * v[0]], F)
    d = -(a[0] * n[0] + a[1] * n[1] + a[2] * n[2])
    return n[0], n[1], n[2], d


def _bisector(S, s, m):
    sx, sy, sz, sw = S[s]
    bx, by, bz, bw = S[m]
    dx, dy, dz = sx - bx, sy - by, sz - bz
    ax, ay, az = sx + bx, sy + by, sz + bz
    dot = ax * dx + ay * dy + az * dz + (bw - sw)
    return dx, dy, dz, -dot / F(2)


def write_from_oracle(path, inst, orc, keep=None):
    """Cell file of the oracle's cells (those with keep[q] true, if given);
    voro_id = site id, id = position in the file."""
    cells = orc.cells if keep is None else orc.cells[np.asarray(keep, bool)]
    V = inst.verts.astype(F)
    S = np.concatenate([inst.sites[:, :3].astype(F),
                        (np.sqrt(inst.sites[:, 3]).astype(F) ** 2)[:, None]],
                       axis=1).astype(F)
    out = [b"RPDC", struct.pack("<ii", 1, cells.size)]
    for q, c in enumerate(cells):
        t, s = int(c["tet"]), int(c["seed"])
        nv, np_, ne = int(c["nb_v"]), int(c["nb_p"]), int(c["nb_e"])
        out.append(struct.pack("<8i", q, s, t, 0, 4, nv, np_, ne))
        out.append(struct.pack("<2f", S[s, 3], F(c["geo"][0])))
        out.append(c["ver"][:nv].astype("<u4").tobytes())
        planes, ids = [], []
        for p in range(np_):
            src = int(c["src"][p])
            if src < 0:
                k = -1 - src
                planes.append((*_tet_plane(V, t, k, inst.tets),
                               F(inst.f_adj[t, k])))
                ids.append((int(inst.f_id[t, k]), -1))
            else:
                planes.append((*_bisector(S, s, src), F(1)))
                ids.append((min(s, src), max(s, src)))
        out.append(np.array(planes, "<f4").tobytes())
        out.append(np.array(ids, "<i4").tobytes())
        e = c["edge"][:ne].astype("<u4")
        out.append(np.stack([e & 255, (e >> 8) & 255, (e >> 16) & 255],
                            axis=1).astype(np.uint8).tobytes())
    with open(path, "wb") as f:
        f.write(b"".join(out))


def write_cube_es2fe(path, inst):
    """The 12 edges of the unit cube as sharp feature edges (type 1): every
    tet edge on one maps (tet, local fid min, local fid max) -> (1, cube
    edge, cube edge); the local fids are the two faces holding the edge
    (face k is opposite vertex k)."""
    V = inst.verts
    on = lambda x: x == 0.0 or x == 1.0  # noqa: E731
    rows = []
    for t in range(inst.n_tet):
        for i, j in EDGES:
            a, b = V[inst.tets[t, i]], V[inst.tets[t, j]]
            ce = -1
            for ax in range(3):
                o1, o2 = (ax + 1) % 3, (ax + 2) % 3
                if on(a[o1]) and on(a[o2]) and a[o1] == b[o1] and \
                        a[o2] == b[o2]:
                    ce = ax * 4 + 2 * int(a[o1]) + int(a[o2])
                    break
            if ce < 0:
                continue
            f0, f1 = sorted(k for k in range(4) if k not in (i, j))
            rows.append((t, f0, f1, 1, ce, ce))
    with open(path, "wb") as f:
        f.write(struct.pack("<i", len(rows)))
        if rows:
            f.write(np.array(rows, "<i4").tobytes())
    return len(rows)


def parse_dump(path):
    """(cell euler {cell: value}, facet euler {(cell, neigh): value},
    sphere -> cells) from a canonical dump."""
    ce, fe, cells = {}, {}, {}
    sphere = None
    for line in open(path):
        w = line.split()
        if w[0] == "ce":
            ce[int(w[1])] = float.fromhex(w[2])
        elif w[0] == "fe":
            fe[(int(w[1]), int(w[2]))] = float.fromhex(w[3])
        elif w[0] == "S":
            sphere = int(w[1])
        elif w[0] == "cells":
            cells[sphere] = [int(x) for x in
                             " ".join(w[1:]).strip("{}").split()]
    return ce, fe, cells


def read_cells(path):
    """Records of a cell file: list of dicts."""
    buf = open(path, "rb").read()
    assert buf[:4] == b"RPDC"
    version, n = struct.unpack_from("<ii", buf, 4)
    assert version == 1
    o = 12
    out = []
    for _ in range(n):
        h = struct.unpack_from("<8i", buf, o)
        o += 32
        weight, vol = struct.unpack_from("<2f", buf, o)
        o += 8
        nv, np_, ne = h[5], h[6], h[7]
        ver = np.frombuffer(buf, "<u4", nv, o)
        o += 4 * nv
        clip = np.frombuffer(buf, "<f4", 5 * np_, o).reshape(np_, 5)
        o += 20 * np_
        id2 = np.frombuffer(buf, "<i4", 2 * np_, o).reshape(np_, 2)
        o += 8 * np_
        e = np.frombuffer(buf, np.uint8, 3 * ne, o).reshape(ne, 3)
        o += 3 * ne
        edge = e[:, 0].astype(np.uint32) | (e[:, 1].astype(np.uint32) << 8) \
            | (e[:, 2].astype(np.uint32) << 16)
        out.append(dict(id=h[0], voro_id=h[1], tet_id=h[2], thread_id=h[3],
                        status=h[4], nb_v=nv, nb_p=np_, nb_e=ne,
                        weight=weight, cell_vol=vol, ver=ver, clip=clip,
                        id2=id2, edge=edge))
    assert o == len(buf)
    return out


def dropin_instance(seed=0):
    """The instance RPD3D_GPU::calculate builds for the 10^3 grid and the
    500 spheres of ``seed`` (as tests/cpp/compat_test.cpp sets them up):
    binary32 vertices and sphere centres, weights (float r)^2, the valid
    spheres numbered in first-encounter order over the ascending map of
    sphere -> neighbour set (triangulation.cxx:442-549), neighbour lists in
    ascending site id.  Returns (instance, site -> sphere id)."""
    from libmat_amd import prep
    verts, tets = prep.grid_mesh(10, 0.0, seed)
    sph = prep.spheres(500, 0, seed, 0.02, 0.08)
    xyzw = sph.copy()
    xyzw[:, 3] = sph[:, 3] * sph[:, 3]
    valid, off, ids = prep.power_neighbors(xyzw, prep.UNIT_BOX)
    rings = {}
    for i in np.nonzero(valid)[0]:
        for j in ids[off[i]:off[i + 1]]:
            rings.setdefault(int(i), set()).add(int(j))
            rings.setdefault(int(j), set()).add(int(i))
    site_of, site2sphere = {}, []

    def number(x):
        if x not in site_of:
            site_of[x] = len(site2sphere)
            site2sphere.append(x)
        return site_of[x]

    nbrs = {}
    for k in sorted(rings):
        s0 = number(k)
        for nb in sorted(rings[k]):
            s1 = number(nb)
            nbrs.setdefault(s0, set()).add(s1)
            nbrs.setdefault(s1, set()).add(s0)
    n = len(site2sphere)
    m = np.array(site2sphere, np.int64)
    sites = np.empty((n, 4), np.float64)
    sites[:, :3] = sph[m, :3].astype(F)
    r = sph[m, 3].astype(F)
    sites[:, 3] = r * r
    knn_off = np.zeros(n + 1, np.int64)
    lists = [sorted(nbrs.get(s, ())) for s in range(n)]
    knn_off[1:] = np.cumsum([len(x) for x in lists])
    knn_ids = np.array([x for lst in lists for x in lst], np.int32)
    vf = verts.astype(F).astype(np.float64)
    v_adj, e_adj, f_adj, f_id, n_sf = prep.tet_adjacency(vf.shape[0], tets)
    inst = prep.Instance(vf, tets, v_adj, e_adj, f_adj, f_id, n_sf, sites,
                         np.ones(n, np.uint32), knn_off, knn_ids,
                         m.astype(np.int32))
    return inst, m
