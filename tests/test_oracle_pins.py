"""Pins of the oracle (CPU only).

The reference ships no tests or golden data, so the oracle is pinned by:
exact rational arithmetic (fractions.Fraction) for the determinant signs,
closed-form diagrams (one sphere, two spheres, eight octant spheres), the
brute-force definition of the restricted power diagram (clip by every
sphere), the partition of each tet by its cells, and the reference's own
CPU sources compiled in place (geogram PSM exact determinant and regular
triangulation, VBW FP64 convex cell) when the reference is mounted.
"""
import os
import random
import subprocess
from fractions import Fraction

import numpy as np
import pytest

import oracle
from libmat_amd import prep
from rpd_compare import digest_oracle, tet_volumes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def frac_det4(m):
    a = [[Fraction(float(x)) for x in row] for row in np.asarray(m).reshape(4, 4)]

    def det(M):
        if len(M) == 1:
            return M[0][0]
        s = Fraction(0)
        for c in range(len(M)):
            minor = [r[:c] + r[c + 1:] for r in M[1:]]
            s += (-1) ** c * M[0][c] * det(minor)
        return s

    d = det(a)
    return (d > 0) - (d < 0)


def det_cases(n=400, seed=0):
    rng = random.Random(seed)
    out = []
    for k in range(n):
        kind = k % 4
        if kind == 0:  # random
            m = [rng.uniform(-1, 1) * 10 ** rng.randint(-3, 2)
                 for _ in range(16)]
        elif kind == 1:  # exactly singular: col3 = col0 + col1 (small ints)
            cols = [[float(rng.randint(-8, 8)) for _ in range(4)]
                    for _ in range(3)]
            cols.append([cols[0][r] + cols[1][r] for r in range(4)])
            m = [cols[c][r] for r in range(4) for c in range(4)]
        elif kind == 2:  # nearly singular: perturb one entry by one ulp
            cols = [[rng.uniform(-1, 1) for _ in range(4)] for _ in range(3)]
            cols.append([cols[0][r] * 0.5 + cols[1][r] * 0.25
                         for r in range(4)])
            m = [cols[c][r] for r in range(4) for c in range(4)]
            i = rng.randrange(16)
            m[i] = np.nextafter(m[i], 2.0)
        else:  # wide exponent range within rows
            m = [rng.uniform(-1, 1) * 2.0 ** rng.randint(-60, 60)
                 for _ in range(16)]
        out.append(np.array(m, np.float64))
    return out


def test_exact_det4_sign_matches_rational_arithmetic():
    for m in det_cases():
        assert oracle.exact_det4_sign(m) == frac_det4(m)


def test_product_exact_predicates_match_rational_arithmetic(tmp_path):
    exe = tmp_path / "pred_test"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                    os.path.join(ROOT, "tests/cpp/pred_test.cpp"), "-o",
                    str(exe)], check=True)
    cases = det_cases(seed=1)
    inp = "\n".join(" ".join(float(x).hex() for x in m) for m in cases)
    out = subprocess.run([str(exe)], input=inp, capture_output=True,
                         text=True, check=True).stdout.split("\n")
    n_filtered = 0
    for m, line in zip(cases, out):
        x, f, w = map(int, line.split())
        s = frac_det4(m)
        assert x == s
        assert f in (0, s)  # the filter is never wrong, only undecided
        n_filtered += f != 0
        a = np.asarray(m).reshape(4, 4)
        sw = frac_det4(np.block([[a[:3, :3], np.zeros((3, 1))],
                                 [np.zeros((1, 3)), np.ones((1, 1))]]))
        assert w == sw
    assert n_filtered > len(cases) // 8


def plane_cases(n_tet, seed):
    """tets of many sizes and offsets (every 7th a sliver, every 11th on a
    1/64 grid) and 24 planes each, at signed distances 0 to 10 tet sizes
    from the tet's first or last corner along random and axis normals of
    scales 1e-6 to 1e6 (every 17th plane of scale 1e-36 or 1e36, outside
    the binary32 table's range); returns the rows and the distances in tet
    sizes"""
    rng = np.random.default_rng(seed)
    deltas = [0.0, 1e-16, 1e-15, 1e-14, 1e-12, 1e-9, 1e-6, 1e-3, 0.03, 0.1,
              0.3, 1.0, 3.0, 10.0]
    rows, dist = [], []
    for it in range(n_tet):
        size = 10.0 ** rng.uniform(-4, 0)
        off = 0.0 if it % 3 == 0 else 10.0 ** rng.uniform(-2, 3) * rng.uniform(-1, 1)
        V = off * np.array([1.0, 1.0, 0.5]) + size * rng.uniform(-1, 1, (4, 3))
        if it % 7 == 1:
            t = 10.0 ** rng.uniform(-15, -2)
            a = rng.uniform()
            b = rng.uniform() * (1 - a)
            V[3] = (V[0] + a * (V[1] - V[0]) + b * (V[2] - V[0]) +
                    t * size * rng.uniform(-1, 1, 3))
        if it % 11 == 2:
            V = np.round(V * 64) / 64
        for j in range(24):
            n = rng.uniform(-1, 1, 3)
            if j % 5 == 0:
                n = np.array([0.0, 0.0, 1.0])
            elif j % 5 == 1:
                n = np.array([1.0, -1.0, 0.0])
            sc = 10.0 ** rng.uniform(-6, 6)
            if (it * 24 + j) % 17 == 0:
                sc = 1e-36 if j % 2 else 1e36
            n = n * sc
            # offset from the corner of largest (or least) n.V, so that
            # grazing planes occur on both sides of both orientations
            hi = rng.uniform() < 0.5
            ext = (max if hi else min)(float(n @ v) for v in V)
            nn = float(np.linalg.norm(n))
            d = deltas[j % 14] * size * nn
            d = d if rng.uniform() < 0.5 else -d
            rows.append(list(V.ravel()) + list(n) + [-(ext + (d if hi else -d))])
            dist.append(abs(d) / (size * nn))
    return rows, dist


def test_binary32_plane_masks_are_sound(tmp_path):
    """The plane masks' test (rpd_mask_negative on the binary32 bisector
    entry rpd_mask_coef): every plane it skips is negative at all four
    corners by the exact predicate on the binary64 plane, and it skips
    nearly all such planes that lie three tet sizes or more beyond a tet
    that is not a sliver and lies within a thousand of its sizes of the
    origin."""
    exe = tmp_path / "mask_test"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wall",
                    "-Werror", os.path.join(ROOT, "tests/cpp/mask_test.cpp"),
                    "-o", str(exe)], check=True)
    rows, dist = plane_cases(3000, seed=5)
    inp = "\n".join(" ".join(float(x).hex() for x in r) for r in rows)
    out = subprocess.run([str(exe)], input=inp, capture_output=True,
                         text=True, check=True).stdout.split("\n")
    assert len(out) >= len(rows)
    n_far = n_far_mask = n_mask = 0
    for i, (r, d, line) in enumerate(zip(rows, dist, out)):
        mask, *x = map(int, line.split())
        if mask:
            assert all(s < 0 for s in x), line
            n_mask += 1
        # efficacy over the tets that are not slivers and lie within 1e3 of
        # their sizes of the origin (binary32 planes resolve a tet size
        # there), planes within the table's range
        V = np.asarray(r[:12]).reshape(4, 3)
        ext = float((V.max(axis=0) - V.min(axis=0)).max())
        scale = max(abs(v) for v in r[12:15])
        if (d >= 3.0 and all(s < 0 for s in x) and 1e-20 < scale < 1e20 and
                (i // 24) % 7 != 1 and np.abs(V).max() <= 1e3 * ext):
            n_far += 1
            n_far_mask += mask
    assert n_far > 1000 and n_far_mask > 0.95 * n_far
    assert n_mask > 1000


@pytest.mark.skipif(not __import__("ref").available(),
                    reason="reference sources not mounted (oracle/_ref)")
def test_exact_det4_sign_matches_reference_psm():
    import ref
    for m in det_cases(seed=2):
        assert oracle.exact_det4_sign(m) == ref.det4_sign(m)


def instance(n, spheres):
    verts, tets = prep.grid_mesh(n, 0.0, 0)
    return prep.build_instance(verts, tets, np.asarray(spheres, np.float64))


def per_seed(orc, inst, key):
    out = {}
    for c in orc.cells:
        out.setdefault(int(inst.site2sphere[c["seed"]]), []).append(c)
    return out


def test_one_sphere_owns_every_tet():
    inst = instance(3, [[0.4, 0.5, 0.6, 0.1]])
    orc = oracle.run(inst)
    assert orc.cells.size == inst.n_tet
    np.testing.assert_allclose(orc.cells["geo"][:, 0], tet_volumes(inst),
                               rtol=1e-13)
    assert (orc.cells["nb_v"] == 4).all() and (orc.cells["nb_p"] == 4).all()


def test_two_spheres_split_at_the_power_bisector():
    # (x-a)^2 - w1 = (x-b)^2 - w2  =>  x0 = (a+b)/2 + (w1-w2)/(2(b-a))
    a, b, r1, r2 = 0.25, 0.75, 0.1, 0.2
    x0 = 0.5 * (a + b) + (r1 ** 2 - r2 ** 2) / (2 * (b - a))
    inst = instance(4, [[a, 0.5, 0.5, r1], [b, 0.5, 0.5, r2]])
    orc = oracle.run(inst)
    c = orc.cells
    v = [c["geo"][c["seed"] == s, 0].sum() for s in (0, 1)]
    np.testing.assert_allclose(v, [x0, 1 - x0], rtol=1e-12)
    m0 = (c["geo"][c["seed"] == 0, 0:1] * c["geo"][c["seed"] == 0, 1:]).sum(0)
    np.testing.assert_allclose(m0 / v[0], [x0 / 2, 0.5, 0.5], atol=1e-12)


def test_octant_spheres_on_grid_planes_exact_path():
    # bisectors x, y, z = 0.5 coincide with tet faces of the 4^3 grid: every
    # vertex on them gives det == 0 and needs the exact fallback
    sph = [[x, y, z, 0.1] for x in (0.25, 0.75) for y in (0.25, 0.75)
           for z in (0.25, 0.75)]
    inst = instance(4, sph)
    orc = oracle.run(inst)
    assert orc.counters[2] > 0  # exact evaluations happened
    c = orc.cells
    for s in range(8):
        sel = c["seed"] == s
        vol = c["geo"][sel, 0].sum()
        cen = (c["geo"][sel, 0:1] * c["geo"][sel, 1:]).sum(0) / vol
        np.testing.assert_allclose(vol, 0.125, rtol=1e-12)
        np.testing.assert_allclose(cen, np.array(sph[s][:3]), atol=1e-12)


def test_knn_clipping_equals_brute_force_clipping():
    # the restricted power diagram by definition: each cell clipped by every
    # other sphere; the neighbour-list path must give the same non-empty
    # (tet, sphere) set and volumes
    verts, tets = prep.grid_mesh(3, 0.1, 5)
    inst = prep.build_instance(verts, tets, prep.spheres(40, 0, 7))
    a = oracle.run(inst)
    b = oracle.run(inst, brute_force=True)
    ka = list(zip(a.cells["tet"], a.cells["seed"]))
    kb = list(zip(b.cells["tet"], b.cells["seed"]))
    assert ka == kb
    np.testing.assert_allclose(a.cells["geo"], b.cells["geo"], rtol=1e-9,
                               atol=1e-15)


@pytest.mark.parametrize("seed", [0, 1])
def test_partition_of_every_tet(seed):
    inst = prep.cube10_instance(seed)
    orc = oracle.run(inst)
    vol = np.zeros(inst.n_tet)
    np.add.at(vol, orc.cells["tet"], orc.cells["geo"][:, 0])
    tv = tet_volumes(inst)
    assert np.max(np.abs(vol - tv) / tv) < 1e-11
    # centroid moments too
    mom = np.zeros((inst.n_tet, 3))
    np.add.at(mom, orc.cells["tet"],
              orc.cells["geo"][:, 0:1] * orc.cells["geo"][:, 1:])
    ct = inst.verts[inst.tets].mean(axis=1)
    np.testing.assert_allclose(mom / tv[:, None], ct, atol=1e-11)


def test_golden_fixtures_seed0_and_1():
    for seed in (0, 1):
        g = np.load(os.path.join(ROOT, "tests/golden",
                                 f"cube10_seed{seed}.npz"))
        inst = prep.cube10_instance(seed)
        orc = oracle.run(inst)
        assert np.array_equal(orc.rel_off, g["rel_off"])
        assert np.array_equal(orc.rel_ids, g["rel_ids"])
        assert orc.cells.size == int(g["n_cells"])
        assert digest_oracle(orc.cells) == str(g["cells_digest"])
        sub = orc.cells[orc.cells["tet"] < 300]
        np.testing.assert_allclose(sub["geo"], g["sub_geo"], rtol=1e-12)


def test_reference_float_arithmetic_agrees_with_exact():
    inst = prep.cube10_instance(0)
    x = oracle.run(inst, mode=0)
    r = oracle.run(inst, mode=1)
    kx = set(zip(x.cells["tet"], x.cells["seed"]))
    kr = set(zip(r.cells["tet"], r.cells["seed"]))
    # report: cells whose existence differs between reference float
    # arithmetic and exact signs (tiny cells at near-degenerate corners)
    diff = len(kx ^ kr)
    print(f"mode R vs X: {diff} of {len(kx)} cells differ")
    assert diff <= 0.001 * len(kx)


@pytest.mark.skipif(not __import__("ref").available(),
                    reason="reference sources not mounted (oracle/_ref)")
def test_geometry_matches_reference_vbw_cells():
    # the reference's own FP64 convex cell (src/CPU/convex_cell.cpp) clips
    # the same planes; volume, barycenter and facet areas must agree
    import ref
    verts, tets = prep.grid_mesh(10, 0.1, 0)
    inst = prep.build_instance(verts, tets, prep.spheres(500, 0, 0))
    orc = oracle.run(inst, tets=np.arange(0, 6000, 37))
    fl = [[2, 1, 3], [0, 2, 3], [1, 0, 3], [0, 1, 2]]
    cells = orc.cells[:300]
    jobs = []
    for c in cells:
        t, s = int(c["tet"]), int(c["seed"])
        V = inst.verts[inst.tets[t]]
        planes = []
        for k in range(int(c["nb_p"])):
            src = int(c["src"][k])
            if src < 0:
                a, b, cc = V[fl[-1 - src]]
                n = np.cross(b - a, cc - a)
                planes.append([*n, -a @ n])
            else:
                S, B = inst.sites[s], inst.sites[src]
                d = S[:3] - B[:3]
                planes.append([*d, -((S[:3] + B[:3]) @ d + B[3] - S[3]) / 2])
        lo, hi = V.min(0) - 1e-3, V.max(0) + 1e-3
        jobs.append((np.r_[lo, hi], np.array(planes)))
    res = ref.vbw_cells_guarded(jobs)
    skipped = sum(r is None for r in res)
    print(f"VBW cross-check: {len(res) - skipped} cells, {skipped} skipped "
          f"(VBW did not terminate)")
    assert skipped <= len(res) // 20
    worst, bad_tets = 0.0, set()
    for c, r in zip(cells, res):
        if r is None:
            continue
        vol, bary, areas = r
        scale = max(c["geo"][0], 1e-30) ** (1 / 3)
        ar = c["area"][:int(c["nb_p"])]
        ok = (abs(vol - c["geo"][0]) <= 1e-9 * max(vol, 1e-12) and
              np.abs(bary - c["geo"][1:]).max() <= 1e-8 * scale + 1e-14 and
              np.abs(areas - ar).max() <= 1e-8 * scale ** 2 + 1e-20)
        if not ok:
            bad_tets.add(int(c["tet"]))
            continue
        worst = max(worst, np.abs(areas - ar).max() / max(scale ** 2, 1e-30))
    print(f"VBW disagrees on cells of {len(bad_tets)} tets; worst relative "
          f"facet-area difference elsewhere {worst:.2e}")
    assert worst < 1e-8
    assert len(bad_tets) <= 0.03 * len(cells)
    # where VBW's inexact sides disagree, the oracle's cells still
    # partition their tet exactly
    if bad_tets:
        tl = np.array(sorted(bad_tets), np.int32)
        full = oracle.run(inst, tets=tl)
        vol = np.zeros(inst.n_tet)
        np.add.at(vol, full.cells["tet"], full.cells["geo"][:, 0])
        tv = tet_volumes(inst)[tl]
        assert np.max(np.abs(vol[tl] - tv) / tv) < 1e-11


@pytest.mark.skipif(not __import__("ref").available(),
                    reason="reference sources not mounted (oracle/_ref)")
def test_neighbour_lists_against_reference_regular_triangulation():
    import ref
    sph = prep.spheres(500, 0, 0, 0.02, 0.08)
    xyzw = sph.copy()
    xyzw[:, 3] **= 2
    off, ids = ref.rt_neighbors(xyzw)
    big = np.array([-100, -100, -100, 101, 101, 101.])
    valid, off2, ids2 = prep.power_neighbors(xyzw, big)
    assert np.array_equal(np.diff(off) > 0, valid > 0)
    extra = 0
    for i in range(500):
        a, b = set(ids[off[i]:off[i + 1]]), set(ids2[off2[i]:off2[i + 1]])
        assert b <= a
        extra += len(a - b)
    assert extra < 0.01 * len(ids)
    # the restricted diagram does not depend on the far hull adjacencies
    # the triangulation adds: same cells with RT lists and with ours
    verts, tets = prep.grid_mesh(10, 0.0, 0)
    inst = prep.build_instance(verts, tets, sph)
    keep = inst.site2sphere
    remap = -np.ones(500, np.int64)
    remap[keep] = np.arange(keep.size)
    rt_lists = [sorted(remap[j] for j in ids[off[i]:off[i + 1]]
                       if remap[j] >= 0) for i in keep]
    inst_rt = prep.Instance(
        inst.verts, inst.tets, inst.v_adj, inst.e_adj, inst.f_adj, inst.f_id,
        inst.n_sf, inst.sites, inst.flags,
        np.r_[0, np.cumsum([len(x) for x in rt_lists])].astype(np.int64),
        np.array([j for x in rt_lists for j in x], np.int32), keep)
    a = oracle.run(inst)
    b = oracle.run(inst_rt)
    assert list(zip(a.cells["tet"], a.cells["seed"])) == \
        list(zip(b.cells["tet"], b.cells["seed"]))
    np.testing.assert_allclose(a.cells["geo"], b.cells["geo"], rtol=1e-9,
                               atol=1e-15)
