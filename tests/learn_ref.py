"""Test helper: numpy restatement of the learned path's proposals and
features (the device kernels k_vertex_owner + k_knn_features).

Per tet: the owners (least power distance, ties to the smaller id) of the
centroid and of the four corners; the proposal pool is the flagged owners
(key 0) and the flagged spheres of the centroid owner's neighbour list and
of each corner owner's list when that owner is not in the centroid owner's
list; a non-owner s has key = min over the five points x of
pd(x, s) - pd(x, owner(x)), evaluated in binary32 around the centroid
exactly as the kernel does (same operations, same order, no contraction),
clamped to the least normal float and truncated to the top 16 bits of its
binary32 pattern (8 exponent and 8 mantissa bits); the 32 least
(truncated key, id) are kept.  Features (bf16): centroid in domain
coordinates, then for the first 8 proposals the centre offset and radius
over the tet radius and the truncated key over its square, zero padding to
48.
"""
import numpy as np

F = np.float32
FLT_MIN = F(1.17549435e-38)


def _pd(c, s):
    dx, dy, dz = c[0] - s[..., 0], c[1] - s[..., 1], c[2] - s[..., 2]
    return ((dx * dx + dy * dy) + dz * dz) - s[..., 3]


def bf16_bits(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def key16(key):
    """the ranking key kept by the kernel: binary32 bits 30..15"""
    return int(np.asarray(key, np.float32).view(np.uint32)) >> 15


def key_of16(k16):
    return np.array([k16 << 15], np.uint32).view(np.float32)[0]


def proposals_and_features(inst, tets):
    """(knn[len(tets), 32], X bits [len(tets), 48]) for the given tets of a
    single resident instance."""
    S = inst.sites
    flags = inst.flags
    off, ids = inst.knn_off, inst.knn_ids
    lo = inst.verts.min(axis=0)
    hi = inst.verts.max(axis=0)
    ext = max(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2])
    inv_ext = 1.0 / ext if ext > 0 else 0.0
    knn = -np.ones((len(tets), 32), np.int32)
    X = np.zeros((len(tets), 48), np.uint16)

    def owner(p):
        return int(np.argmin(_pd(p, S)))  # first minimum: the smaller id

    for row, t in enumerate(tets):
        V = inst.verts[inst.tets[t]]
        C = (((V[0] + V[1]) + V[2]) + V[3]) * 0.25
        D = (V - C).astype(F)
        rho2f = F(0)
        for k in range(4):
            rho2f = max(rho2f, (D[k, 0] * D[k, 0] + D[k, 1] * D[k, 1]) +
                        D[k, 2] * D[k, 2])
        own = [owner(C)] + [owner(V[k]) for k in range(4)]
        best = _pd(C, S[own[0]])
        qx = [F(-best)]
        for k in range(4):
            dd = (D[k, 0] * D[k, 0] + D[k, 1] * D[k, 1]) + D[k, 2] * D[k, 2]
            qx.append(dd - F(_pd(V[k], S[own[k + 1]])))
        ring0 = set(ids[off[own[0]]:off[own[0] + 1]].tolist())
        pool = {}
        for o in own:
            if flags[o] == 1:
                pool[o] = F(0)
        scan = [own[0]] + [o for o in own[1:] if o not in ring0]
        for o in dict.fromkeys(scan):
            for s in ids[off[o]:off[o + 1]].tolist():
                if flags[s] != 1 or s in own or s in pool:
                    continue
                sp = S[s]
                u = (sp[:3] - C).astype(F)
                base = F((((sp[0] - C[0]) * (sp[0] - C[0]) +
                           (sp[1] - C[1]) * (sp[1] - C[1])) +
                          (sp[2] - C[2]) * (sp[2] - C[2])) - sp[3])
                key = base + qx[0]
                for k in range(4):
                    du = (D[k, 0] * u[0] + D[k, 1] * u[1]) + D[k, 2] * u[2]
                    key = min(key, (base + qx[1 + k]) - F(2) * du)
                pool[s] = max(key, FLT_MIN)
        pool = {s: key16(k) for s, k in pool.items()}
        order = sorted(pool.items(), key=lambda kv: (kv[1], kv[0]))[:32]
        for q, (s, _) in enumerate(order):
            knn[row, q] = s
        f = np.zeros(48, F)
        f[0:3] = ((C - lo) * inv_ext).astype(F)
        inv_r = F(1) / np.sqrt(rho2f) if rho2f > 0 else F(0)
        for k, (s, key) in enumerate(order[:8]):
            sp = S[s]
            f[3 + 5 * k + 0] = F(sp[0] - C[0]) * inv_r
            f[3 + 5 * k + 1] = F(sp[1] - C[1]) * inv_r
            f[3 + 5 * k + 2] = F(sp[2] - C[2]) * inv_r
            f[3 + 5 * k + 3] = F(np.sqrt(sp[3])) * inv_r
            f[3 + 5 * k + 4] = key_of16(key) * (inv_r * inv_r)
        X[row] = bf16_bits(f)
    return knn, X
