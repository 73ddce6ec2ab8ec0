"""Writes the 10^3-grid golden fixtures (seeds 0-3) from the oracle.

The reference ships no tests, data or golden vectors, so these fixtures are
produced by this repository's oracle (itself pinned by the closed-form,
exact-arithmetic and reference-source cross-checks in
tests/test_oracle_pins.py).  Each file holds: an input digest, the full
tet -> candidate-site relation, the combinatorial digest of every emitted
cell, and the complete cells (combinatorics + FP64 volume and centroid) of
the first 300 tets.
Usage: python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]


def input_digest(inst):
    h = hashlib.sha256()
    for a in (inst.verts, inst.tets, inst.v_adj, inst.e_adj, inst.f_adj,
              inst.f_id, inst.sites, inst.flags, inst.knn_off, inst.knn_ids):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main():
    import oracle
    from libmat_amd import prep
    from rpd_compare import digest_oracle
    for seed in range(4):
        inst = prep.cube10_instance(seed)
        orc = oracle.run(inst, mode=0)
        c = orc.cells
        sub = c[c["tet"] < 300]
        ver = np.concatenate([x["ver"][:x["nb_v"]] for x in sub])
        src = np.concatenate([x["src"][:x["nb_p"]] for x in sub])
        edge = np.concatenate([x["edge"][:x["nb_e"]] for x in sub])
        np.savez_compressed(
            os.path.join(HERE, f"cube10_seed{seed}.npz"),
            input_digest=np.array(input_digest(inst)),
            cells_digest=np.array(digest_oracle(c)),
            n_cells=np.array(c.size),
            rel_off=orc.rel_off, rel_ids=orc.rel_ids,
            sub_tet=sub["tet"], sub_seed=sub["seed"], sub_nb_v=sub["nb_v"],
            sub_nb_p=sub["nb_p"], sub_nb_e=sub["nb_e"], sub_ver=ver,
            sub_src=src, sub_edge=edge, sub_geo=sub["geo"])
        print(seed, c.size)


if __name__ == "__main__":
    main()
