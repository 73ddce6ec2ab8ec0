"""Learned proposals on the GPU: model gradients and inference logits
against the CPU oracle, proposals and features against their restatement,
bit-exact learned runs, training gain over the untrained model and over
top-k proposals."""
import numpy as np
import pytest

import mlp_oracle as mo
import oracle
from libmat_amd import learn, prep
from rpd_compare import compare_with_oracle

pytestmark = pytest.mark.gpu


def _same_result(a, b):
    for f in ("cells", "words", "geom", "areas"):
        x, y = getattr(a, f), getattr(b, f)
        if f == "cells":
            for k in ("tet", "seed", "nb_v", "nb_p", "nb_e", "status",
                      "word_off", "area_off"):
                assert np.array_equal(x[k], y[k]), (f, k)
        else:
            assert np.array_equal(x, y), f


def test_mlp_gradient_matches_oracle(ctx):
    rng = np.random.default_rng(5)
    n = 203
    params = mo.xavier(7).astype(np.float32)
    Xb = mo.bf16_bits(rng.normal(size=(n, mo.IN)).astype(np.float32))
    Y = (rng.random((n, mo.OUT)) < 0.25).astype(np.float32)
    M = (rng.random((n, mo.OUT)) < 0.9).astype(np.float32)
    g_gpu, l_gpu = ctx.mlp_grad(params, Xb, Y, M, 4.0)
    l_orc, g_orc = mo.loss_and_grad(params.astype(np.float64),
                                    mo.from_bits(Xb), Y, M, 4.0, True)
    print(f"[mlp] loss gpu {l_gpu:.6f} oracle {l_orc:.6f}", flush=True)
    assert abs(l_gpu - l_orc) <= 1e-4 * abs(l_orc)
    o = 0
    for name, shape in mo.SHAPES:
        k = int(np.prod(shape))
        a, b = g_gpu[o:o + k], g_orc[o:o + k]
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-12)
        cos = a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30)
        print(f"[mlp] {name}: max rel err {err:.2e} cosine {cos:.6f}",
              flush=True)
        assert err < 2e-2 and cos > 0.9995, name
        o += k


def test_proposals_start_with_the_owners(ctx):
    # the owners (least power distance) of the centroid and of the 4
    # corners, by brute force, lead every proposal list (key 0, by id)
    inst = prep.cube10_instance(0)
    ctx.upload(inst)
    knn, X = ctx.features()
    V = inst.verts[inst.tets]
    pts = np.concatenate([V.mean(axis=1)[:, None, :], V], axis=1)  # (nt,5,3)
    pd = ((pts[:, :, None, :] - inst.sites[None, None, :, :3]) ** 2).sum(-1) \
        - inst.sites[None, None, :, 3]
    owners = np.argmin(pd, axis=2)
    for t in range(inst.n_tet):
        U = sorted(set(owners[t].tolist()))
        assert knn[t][:len(U)].tolist() == U, t
        ids = knn[t][knn[t] >= 0]
        assert len(set(ids.tolist())) == ids.size
    assert X.shape == (inst.n_tet, 48) and (X[:, 43:] == 0).all()


def test_mlp_inference_matches_oracle_forward(ctx):
    # k_mlp_infer's logits on the resident features against the binary64
    # forward pass with the device's bf16 rounding points; the residual is
    # fp32 accumulation order (and a hidden activation rounded to the other
    # bf16 neighbour where its fp32 value lands next to a tie)
    inst = prep.cube10_instance(2)
    ctx.upload(inst)
    knn, X = ctx.features()
    params = mo.xavier(9).astype(np.float32)
    ctx.mlp_set(params)
    z_gpu = ctx.scores().astype(np.float64)
    z_orc = mo.forward(params.astype(np.float64), mo.from_bits(X))
    err = np.abs(z_gpu - z_orc)
    scale = np.abs(z_orc).max()
    frac_close = float((err <= 1e-4 * (1.0 + np.abs(z_orc))).mean())
    cos = float((z_gpu * z_orc).sum() /
                (np.linalg.norm(z_gpu) * np.linalg.norm(z_orc)))
    print(f"[mlp] inference: max abs err {err.max():.2e} (logit scale "
          f"{scale:.2f}), within 1e-4 rel {frac_close:.5f}, cosine {cos:.8f}",
          flush=True)
    assert err.max() <= 1e-2 * scale
    assert frac_close >= 0.995 and cos > 0.99999
    # the threshold mask the learned path uses is the logits' own
    thr = float(np.median(z_gpu))
    assert ((z_gpu > thr) == (ctx.scores() > thr)).all()


@pytest.mark.parametrize("case", ["cube10", "many_sites"])
def test_features_match_restatement(ctx, case):
    # proposals (ids and order) and the bf16 feature rows of k_knn_features
    # against the numpy restatement of the same binary32 arithmetic; an
    # instance of more than 65,536 sites takes the kernel's 64-bit entries
    from learn_ref import proposals_and_features
    if case == "cube10":
        inst, n_sample = prep.cube10_instance(3), 800
    else:
        verts, tets = prep.grid_mesh(6, 0.1, 2)
        inst = prep.build_instance(verts, tets,
                                   prep.spheres(90000, 0, 4, 0.001, 0.002))
        n_sample = 150
        assert inst.n_site > 65536
    print(f"[features] {case}: n_site={inst.n_site}", flush=True)
    ctx.upload(inst)
    knn, X = ctx.features()
    sample = np.sort(np.random.default_rng(3).choice(inst.n_tet, n_sample,
                                                     replace=False))
    knn_r, X_r = proposals_and_features(inst, sample)
    bad_ids = np.nonzero((knn[sample] != knn_r).any(axis=1))[0]
    bad_x = np.nonzero((X[sample] != X_r).any(axis=1))[0]
    print(f"[features] {sample.size} tets: proposal rows differing "
          f"{bad_ids.size}, feature rows differing {bad_x.size}", flush=True)
    if bad_ids.size:
        t = bad_ids[0]
        print("gpu", knn[sample[t]].tolist(), "\nref", knn_r[t].tolist())
    if bad_x.size:
        t = bad_x[0]
        print("gpu", X[sample[t]].tolist(), "\nref", X_r[t].tolist())
    assert bad_ids.size == 0 and bad_x.size == 0


@pytest.mark.parametrize("mode", ["untrained", "topk"])
def test_learned_runs_match_plain(ctx, mode):
    inst = prep.cube10_instance(1)
    orc = oracle.run(inst)
    ctx.upload(inst)
    ctx.run()
    plain = ctx.download()
    if mode == "untrained":
        ctx.mlp_init(3)
        st = ctx.run(learned=1, logit_threshold=0.0)
    else:
        st = ctx.run(learned=2, proposal_k=4)
    res = ctx.download()
    print(f"[learned] {mode}: first-round candidates {st['learned'][0]}, "
          f"re-run tets {st['learned'][1]} of {inst.n_tet} (+"
          f"{st['learned'][2]} candidates, {st['learned'][3]} on the full "
          f"relation)", flush=True)
    _same_result(res, plain)
    compare_with_oracle(res, orc)


def test_learned_batch_matches_plain_batch(ctx):
    # uneven instances in one batch, uniform and clustered spheres: the
    # learned and top-k pipelines emit the plain batch's cells exactly
    insts = [prep.cube10_instance(5), prep.cube10_instance(6)]
    verts, tets = prep.grid_mesh(12, 0.1, 8)
    insts.append(prep.build_instance(verts, tets, prep.spheres(2500, 2, 8)))
    verts, tets = prep.grid_mesh(7, 0.1, 9)
    insts.append(prep.build_instance(verts, tets, prep.spheres(300, 1, 9)))
    ctx.upload_batch(insts)
    ctx.run()
    plain = ctx.download()
    ctx.mlp_init(4)
    for kw in (dict(learned=1, logit_threshold=0.0),
               dict(learned=1, logit_threshold=2.0),
               dict(learned=2, proposal_k=6)):
        st = ctx.run(**kw)
        print(f"[learned] batch {kw}: first-round {st['learned'][0]}, "
              f"re-run tets {st['learned'][1]}, full relation "
              f"{st['learned'][3]}", flush=True)
        _same_result(ctx.download(), plain)


def test_training_gains_over_untrained_and_topk(ctx):
    train = [prep.cube10_instance(s) for s in range(16)]
    params, loss, thr = learn.train_proposals(ctx, train, epochs=6)
    print(f"[train] loss per epoch {np.round(loss, 4).tolist()} "
          f"threshold {thr:.3f}", flush=True)
    assert loss[-1] < 0.8 * loss[0]
    test = [prep.cube10_instance(s) for s in (1000, 1001)]
    ctx.upload_batch(test)
    ctx.run()
    plain = ctx.download()
    ctx.features(download=False)
    Y, M = ctx.labels()
    z = ctx.scores()
    prec, rec = learn.proposal_quality(z, Y, M, thr)
    st = ctx.run(learned=1, logit_threshold=thr)
    _same_result(ctx.download(), plain)
    fb_learned = st["learned"][1]
    ctx.mlp_init(3)
    st0 = ctx.run(learned=1, logit_threshold=0.0)
    fb_untrained = st0["learned"][1]
    # top-k at the smallest k with at least as many first-round candidates
    # as the learned path: the model must leave fewer tets open
    for k in range(1, 33):
        st1 = ctx.run(learned=2, proposal_k=k)
        if st1["learned"][0] >= st["learned"][0]:
            break
    _same_result(ctx.download(), plain)
    fb_topk = st1["learned"][1]
    n = sum(i.n_tet for i in test)
    print(f"[train] held-out precision {prec:.3f} recall {rec:.4f}; fallback "
          f"tets learned {fb_learned}, untrained {fb_untrained}, top-{k} "
          f"{fb_topk} of {n}; first-round candidates learned "
          f"{st['learned'][0]}, top-{k} {st1['learned'][0]}", flush=True)
    assert rec > 0.98
    assert fb_learned < fb_untrained
    assert st["learned"][0] <= st1["learned"][0]
    assert fb_learned < fb_topk
