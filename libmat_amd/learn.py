"""Training and evaluation of the learned proposal model.

The model scores, per tet, the 32 proposals of rpd_features (the owner of
the tet centroid and its power-nearest neighbours); a plain run of the same
resident batch labels them (rpd_labels).  Training runs entirely on the
device (rpd_mlp_train: the project's BF16-MFMA forward/backward kernels and
Adam).  The logit threshold is chosen on the training batch for a target
recall of the non-owner proposals (the owner is always clipped).
"""
import numpy as np


def train_proposals(ctx, instances, epochs=6, batch=1024, lr=2e-3,
                    pos_weight=4.0, seed=0, init_seed=0, recall=0.995):
    """Labels from a plain run of ``instances`` (one resident batch), then
    training; returns (params, loss per epoch, logit threshold)."""
    ctx.upload_batch(instances)
    ctx.run()
    ctx.features(download=False)
    ctx.labels(download=False)
    ctx.mlp_init(init_seed)
    loss = ctx.train(epochs=epochs, batch=batch, lr=lr,
                     pos_weight=pos_weight, seed=seed)
    Y, M = ctx.labels()
    thr = choose_threshold(ctx.scores(), Y, M, recall)
    return ctx.mlp_get(), loss, thr


def choose_threshold(z, Y, M, recall=0.995):
    """Largest logit threshold keeping ``recall`` of the positive non-owner
    proposals."""
    pos = z[:, 1:][(Y[:, 1:] > 0.5) & (M[:, 1:] > 0.5)]
    if pos.size == 0:
        return 0.0
    return float(np.quantile(pos, 1.0 - recall))


def proposal_quality(z, Y, M, thr):
    """Precision and recall of the clipped proposals (owner always)."""
    sel = z > thr
    sel[:, 0] = True
    sel &= M > 0.5
    pos = (Y > 0.5) & (M > 0.5)
    tp = (sel & pos).sum()
    return (float(tp / max(1, sel.sum())), float(tp / max(1, pos.sum())))
