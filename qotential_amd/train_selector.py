"""Train the learned level selector on the GPU and write its model file.

Training set: near-surface targets on random stellarators (R0 = 1, minor radius
in [0.2, 0.35], three extra modes of amplitude at most 0.05, nfp in {3, 4, 5}),
orders cycling through 8, 6, 8, 10, every pair labelled with the coarsest
serving that agrees with the ball rule's weights to the label tolerance
(qotential_amd.selector.classes_from_errors).  Training runs the library's HIP
kernels (qol_selector_train: forward, backward, Adam).

  python -m qotential_amd.train_selector --out qotential_amd/selector_model.npz
"""
import argparse
import sys
import time

import numpy as np

from . import selector as S
from . import workload as W

ORDERS = (8, 6, 8, 10)


def build_set(nsurf, pairs, seed, log=print):
    rng = np.random.default_rng(seed)
    feats, labs, rules = [], [], []
    per = pairs // nsurf
    for i in range(nsurf):
        modes, nfp = S.random_stellarator(rng)
        p = ORDERS[i % len(ORDERS)]
        geo = W.geometry(12, 12 * nfp, p, modes=modes, nfp=nfp)
        nt = max(200, per // 10)
        t0 = time.perf_counter()
        f, c, r = S.surface_pairs(geo, nt, seed=1000 + i)
        # keep about `per` pairs of this surface
        if len(c) > per:
            keep = np.sort(rng.choice(len(c), per, replace=False))
            f, c, r = f[keep], c[keep], r[keep]
        feats.append(f); labs.append(c); rules.append(r)
        log(f"[selector] surface {i + 1}/{nsurf}: p={p} nfp={nfp} {geo['ntri']} triangles, {len(c)} pairs, "
            f"classes {np.bincount(c, minlength=S.N_CLASS).tolist()} ({time.perf_counter() - t0:.1f}s)")
    return np.concatenate(feats), np.concatenate(labs), np.concatenate(rules)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--surfaces", type=int, default=32)
    ap.add_argument("--pairs", type=int, default=2000000)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=str(S.MODEL_FILE))
    a = ap.parse_args(argv)
    log = lambda *x: print(*x, flush=True)
    feat, lab, rule = build_set(a.surfaces, a.pairs, a.seed, log)
    rng = np.random.default_rng(a.seed + 1)
    hold = rng.random(len(lab)) < 0.05
    norm = S.feature_norm(feat[~hold])
    cw = S.class_weights(lab[~hold])
    sel = S.Selector(S.init_params(a.seed), norm)
    t0 = time.perf_counter()
    loss = sel.train(feat[~hold], lab[~hold], cw, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed)
    log(f"[selector] trained {a.epochs} epochs on {int((~hold).sum())} pairs in {time.perf_counter() - t0:.1f}s; "
        f"loss per epoch {np.round(loss, 4).tolist()}")
    pr = sel.predict(feat[hold])
    pred = pr.argmax(1)
    conf = pr.max(1)
    acc = float((pred == lab[hold]).mean())
    sure = conf >= 0.9
    coarse = (S.serve_rank(pred) < S.serve_rank(lab[hold])) & sure
    log(f"[selector] held-out: accuracy {acc:.4f}, confident {sure.mean():.4f}, "
        f"confident and coarser than the label {coarse.mean():.2e}")
    sel.save(a.out, loss=loss, holdout_accuracy=acc, class_counts=np.bincount(lab, minlength=S.N_CLASS),
             label_tol=S.LABEL_TOL, surfaces=a.surfaces, pairs=len(lab))
    log(f"[selector] wrote {a.out}")
    # an unseen surface: the ball rule, the rule one level finer, the selector
    modes, nfp = S.random_stellarator(np.random.default_rng(a.seed + 12345))
    geo = W.geometry(16, 16 * nfp, 8, modes=modes, nfp=nfp)
    for name, r in S.ablation(sel, geo).items():
        log(f"[selector] unseen surface ({geo['ntri']} triangles, p=8) {name}: {r}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
