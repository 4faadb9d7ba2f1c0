"""Learned level selector: training data, training on the GPU, model files.

The selector is an MLP 10 -> 64 -> 64 -> 7 on per-pair geometry features
(qol_engine_pair_features) whose forward, backward, Adam and inference run as
HIP kernels in libqotential_hip.so (qol_selector_*).  It predicts, for each
(target, panel) pair, the coarsest way to serve the pair whose correction is
within a tolerance of the converged serving (level 4 for the pairs the
reference's ball rule, src/patch_refine_mod.f90:541-613, serves by a level; RRQ
for the rest) on smooth densities.  The ball rule serves some pairs at a level
whose error is well above that of the finer levels, and others finer than they
need; the selector moves both toward the tolerance:

  class 0   the panel's own nodes (no correction)
  class 1-4 direct quadrature on refinement level 1-4
  class 5   RRQ
  class 6   direct on level 4 where the ball rule asks for RRQ

Training labels come from the engine itself: every pair is evaluated with
each forced serving (qol_engine_force_label) and its correction on point-charge
densities compared with the converged serving's.  Below the confidence
threshold the ball rule labels the pair.
"""
import ctypes as C
from pathlib import Path

import numpy as np

from ._lib import QolError, check, dbl, i64, lib, ptr

N_FEAT, N_HID, N_CLASS = 10, 64, 7
N_W = N_HID * N_FEAT + N_HID * N_HID + N_CLASS * N_HID
N_PARAM = N_W + 2 * N_HID + N_CLASS
MODEL_FILE = Path(__file__).resolve().parent / "selector_model.npz"
# a pair's correction error against the converged serving, relative to the density's scale
LABEL_TOL = 1e-10


def init_params(seed=0):
    """He-normal weights, zero biases"""
    rng = np.random.default_rng(seed)
    w1 = rng.standard_normal((N_HID, N_FEAT)) * np.sqrt(2.0 / N_FEAT)
    w2 = rng.standard_normal((N_HID, N_HID)) * np.sqrt(2.0 / N_HID)
    w3 = rng.standard_normal((N_CLASS, N_HID)) * np.sqrt(2.0 / N_HID)
    return np.concatenate([w1.ravel(), w2.ravel(), w3.ravel(), np.zeros(2 * N_HID + N_CLASS)]).astype(np.float32)


def feature_norm(feat):
    """feature means then 1/std (std floored), float32[20]"""
    mu = feat.mean(0)
    sd = np.maximum(feat.std(0), 1e-3)
    return np.concatenate([mu, 1.0 / sd]).astype(np.float32)


class Selector:
    """A selector model on the GPU (qol_selector_create)."""

    def __init__(self, params=None, norm=None):
        h = C.c_void_p()
        check(lib().qol_selector_create(C.byref(h)), "qol_selector_create")
        self.handle = h
        self.set_params(init_params() if params is None else params,
                        np.concatenate([np.zeros(N_FEAT), np.ones(N_FEAT)]) if norm is None else norm)

    def close(self):
        if getattr(self, "handle", None):
            lib().qol_selector_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params, norm):
        params = np.ascontiguousarray(params, dtype=np.float32)
        norm = np.ascontiguousarray(norm, dtype=np.float32)
        if params.size != N_PARAM or norm.size != 2 * N_FEAT:
            raise QolError("selector parameter sizes")
        check(lib().qol_selector_set_params(self.handle, ptr(params), ptr(norm)), "qol_selector_set_params")

    def params(self):
        p = np.zeros(N_PARAM, dtype=np.float32)
        n = np.zeros(2 * N_FEAT, dtype=np.float32)
        check(lib().qol_selector_get_params(self.handle, ptr(p), ptr(n)), "qol_selector_get_params")
        return p, n

    def grad(self, feat, labels, class_w):
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        labels = np.ascontiguousarray(labels, dtype=np.int64)
        cw = np.ascontiguousarray(class_w, dtype=np.float32)
        g = np.zeros(N_PARAM, dtype=np.float32)
        loss = dbl(0.0)
        check(lib().qol_selector_grad(self.handle, i64(len(labels)), ptr(feat), ptr(labels), ptr(cw), ptr(g),
                                      C.byref(loss)), "qol_selector_grad")
        return g, loss.value

    def train(self, feat, labels, class_w, epochs=10, batch=4096, lr=2e-3, seed=0):
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        labels = np.ascontiguousarray(labels, dtype=np.int64)
        cw = np.ascontiguousarray(class_w, dtype=np.float32)
        le = np.zeros(epochs)
        check(lib().qol_selector_train(self.handle, i64(len(labels)), ptr(feat), ptr(labels), ptr(cw), i64(epochs),
                                       i64(batch), dbl(lr), C.c_uint64(seed), ptr(le)), "qol_selector_train")
        return le

    def predict(self, feat):
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        pr = np.zeros((max(len(feat), 1), N_CLASS), dtype=np.float32)
        check(lib().qol_selector_predict(self.handle, i64(len(feat)), ptr(feat), ptr(pr)), "qol_selector_predict")
        return pr[: len(feat)]

    def save(self, path, **meta):
        p, n = self.params()
        np.savez(path, params=p, norm=n, **meta)

    @classmethod
    def load(cls, path=MODEL_FILE):
        d = np.load(path, allow_pickle=False)
        return cls(d["params"], d["norm"])


def random_stellarator(rng):
    """VMEC modes of a random stellarator: R0 = 1, minor radius in [0.2, 0.35],
    three extra modes of amplitude at most 0.05; nfp in {3, 4, 5}"""
    a = rng.uniform(0.2, 0.35)
    modes = [[0, 0, 1.0, 0.0], [1, 0, a, a]]
    used = set()
    while len(modes) < 5:
        m, n = int(rng.integers(1, 3)), int(rng.integers(-1, 2))
        if (m, n) in used or (m, n) == (1, 0):
            continue
        used.add((m, n))
        modes.append([m, n, rng.uniform(-0.05, 0.05), rng.uniform(-0.05, 0.05)])
    return np.array(modes, dtype=np.float64), int(rng.choice([3, 4, 5]))


def serving_errors(engine, densities):
    """Per pair and forced serving c = 0..5 (own nodes, levels 1..4, RRQ), the
    largest error of the pair's correction over the given densities, against
    the converged serving (level 4 for pairs the ball rule serves by a level,
    RRQ for the rest), relative to each density's largest value on the
    surface; and the ball rule's labels.  densities: list of (ub, ubn) at the
    surface nodes."""
    engine.force_label(-1)
    engine.eval(exterior=True, side_tol=1e-10)
    rule = engine.weights()["label"]
    W = []
    for c in range(6):
        engine.force_label(c)
        engine.eval(exterior=True, side_tol=1e-10)
        w = engine.weights()
        W.append((w["As"], w["Ad"]))
    engine.force_label(-1)
    offs, _ = engine.pairs()
    pan = np.repeat(np.arange(len(offs) - 1), np.diff(offs))
    h = W[0][0].shape[1]
    rows = pan[:, None] * h + np.arange(h)[None, :]
    conv = np.where(rule == 5, 5, 4)
    err = np.zeros((len(rule), 6))
    for ub, ubn in densities:
        sc = max(np.abs(ub).max(), np.abs(ubn).max())
        a, b = ubn[rows], ub[rows]
        val = np.stack([(As * a).sum(1) - (Ad * b).sum(1) for As, Ad in W], 1)
        ref = val[np.arange(len(rule)), conv]
        err = np.maximum(err, np.abs(val - ref[:, None]) / sc)
    return err, rule


def classes_from_errors(err, rule, tol=LABEL_TOL):
    """the coarsest serving whose error is within tol: 0..4 by level, 6 for
    level 4 where the rule asks for RRQ, else 5 (RRQ)"""
    ok = err[:, :5] <= tol
    first = np.where(ok.any(1), ok.argmax(1), 5)
    return np.where((first == 4) & (rule == 5), 6, first).astype(np.int64)


def surface_pairs(geo, nt, seed, umin=-7.0, umax=-0.5, ndens=3):
    """(features, classes, rule labels) of the pairs of nt near-surface targets;
    the test densities are point-charge fields (8 charges 0.4 outside)"""
    from . import workload as W
    from .engine import Engine
    tx, _, _ = W.targets(geo, nt, umin, umax, sides=0, seed=seed)
    dens = []
    for i in range(ndens):
        pos, q = W.point_charges(geo, nchg=8, offset=0.4, seed=seed * 31 + i)
        ub, gu = W.charge_field(geo["sx"], pos, q)
        dens.append((ub, (geo["snx"] * gu).sum(1)))
    e = Engine(geo)
    e.set_targets(tx)
    e.nearlist()
    feat = e.pair_features()
    err, rule = serving_errors(e, dens)
    e.close()
    return feat, classes_from_errors(err, rule), rule


def class_weights(labels):
    """inverse square-root frequency, normalised to mean 1 over the samples"""
    cnt = np.bincount(labels, minlength=N_CLASS).astype(np.float64)
    w = np.where(cnt > 0, 1.0 / np.sqrt(np.maximum(cnt, 1.0)), 0.0)
    w *= len(labels) / max((w[labels]).sum(), 1e-300)
    return w.astype(np.float32)


def serve_rank(cls):
    """cost order of the servings: own nodes < level 1 < ... < level 4 (classes 4
    and 6) < RRQ"""
    cls = np.asarray(cls)
    return np.where(cls == 6, 4, np.where(cls == 5, 5, cls))


def direct_flops(p, labels):
    """algorithmic FLOPs of the direct level weights for a label vector:
    per pair served at level l, nl(l) (19 + 4h)"""
    from .engine import order_info
    info = order_info(p)
    cnt = np.bincount(np.asarray(labels), minlength=7)
    return float(sum(cnt[l] * int(info["nl"][l - 1]) * (19 + 4 * info["h"]) for l in range(1, 5)))


def ablation(selector, geo, nt=20000, seed=99, threshold=0.9, steps=5):
    """The selector against the method without it on one surface: the ball rule,
    the ball rule with every level pair one level finer, and the selector (ball
    rule below the confidence threshold).  Per variant: labels, direct FLOPs,
    ms per step (eval + corrections), Green's identity error (max and 99.9th
    percentile over the targets, point-charge density)."""
    import time
    from . import workload as W
    from .engine import Engine
    tx, _, dist = W.targets(geo, nt, -6.0, -1.0, sides=0, seed=seed)
    pos, q = W.point_charges(geo, nchg=8, offset=0.4, seed=seed + 1)
    ub, gu = W.charge_field(geo["sx"], pos, q)
    ubn = (geo["snx"] * gu).sum(1)
    uex, _ = W.charge_field(tx, pos, q)
    uref = np.where(dist < 0.0, uex, 0.0)
    e = Engine(geo)
    e.set_targets(tx)
    e.nearlist()
    far = e.far(ub, ubn)
    out = {}
    for name in ("rule", "rule+1", "selector"):
        if name == "rule":
            e.set_selector(None)
        elif name == "rule+1":
            e.refine_levels(1)
        else:
            e.set_selector(selector, threshold)
        e.eval()
        u, _ = e.apply(ub, ubn, far)
        lab = e.labels()
        err = np.abs(u - uref) / np.abs(uex).max()
        e.apply(ub, ubn, far)
        e.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            e.eval()
            e.apply_resident()
        e.sync()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        out[name] = dict(labels=np.bincount(lab, minlength=6).tolist(), direct_gflop=direct_flops(geo["p"], lab) / 1e9,
                         ms=ms, grf_max=float(err.max()), grf_p999=float(np.quantile(err, 0.999)),
                         grf_median=float(np.median(err)))
    e.set_selector(None)
    e.close()
    return out
