"""The learned level selector's HIP kernels: gradients and probabilities
against a PyTorch FP32 reference of the same network (BF16-rounded weights,
FP32 arithmetic), training that lowers the loss, and forced servings in the
engine (the selector's training labels come from them)."""
import numpy as np
import pytest
import torch

from qotential_amd import selector as S
from qotential_amd import workload as W
from qotential_amd.engine import Engine

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def bf16(x):
    """round float32 to the nearest BF16 (ties to even), as float32"""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def torch_reference(params, norm, feat, labels, cw):
    nh, nf, nc = S.N_HID, S.N_FEAT, S.N_CLASS
    w = bf16(params[:S.N_W])
    o1, o2 = nh * nf, nh * nf + nh * nh
    t = lambda a: torch.tensor(np.array(a, dtype=np.float32), requires_grad=True)
    W1, W2, W3 = t(w[:o1].reshape(nh, nf)), t(w[o1:o2].reshape(nh, nh)), t(w[o2:].reshape(nc, nh))
    b = params[S.N_W:]
    b1, b2, b3 = t(b[:nh]), t(b[nh:2 * nh]), t(b[2 * nh:])
    x = (torch.tensor(feat) - torch.tensor(norm[:nf])) * torch.tensor(norm[nf:])
    h1 = torch.relu(x @ W1.T + b1)
    h2 = torch.relu(h1 @ W2.T + b2)
    lg = h2 @ W3.T + b3
    y = torch.tensor(labels)
    per = torch.logsumexp(lg, 1) - lg[torch.arange(len(y)), y]
    loss = (torch.tensor(cw)[y] * per).mean()
    loss.backward()
    g = torch.cat([W1.grad.ravel(), W2.grad.ravel(), W3.grad.ravel(), b1.grad, b2.grad, b3.grad]).numpy()
    return g, float(loss), torch.softmax(lg, 1).detach().numpy()


def test_selector_gradient_and_probabilities_vs_torch():
    rng = np.random.default_rng(7)
    n = 777
    feat = (rng.standard_normal((n, S.N_FEAT)) * 2 + 0.5).astype(np.float32)
    labels = rng.integers(0, S.N_CLASS, n)
    cw = rng.uniform(0.5, 2.0, S.N_CLASS).astype(np.float32)
    params = S.init_params(3)
    params[S.N_W:] = rng.standard_normal(S.N_PARAM - S.N_W).astype(np.float32) * 0.1
    norm = S.feature_norm(feat)
    sel = S.Selector(params, norm)
    g, loss = sel.grad(feat, labels, cw)
    pr = sel.predict(feat)
    sel.close()
    g_ref, loss_ref, pr_ref = torch_reference(params, norm, feat, labels, cw)
    rel = np.abs(g - g_ref).max() / np.abs(g_ref).max()
    say(f"gradient vs torch: max rel {rel:.2e}; loss {loss:.6f} vs {loss_ref:.6f}")
    assert rel < 1e-4
    assert abs(loss - loss_ref) < 1e-5 * abs(loss_ref)
    assert np.abs(pr - pr_ref).max() < 1e-5


@pytest.fixture(scope="module")
def small_set():
    geo = W.geometry(12, 36, 6, modes=W.VMEC_BENCH)
    feat, cls, rule = S.surface_pairs(geo, 3000, seed=21)
    return geo, feat, cls, rule


def test_forced_servings_and_labels(small_set):
    geo, feat, cls, rule = small_set
    say("pairs", len(cls), "classes", np.bincount(cls, minlength=S.N_CLASS).tolist())
    assert np.all(np.isfinite(feat)) and feat.shape == (len(cls), S.N_FEAT)
    # RRQ pairs are never served by levels 1..3, and level 4 for them is class 6
    assert not np.isin(cls[rule == 5], (1, 2, 3, 4)).any()
    assert not (cls[rule <= 4] == 6).any()
    tx, _, _ = W.targets(geo, 3000, -7.0, -0.5, sides=0, seed=21)
    pos, q = W.point_charges(geo, seed=4)
    ub, gu = W.charge_field(geo["sx"], pos, q)
    ubn = (geo["snx"] * gu).sum(1)
    e = Engine(geo)
    e.set_targets(tx)
    e.nearlist()
    far = e.far(ub, ubn)
    e.force_label(0)
    e.eval()
    u0, val0 = e.apply(ub, ubn, far)
    assert not val0.any() and np.array_equal(u0, far)
    e.force_label(2)
    e.eval()
    assert (e.weights()["label"] == 2).all()
    e.close()


def test_selector_training_lowers_loss(small_set):
    geo, feat, cls, rule = small_set
    norm = S.feature_norm(feat)
    cw = S.class_weights(cls)
    sel = S.Selector(S.init_params(5), norm)
    loss = sel.train(feat, cls, cw, epochs=40, batch=512, lr=3e-3, seed=1)
    pred = sel.predict(feat).argmax(1)
    sel.close()
    acc = float((pred == cls).mean())
    say(f"loss {loss[0]:.3f} -> {loss[-1]:.3f}, train accuracy {acc:.3f}")
    assert loss[-1] < 0.5 * loss[0]
    assert acc > 0.8


def test_served_labels_follow_the_network():
    """The classification runs the network on the matrix cores, 64 pairs per
    wave in groups of 16 (ragged at each panel's end).  Every pair's label must
    be the FP32 network's (PyTorch, same BF16 weights) on the engine's own
    features of the pair: the class of the largest probability where that
    probability reaches the threshold (class 6 served by level 4), else the
    ball rule's.  Pairs within 1e-5 of a tie or of the threshold can fall
    either way under another summation order."""
    geo = W.geometry(12, 36, 8, modes=W.VMEC_BENCH)
    tx, _, _ = W.targets(geo, 20000, -7.0, -0.5, sides=0, seed=5)
    sel = S.Selector.load()
    params, norm = sel.params()
    thr = 0.9
    e = Engine(geo)
    e.set_targets(tx)
    e.nearlist()
    feat = e.pair_features()
    e.eval()
    rule = e.labels()
    e.set_selector(sel, thr)
    e.eval()
    lab = e.labels()
    e.set_selector(None)
    e.close()
    sel.close()
    _, _, pr = torch_reference(params, norm, feat, np.zeros(len(feat), dtype=np.int64),
                               np.ones(S.N_CLASS, dtype=np.float32))
    top = np.sort(pr, 1)
    best = pr.argmax(1)
    want = np.where(pr.max(1) >= thr, np.where(best == 6, 4, best), rule)
    tie = (top[:, -1] - top[:, -2] < 1e-5) | (np.abs(top[:, -1] - thr) < 1e-5)
    diff = lab != want
    say(f"{len(lab)} pairs, {int((lab != rule).sum())} moved off the ball rule, "
        f"{int(tie.sum())} within 1e-5 of a tie or the threshold, {int(diff.sum())} differ from torch")
    assert len(lab) > 20000 and (lab != rule).mean() > 0.01
    assert not diff[~tie].any()
    assert tie.mean() < 1e-3
