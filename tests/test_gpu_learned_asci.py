"""Learned ASCI pre-screen end to end on the GPU (H2O cc-pVDZ).

The network is trained by the project's own BF16 kernels on
(candidate, log|rv|) pairs exported from plain asci_search steps; the
screened search keeps 5% of the candidates by prediction and lets the exact
scores pick the top-k. Reported: recall against the plain top-k, the
variational energy difference after re-diagonalisation, and the gain over
a pre-screen without training (random network of the same shape).
"""
import numpy as np
import pytest

from macis_amd import _capi
from macis_amd import dets as D

from conftest import announce

pytestmark = pytest.mark.gpu
EPS = np.finfo(float).eps


def grow(ctx, W, hf, norb, sizes, E0):
    """Plain GPU ASCI iterations (asci_iter) to the listed sizes."""
    wfn, X = hf.copy(), np.array([1.0])
    s = _capi.ASCISettings.defaults()
    for nd in sizes:
        order = np.argsort(-np.abs(X), kind="stable")
        wfn, X = wfn[order], X[order]
        nc = min(100, len(wfn))
        out = ctx.empty((nd, W), np.uint64)
        n = ctx.asci_search(W, s, nd, ctx.to_device(wfn[:nc]), nc, E0, ctx.to_device(X[:nc]),
                            norb, out)
        wfn = out.numpy()[:n]
        dX = ctx.empty(n, np.float64).zero()
        E0, _, _ = ctx.selected_ci_diag(W, ctx.to_device(wfn), n, dX)
        X = dX.numpy()
    return E0, wfn, X


def diag(ctx, W, dets):
    dX = ctx.empty(len(dets), np.float64).zero()
    E, _, _ = ctx.selected_ci_diag(W, ctx.to_device(dets), len(dets), dX)
    return E


def test_learned_prescreen_recall_energy_and_gain(ctx, h2o):
    announce("learned ASCI pre-screen on H2O")
    norb, _, T, V = h2o
    ctx.set_integrals(norb, T, V)
    W = 1
    hf = D.canonical_hf_determinant(W, 5, 5)
    s = _capi.ASCISettings.defaults()
    E_hf = None
    # training data from two plain steps
    E1, w1, X1 = grow(ctx, W, hf, norb, [100, 800], 0.0 if E_hf else -85.0)
    feats, targ = [], []
    for wfn, X, E0 in [(w1, X1, E1)]:
        order = np.argsort(-np.abs(X), kind="stable")
        core, Cc = wfn[order[:100]], X[order[:100]]
        dd, da, n = ctx.asci_candidates(W, s, ctx.to_device(core), 100, E0, ctx.to_device(Cc), norb,
                                        4_000_000)
        feats.append(dd.numpy()[:n])
        targ.append(np.log(da.numpy()[:n]))
    Xd = np.concatenate(feats)
    y = np.concatenate(targ).astype(np.float32)
    mu, sd = y.mean(), y.std()
    yn = ((y - mu) / sd).astype(np.float32)
    net = _capi.MLP(ctx, seed=3)
    untrained = _capi.MLP(ctx, seed=3)
    rng = np.random.default_rng(0)
    dX_all, dy_all = ctx.to_device(Xd), ctx.to_device(yn)
    first = net.train_step(W, dX_all, dy_all, min(len(yn), 65536), 0.0)
    bs = 8192
    for epoch in range(6):
        perm = rng.permutation(len(yn))
        for b0 in range(0, len(yn) - bs + 1, bs):
            idx = perm[b0:b0 + bs]
            net.train_step(W, ctx.to_device(Xd[idx]), ctx.to_device(yn[idx]), bs, 1e-3)
    last = net.train_step(W, dX_all, dy_all, min(len(yn), 65536), 0.0)
    assert last < 0.5 * first, (first, last)

    # next step: plain vs screened from the same core
    order = np.argsort(-np.abs(X1), kind="stable")
    core, Cc = w1[order[:100]], X1[order[:100]]
    nd = 6400
    dcore, dC = ctx.to_device(core), ctx.to_device(Cc)
    out_p = ctx.empty((nd, W), np.uint64)
    n_p = ctx.asci_search(W, s, nd, dcore, 100, E1, dC, norb, out_p)
    t_plain = ctx.timings()["asci_total_ms"]
    plain = out_p.numpy()[:n_p]
    res = {}
    for name, m in [("trained", net), ("untrained", untrained)]:
        out = ctx.empty((nd, W), np.uint64)
        n = ctx.asci_search_screened(W, s, m, 0.05, nd, dcore, 100, E1, dC, norb, out)
        t = ctx.timings()
        got = out.numpy()[:n]
        recall = len({tuple(r) for r in got.tolist()} & {tuple(r) for r in plain.tolist()}) / n_p
        res[name] = (recall, got, t["asci_total_ms"], t["asci_nn_ms"])
    E_plain = diag(ctx, W, plain)
    E_nn = diag(ctx, W, res["trained"][1])
    E_rand = diag(ctx, W, res["untrained"][1])
    print(f"plain {t_plain:.2f} ms; trained recall {res['trained'][0]:.4f} "
          f"({res['trained'][2]:.2f} ms, nn {res['trained'][3]:.2f} ms); "
          f"untrained recall {res['untrained'][0]:.4f}; dE trained {E_nn - E_plain:.3e}, "
          f"untrained {E_rand - E_plain:.3e}")
    # the trained network beats the untrained one, and the variational
    # energy stays within 0.1 mHa of the plain search
    assert res["trained"][0] > res["untrained"][0]
    assert E_nn - E_plain < E_rand - E_plain
    assert abs(E_nn - E_plain) <= 1e-4
