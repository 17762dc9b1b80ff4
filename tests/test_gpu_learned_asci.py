"""Learned ASCI pre-screen end to end on the GPU (H2O cc-pVDZ).

Plain GPU ASCI steps grow the wavefunction HF -> 100 -> 800 -> 6400
determinants (asci_iter, iteration.hpp:16-67); every step also exports its
scored candidates (determinant, accumulated |rv|). The network is trained
by the project's own BF16 kernels on those (candidate, log|rv|) pairs (the
bottom 70% of targets clipped to their 70th percentile, so the network
spends its capacity on ranking the candidates that can enter the top-k) and
then screens the NEXT step, a refinement step at 6400 determinants from the
6400-determinant wavefunction that it has never seen: the best-predicted 5%
of the candidates survive and their exact scores pick the top-k. Reported:
recall against the plain top-k, the variational energy difference after
re-diagonalisation (bar: 0.1 mHa), the gain over a network of the same
shape without training, and the step times.
"""
import numpy as np
import pytest

from macis_amd import _capi
from macis_amd import dets as D

from conftest import announce

pytestmark = pytest.mark.gpu
KEEP = 0.05
EPOCHS, BATCH, LR = 20, 4096, 1e-3
CLIP_Q = 0.7  # targets below this quantile are clipped: the ranking of the top matters


def core_of(wfn, X, ncore=100):
    order = np.argsort(-np.abs(X), kind="stable")
    return wfn[order[:ncore]], X[order[:ncore]]


def diag(ctx, W, dets):
    dX = ctx.empty(len(dets), np.float64).zero()
    E, _, _ = ctx.selected_ci_diag(W, ctx.to_device(dets), len(dets), dX)
    return E, dX.numpy()


def train(ctx, net, W, dets, y, epochs, batch, lr, seed=0):
    """Adam epochs over shuffled mini-batches; each epoch's permutation is
    uploaded once and the steps walk through it on the device."""
    rng = np.random.default_rng(seed)
    n = len(y)
    dd = ctx.empty((n, W), np.uint64)
    dy = ctx.empty(n, np.float32)
    for _ in range(epochs):
        perm = rng.permutation(n)
        dd.upload(dets[perm])
        dy.upload(y[perm])
        for b0 in range(0, n - batch + 1, batch):
            net.train_step(W, dd.ptr + 8 * W * b0, dy.ptr + 4 * b0, batch, lr, want_loss=False)


def test_learned_prescreen_recall_energy_and_gain(ctx, h2o):
    announce("learned ASCI pre-screen on H2O")
    norb, _, T, V = h2o
    ctx.set_integrals(norb, T, V)
    W = 1
    s = _capi.ASCISettings.defaults()
    hf = D.canonical_hf_determinant(W, 5, 5)
    dE = ctx.empty(1, np.float64)
    ctx.ham_diagonal(W, ctx.to_device(hf), 1, dE)
    E0 = float(dE.numpy()[0])  # the HF energy starts the first step
    wfn, X = hf.copy(), np.array([1.0])
    feats, targ = [], []
    for nd in [100, 800, 6400]:  # plain steps, candidates kept as training data
        core, Cc = core_of(wfn, X)
        dcore, dC = ctx.to_device(core), ctx.to_device(Cc)
        dd, da, n = ctx.asci_candidates(W, s, dcore, len(core), E0, dC, norb, 4_000_000)
        feats.append(dd.numpy()[:n])
        targ.append(np.log(da.numpy()[:n]))
        out = ctx.empty((nd, W), np.uint64)
        n = ctx.asci_search(W, s, nd, dcore, len(core), E0, dC, norb, out)
        wfn = out.numpy()[:n]
        E0, X = diag(ctx, W, wfn)
    Xd = np.concatenate(feats)
    y = np.concatenate(targ).astype(np.float32)
    y = np.maximum(y, np.quantile(y, CLIP_Q))
    yn = ((y - y.mean()) / y.std()).astype(np.float32)
    net = _capi.MLP(ctx, seed=3)
    untrained = _capi.MLP(ctx, seed=3)
    dX_all, dy_all = ctx.to_device(Xd[:65536]), ctx.to_device(yn[:65536])
    first = net.train_step(W, dX_all, dy_all, min(len(yn), 65536), 0.0)
    train(ctx, net, W, Xd, yn, EPOCHS, BATCH, LR)
    last = net.train_step(W, dX_all, dy_all, min(len(yn), 65536), 0.0)
    assert last < 0.5 * first, (first, last)

    # the unseen step: plain vs screened from the same core
    nd = len(wfn)
    core, Cc = core_of(wfn, X)
    dcore, dC = ctx.to_device(core), ctx.to_device(Cc)
    out_p = ctx.empty((nd, W), np.uint64)
    n_p = ctx.asci_search(W, s, nd, dcore, len(core), E0, dC, norb, out_p)
    t_plain = ctx.timings()["asci_total_ms"]
    plain = out_p.numpy()[:n_p]
    res = {}
    for name, m in [("trained", net), ("untrained", untrained)]:
        out = ctx.empty((nd, W), np.uint64)
        n = ctx.asci_search_screened(W, s, m, KEEP, nd, dcore, len(core), E0, dC, norb, out)
        t = ctx.timings()
        got = out.numpy()[:n]
        assert n == n_p
        recall = len(set(got[:, 0].tolist()) & set(plain[:, 0].tolist())) / n_p
        res[name] = (recall, got, t["asci_total_ms"], t["asci_nn_ms"], t["asci_nsurvivors"])
    E_plain = diag(ctx, W, plain)[0]
    E_nn = diag(ctx, W, res["trained"][1])[0]
    E_rand = diag(ctx, W, res["untrained"][1])[0]
    print(f"\n{len(yn)} training pairs; plain step {t_plain:.2f} ms; trained: recall "
          f"{res['trained'][0]:.4f}, step {res['trained'][2]:.2f} ms (network "
          f"{res['trained'][3]:.2f} ms, {res['trained'][4]} survivors); untrained recall "
          f"{res['untrained'][0]:.4f}; dE trained {E_nn - E_plain:+.3e} Ha, untrained "
          f"{E_rand - E_plain:+.3e} Ha")
    # the trained network beats the untrained one, and the variational
    # energy stays within 0.1 mHa of the plain search
    assert res["trained"][0] > res["untrained"][0] + 0.2
    assert E_nn - E_plain < E_rand - E_plain
    assert abs(E_nn - E_plain) <= 1e-4
