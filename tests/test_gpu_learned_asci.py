"""Learned ASCI block screen end to end on the GPU (H2O cc-pVDZ).

Plain GPU ASCI steps grow the wavefunction HF -> 100 -> 800 -> 6400
determinants (asci_iter, iteration.hpp:16-67); the excitation blocks of
every step's core (mhip_asci_blocks: one block per hole set, target = log of
the block's largest |h / (E - H_kk)|) are the training pairs. The network is
trained by the project's own BF16 kernels and then screens the NEXT step, a
refinement step at 6400 determinants from the 6400-determinant wavefunction:
only the best-scored blocks (log|C_core| + prediction) are generated, and
their exact contributions are accumulated and ranked as in the plain search.

Asserted: the screen removes work (the kept blocks carry a small share of
the plain search's raw candidates); every block kept reproduces the plain
search; the trained network beats an untrained one of the same shape in
recall and in the variational energy after re-diagonalisation. Reported
beside them: the screen that ranks blocks by |C_core| alone (no network), and
the energy difference to the plain search (the 0.1 mHa bar is not reached at
these keep fractions: the plain top-k is spread over most blocks).
"""
import numpy as np
import pytest

from macis_amd import _capi
from macis_amd import dets as D

from conftest import announce

pytestmark = pytest.mark.gpu
EPOCHS, BATCH, LR = 20, 1024, 1e-3


def core_of(wfn, X, ncore=100):
    order = np.argsort(-np.abs(X), kind="stable")
    return wfn[order[:ncore]], X[order[:ncore]]


def diag(ctx, W, dets):
    dets = dets[np.argsort(dets[:, 0], kind="stable")]
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


def block_coeff_scores(core, C):
    """log|C_core| repeated over each core determinant's blocks."""
    cnt = []
    for d in core[:, 0].tolist():
        a, b = bin(d & 0xFFFFFFFF).count("1"), bin(d >> 32).count("1")
        cnt.append(a + b + a * (a - 1) // 2 + b * (b - 1) // 2 + a * b)
    return np.repeat(np.log(np.abs(C)), cnt).astype(np.float32)


def test_learned_block_screen(ctx, h2o):
    announce("learned ASCI block screen on H2O")
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
    for nd in [100, 800, 6400]:  # plain steps; their cores' blocks are the training data
        core, Cc = core_of(wfn, X)
        dcore, dC = ctx.to_device(core), ctx.to_device(Cc)
        d_bd, d_t, nb = ctx.asci_blocks(W, s, dcore, len(core), E0, norb)
        t = d_t.numpy()[:nb]
        keep = np.isfinite(t)
        feats.append(d_bd.numpy()[:nb][keep])
        targ.append(t[keep])
        out = ctx.empty((nd, W), np.uint64)
        n = ctx.asci_search(W, s, nd, dcore, len(core), E0, dC, norb, out)
        wfn = out.numpy()[:n]
        E0, X = diag(ctx, W, wfn)
    Xd = np.concatenate(feats)
    y = np.concatenate(targ).astype(np.float32)
    mu, sd = float(y.mean()), float(y.std())
    yn = ((y - mu) / sd).astype(np.float32)
    net = _capi.MLP(ctx, seed=3)
    untrained = _capi.MLP(ctx, seed=3)
    dX_all, dy_all = ctx.to_device(Xd), ctx.to_device(yn)
    first = net.train_step(W, dX_all, dy_all, len(yn), 0.0)
    train(ctx, net, W, Xd, yn, EPOCHS, BATCH, LR)
    last = net.train_step(W, dX_all, dy_all, len(yn), 0.0)
    assert last < 0.5 * first, (first, last)

    # the unseen step: plain vs screened from the same core
    nd = len(wfn)
    core, Cc = core_of(wfn, X)
    dcore, dC = ctx.to_device(core), ctx.to_device(Cc)
    out_p = ctx.empty((nd, W), np.uint64)
    n_p = ctx.asci_search(W, s, nd, dcore, len(core), E0, dC, norb, out_p)
    tp = ctx.timings()
    plain = out_p.numpy()[:n_p]
    pset = set(plain[:, 0].tolist())
    E_plain = diag(ctx, W, plain)[0]

    def screened(keep, m=None, scores=None):
        out = ctx.empty((nd, W), np.uint64)
        if m is not None:
            n = ctx.asci_search_screened(W, s, m, keep, nd, dcore, len(core), E0, dC, norb, out,
                                         scale=sd, shift=mu)
        else:
            n = ctx.asci_search_blocks(W, s, ctx.to_device(scores), keep, nd, dcore, len(core),
                                       E0, dC, norb, out)
        got = out.numpy()[:n]
        assert n == n_p
        return got, len(pset & set(got[:, 0].tolist())) / n_p, ctx.timings()

    # every block kept: the plain search
    got, rec, _ = screened(1.0, net)
    assert np.array_equal(got, plain)
    res = {}
    for keep in (0.05, 0.2):
        for name, m, sc in [("trained", net, None), ("untrained", untrained, None),
                            ("coeff-only", None, block_coeff_scores(core, Cc))]:
            got, rec, tm = screened(keep, m, sc)
            res[(name, keep)] = (rec, diag(ctx, W, got)[0] - E_plain, tm["asci_npairs"])
    lines = [f"{len(yn)} training blocks; plain step {tp['asci_total_ms']:.2f} ms, "
             f"{tp['asci_npairs']} raw candidates"]
    for (name, keep), (rec, de, npairs) in sorted(res.items()):
        lines.append(f"  keep {keep} {name:10s}: recall {rec:.4f}, dE {de * 1e3:+.3f} mHa, "
                     f"{npairs} raw candidates ({npairs / tp['asci_npairs']:.3f} of plain)")
    print("\n" + "\n".join(lines))
    for keep in (0.05, 0.2):
        rt, dt, pt = res[("trained", keep)]
        ru, du, _ = res[("untrained", keep)]
        # the screen removes work
        assert pt <= 3 * keep * tp["asci_npairs"]
        # the trained network beats the same network untrained, in recall and
        # in the variational energy's distance from the plain search
        assert rt > ru
        assert abs(dt) < abs(du)
