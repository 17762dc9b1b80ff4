"""How far a screen that acts before the accumulation can go (oracle, CPU).

asci_search ranks determinants by their ACCUMULATED score, the sum of the
contributions of every core determinant (determinant_search.hpp:498-629).
A screen that drops candidates before that sum is formed (the block screen,
or any per-candidate predictor of the raw score) decides on partial
information. On H2O cc-pVDZ (core 100 of the 6400-determinant wavefunction,
search for 6400) these tests pin three facts that bound every such screen:

- the plain top-k is spread over most excitation blocks: more than half of
  the blocks hold at least one of its determinants, so a screen keeping 5 %
  of the blocks cannot recover the top-k however it ranks them;
- even the exact raw score of every contribution does not suffice: keeping
  the 5 % largest |rv| contributions and accumulating only those recovers
  well under 90 % of the top-k;
- ranking blocks by |C_core| alone recovers less than 70 % at 5 %.
"""
import numpy as np
import pytest

from macis_amd import dets as D
from oracle import oracle as O


def _occ(x):
    return [i for i in range(64) if (int(x) >> i) & 1]


def _ctz(x):
    return np.log2((x & (~x + np.uint64(1))).astype(np.float64)).astype(np.int64)


def _block_ids(ca, cb, sa, sb):
    """Block number of each candidate (sa, sb) of the core det (ca, cb), in
    the order [alpha singles | beta singles | aa pairs | bb pairs | ab pairs]."""
    oa, ob = _occ(ca), _occ(cb)
    na, nb = len(oa), len(ob)
    pa, pb = na * (na - 1) // 2, nb * (nb - 1) // 2
    pos_a = np.full(64, -1)
    pos_a[oa] = np.arange(na)
    pos_b = np.full(64, -1)
    pos_b[ob] = np.arange(nb)
    ha, hb = np.uint64(ca) & ~sa, np.uint64(cb) & ~sb
    ka, kb = np.bitwise_count(ha), np.bitwise_count(hb)

    def lex(x, y, m):
        return x * m - x * (x + 1) // 2 + (y - x - 1)

    i1 = np.where(ka > 0, pos_a[_ctz(np.where(ka > 0, ha, 1))], 0)
    i2 = np.where(ka > 1, pos_a[_ctz(np.where(ka > 1, ha & (ha - np.uint64(1)), 1))], 0)
    j1 = np.where(kb > 0, pos_b[_ctz(np.where(kb > 0, hb, 1))], 0)
    j2 = np.where(kb > 1, pos_b[_ctz(np.where(kb > 1, hb & (hb - np.uint64(1)), 1))], 0)
    out = np.select(
        [(ka == 1) & (kb == 0), (ka == 0) & (kb == 1), ka == 2, kb == 2],
        [i1, na + j1, na + nb + lex(i1, i2, na), na + nb + pa + lex(j1, j2, nb)],
        na + nb + pa + pb + i1 * nb + j1)
    return out, na + nb + pa + pb + na * nb


@pytest.fixture(scope="module")
def step(h2o_oracle):
    orc = h2o_oracle
    s = O.ASCISettings.defaults()
    hf = D.canonical_hf_determinant(1, 5, 5)
    E = orc.matrix_element(1, hf, hf)
    w, X = hf, np.array([1.0])
    for nd in (100, 800, 6400):
        E, w, X = orc.asci_iter(1, s, nd, E, w, X)
    order = np.argsort(-np.abs(X), kind="stable")
    core, C = w[order[:100]], X[order[:100]]
    nd = 6400
    plain, _ = orc.asci_search(1, s, nd, core, E, C)
    top = np.unique(plain[: nd - len(core), 0])
    # every contribution with its core det's block number (global)
    states, rvs, bids, base = [], [], [], 0
    for c in range(len(core)):
        st, rv = orc.asci_contributions(1, s, core[c:c + 1], E, C[c:c + 1])
        d = int(core[c, 0])
        b, nbk = _block_ids(d & 0xFFFFFFFF, d >> 32, st[:, 0] & np.uint64(0xFFFFFFFF),
                            st[:, 0] >> np.uint64(32))
        states.append(st[:, 0])
        rvs.append(rv)
        bids.append(base + b)
        base += nbk
    return (core, C, top, np.concatenate(states), np.concatenate(rvs), np.concatenate(bids), base,
            E)


def _blocks_per_core(core):
    return [_block_ids(int(d) & 0xFFFFFFFF, int(d) >> 32, np.zeros(0, np.uint64),
                       np.zeros(0, np.uint64))[1] for d in core[:, 0]]


def _recall(step, keep_mask):
    core, C, top, states, rv, bids, nblk, _ = step
    keys, inv = np.unique(states[keep_mask], return_inverse=True)
    acc = np.abs(np.bincount(inv, weights=rv[keep_mask]))
    sel = ~np.isin(keys, core[:, 0]) & (acc != 0)
    keys, acc = keys[sel], acc[sel]
    best = keys[np.lexsort((keys, -acc))[: len(top)]]
    return np.isin(best, top).sum() / len(top)


def test_every_block_kept_recovers_everything(step):
    assert _recall(step, np.ones(len(step[4]), bool)) == 1.0


def test_top_k_is_spread_over_most_blocks(step):
    core, C, top, states, rv, bids, nblk, _ = step
    holding = np.unique(bids[np.isin(states, top)])
    assert len(holding) / nblk > 0.5


def test_exact_raw_scores_do_not_suffice(step):
    rv = step[4]
    k = int(np.ceil(0.05 * len(rv)))
    m = np.zeros(len(rv), bool)
    m[np.argsort(-np.abs(rv), kind="stable")[:k]] = True
    assert _recall(step, m) < 0.9


def test_block_numbers_agree_with_the_oracle_blocks(step, h2o_oracle):
    core, C, top, states, rv, bids, nblk, E = step
    _, t = h2o_oracle.asci_blocks(1, O.ASCISettings.defaults(), core, E)
    assert len(t) == nblk
    mx = np.full(nblk, -np.inf)
    np.maximum.at(mx, bids, np.log(np.abs(rv)))
    lc = np.repeat(np.log(np.abs(C)), _blocks_per_core(core))
    np.testing.assert_allclose(mx, t + lc, rtol=1e-12, atol=1e-12)


def test_coefficient_ranking_of_blocks(step):
    core, C, top, states, rv, bids, nblk, _ = step
    score = np.repeat(np.log(np.abs(C)), _blocks_per_core(core))
    kept = np.zeros(nblk, bool)
    kept[np.argsort(-score, kind="stable")[: int(np.ceil(0.05 * nblk))]] = True
    assert _recall(step, kept[bids]) < 0.7
