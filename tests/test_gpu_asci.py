"""GPU parity of the ASCI search (K6-K8) against the oracle restatement of
asci_search (determinant_search.hpp:394-663).

Top-k membership is compared as sets; determinants may only differ inside a
tie band at the k-th |rv| (the accumulation order of duplicate scores is
implementation-defined in the reference as well, determinant_sort.hpp:50-56).
"""
import numpy as np
import pytest

from macis_amd import _capi
from macis_amd import dets as D
from macis_amd.synthetic import seeded_integrals, selected_ci_dets
from oracle import oracle as O

from conftest import announce

pytestmark = pytest.mark.gpu


def _rows(a):
    return {tuple(r) for r in np.asarray(a).reshape(len(a), -1).tolist()}


def run_both(ctx, orc, W, norb, core, C, E0, ndets_max, **kw):
    s_orc = O.ASCISettings.defaults(**kw)
    s_gpu = _capi.ASCISettings.defaults(**kw)
    ref, ref_rv = orc.asci_search(W, s_orc, ndets_max, core, E0, C)
    out = ctx.empty((ndets_max, W), np.uint64)
    n = ctx.asci_search(W, s_gpu, ndets_max, ctx.to_device(core), len(core), E0,
                        ctx.to_device(C), norb, out)
    got = out.numpy()[:n]
    return got, ref, ref_rv


def check_topk(got, ref, ref_rv, ncore, rel=1e-10):
    assert len(got) == len(ref)
    # core dets last, in input order
    assert np.array_equal(got[len(got) - ncore:], ref[len(ref) - ncore:])
    g, r = _rows(got[: len(got) - ncore]), _rows(ref[: len(ref) - ncore])
    diff = g ^ r
    if diff:
        kth = np.abs(ref_rv).min()
        band = {tuple(x) for x, v in zip(ref[: len(ref) - ncore].tolist(), ref_rv)
                if abs(abs(v) - kth) <= rel * kth}
        assert len(diff) <= 2 * len(band) and (r - g) <= band, "top-k differs outside the tie band"
    return len(diff)


def test_asci_search_h2o_from_hf(ctx, h2o, h2o_oracle):
    announce("asci_search H2O from HF, ndets_max 100")
    norb, _, T, V = h2o
    ctx.set_integrals(norb, T, V)
    hf = D.canonical_hf_determinant(1, 5, 5)
    E0 = h2o_oracle.matrix_element(1, hf, hf)
    got, ref, ref_rv = run_both(ctx, h2o_oracle, 1, norb, hf, np.array([1.0]), E0, 100)
    check_topk(got, ref, ref_rv, 1)
    t = ctx.timings()
    assert t["asci_npairs"] > 0 and t["asci_nunique"] <= t["asci_npairs"]


def test_asci_search_h2o_grown_core(ctx, h2o, h2o_oracle):
    announce("asci_search H2O, 100 core dets of an 800-det wavefunction")
    norb, _, T, V = h2o
    ctx.set_integrals(norb, T, V)
    hf = D.canonical_hf_determinant(1, 5, 5)
    E0 = h2o_oracle.matrix_element(1, hf, hf)
    s = O.ASCISettings.defaults()
    E1, w, X = h2o_oracle.asci_iter(1, s, 100, E0, hf, np.array([1.0]))
    E2, w, X = h2o_oracle.asci_iter(1, s, 800, E1, w, X)
    order = np.argsort(-np.abs(X), kind="stable")
    core, C = w[order[:100]], X[order[:100]]
    got, ref, ref_rv = run_both(ctx, h2o_oracle, 1, norb, core, C, E2, 6400)
    check_topk(got, ref, ref_rv, 100)
    # just_singles path
    got, ref, ref_rv = run_both(ctx, h2o_oracle, 1, norb, core, C, E2, 1000, just_singles=1)
    check_topk(got, ref, ref_rv, 100)


def test_asci_search_128bit(ctx):
    announce("asci_search 64 orbitals 20a20b W=2, 30 core dets")
    n = 64
    T, V = seeded_integrals(n, 31)
    ctx.set_integrals(n, T, V)
    orc = O.Ham(T, V)
    core = selected_ci_dets(n, 20, 20, 30, seed=4, T=T)
    rng = np.random.default_rng(5)
    C = rng.standard_normal(30)
    C /= np.linalg.norm(C)
    E0 = orc.matrix_element(2, core[0], core[0]) - 1.0
    got, ref, ref_rv = run_both(ctx, orc, 2, n, core, C, E0, 20000)
    check_topk(got, ref, ref_rv, 30)


def test_asci_search_everything_fits(ctx, h2o, h2o_oracle):
    announce("asci_search with ndets_max above the candidate count")
    norb, _, T, V = h2o
    ctx.set_integrals(norb, T, V)
    hf = D.canonical_hf_determinant(1, 5, 5)
    E0 = h2o_oracle.matrix_element(1, hf, hf)
    got, ref, _ = run_both(ctx, h2o_oracle, 1, norb, hf, np.array([1.0]), E0, 100000)
    assert np.array_equal(got, ref)  # det-sorted order when no top-k is needed


def _prune_events(counts, tiny, pmax):
    """Core-det indices of the reference's prune events up to the first one
    whose stage 1 leaves the list above pair_size_max (stage 2), from the
    per-core-det pair and small-score counts (determinant_search.hpp:120-137)."""
    S, since, events = 0, 0, []
    for d, (c, t) in enumerate(zip(counts, tiny)):
        S += c
        since += t
        if S <= pmax:
            continue
        S -= since
        since = 0
        events.append(d)
        if S > pmax:
            return events, True
    return events, False


PRUNE_CASES = [(0.6, 0.6), (0.4, 0.6), (0.25, 0.3), (0.15, 0.5), (0.08, 0.3)]


@pytest.mark.parametrize("stream", ["auto", "0", "1"])
@pytest.mark.parametrize("frac,q", PRUNE_CASES)
def test_asci_search_pair_size_max_pruning(ctx, h2o, h2o_oracle, frac, q, stream, monkeypatch):
    """pair_size_max below the raw candidate count: the GPU replays the
    reference's prunes (stage 1: erase |rv| <= rv_prune_tol; stage 2: sort
    and accumulate; later prunes test accumulated scores) and its top-k
    matches the oracle's asci_search, which runs the reference loop. The
    cases cover prunes with stage 1 only and with stage 2, and the pruning
    must change the result against the unpruned search. Stage 2 runs over
    the slots of its event or, for events over a large share of the pairs,
    streams the sorted list (MHIP_PRUNE_STREAM forces either); the result is
    the same."""
    announce(f"asci_search pruning pair_size_max = {frac} x pairs, tol = q{q}, stream {stream}")
    if stream != "auto":
        monkeypatch.setenv("MHIP_PRUNE_STREAM", stream)
    norb, _, T, V = h2o
    ctx.set_integrals(norb, T, V)
    hf = D.canonical_hf_determinant(1, 5, 5)
    E0 = h2o_oracle.matrix_element(1, hf, hf)
    s = O.ASCISettings.defaults()
    E1, w, X = h2o_oracle.asci_iter(1, s, 100, E0, hf, np.array([1.0]))
    E2, w, X = h2o_oracle.asci_iter(1, s, 800, E1, w, X)
    order = np.argsort(-np.abs(X), kind="stable")
    core, C = w[order[:100]], X[order[:100]]
    per = [h2o_oracle.asci_contributions(1, s, core[i:i + 1], E2, C[i:i + 1])[1]
           for i in range(len(core))]
    counts = [len(r) for r in per]
    rv_all = np.abs(np.concatenate(per))
    tol = float(np.quantile(rv_all, q))
    pmax = int(frac * len(rv_all))
    tiny = [int((np.abs(r) <= tol).sum()) for r in per]
    events, stage2 = _prune_events(counts, tiny, pmax)
    assert events, "the case must prune"
    got, ref, ref_rv = run_both(ctx, h2o_oracle, 1, norb, core, C, E2, 3000,
                                pair_size_max=pmax, rv_prune_tol=tol)
    check_topk(got, ref, ref_rv, 100)
    assert ctx.timings()["asci_npairs"] == len(rv_all)
    full, _, _ = run_both(ctx, h2o_oracle, 1, norb, core, C, E2, 3000)
    assert _rows(full) != _rows(got), "pruning left the top-k unchanged"


def test_asci_pruning_cases_cover_both_stages(h2o, h2o_oracle):
    """The parameter sets above include prunes with and without stage 2."""
    norb, _, T, V = h2o
    hf = D.canonical_hf_determinant(1, 5, 5)
    E0 = h2o_oracle.matrix_element(1, hf, hf)
    s = O.ASCISettings.defaults()
    E1, w, X = h2o_oracle.asci_iter(1, s, 100, E0, hf, np.array([1.0]))
    E2, w, X = h2o_oracle.asci_iter(1, s, 800, E1, w, X)
    order = np.argsort(-np.abs(X), kind="stable")
    core, C = w[order[:100]], X[order[:100]]
    per = [h2o_oracle.asci_contributions(1, s, core[i:i + 1], E2, C[i:i + 1])[1]
           for i in range(len(core))]
    rv_all = np.abs(np.concatenate(per))
    seen = set()
    for frac, q in PRUNE_CASES:
        tol = float(np.quantile(rv_all, q))
        tiny = [int((np.abs(r) <= tol).sum()) for r in per]
        events, stage2 = _prune_events([len(r) for r in per], tiny, int(frac * len(rv_all)))
        assert events
        seen.add(stage2)
    assert seen == {False, True}


def _as_void(a):
    """Rows of a (n, W) uint64 array as single opaque items (for unique/isin)."""
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a.view(np.dtype((np.void, 8 * a.shape[1]))).ravel()


def _accumulate(states, rv):
    """Unique determinants (as opaque rows, sorted), the sum of their scores
    in generation order and the sum of the terms' magnitudes."""
    keys, inv = np.unique(_as_void(states), return_inverse=True)
    inv = inv.reshape(-1)
    return keys, np.bincount(inv, weights=rv), np.bincount(inv, weights=np.abs(rv))


@pytest.mark.parametrize("W", [1, 2])
def test_asci_candidate_scores_match_oracle(ctx, h2o, h2o_oracle, W):
    """mhip_asci_candidates' accumulated |rv| per unique non-core
    determinant equals the oracle's asci_contributions
    (determinant_contributions.hpp:25-160) summed per determinant
    (determinant_sort.hpp:39-83): the same determinants, and each score to
    1e-12 of the sum of its terms' magnitudes (the order of a sum is the
    only freedom)."""
    if W == 1:
        announce("asci candidate scores, H2O 100-det core, W=1")
        norb, _, T, V = h2o
        orc = h2o_oracle
        hf = D.canonical_hf_determinant(1, 5, 5)
        E0 = orc.matrix_element(1, hf, hf)
        s = O.ASCISettings.defaults()
        E1, w, X = orc.asci_iter(1, s, 100, E0, hf, np.array([1.0]))
        E2, w, X = orc.asci_iter(1, s, 800, E1, w, X)
        order = np.argsort(-np.abs(X), kind="stable")
        core, C, E = w[order[:100]], X[order[:100]], E2
    else:
        announce("asci candidate scores, 64 orbitals 20a20b, 12 core dets, W=2")
        norb = 64
        T, V = seeded_integrals(norb, 31)
        orc = O.Ham(T, V)
        core = selected_ci_dets(norb, 20, 20, 12, seed=9, T=T)
        rng = np.random.default_rng(11)
        C = rng.standard_normal(len(core))
        C /= np.linalg.norm(C)
        E = orc.matrix_element(2, core[0], core[0]) - 0.5
    ctx.set_integrals(norb, T, V)
    states, rv = orc.asci_contributions(W, O.ASCISettings.defaults(), core, E, C)
    keys, acc, mag = _accumulate(states, rv)
    keep = ~np.isin(keys, _as_void(core)) & (acc != 0.0)
    keys, ref, mag = keys[keep], np.abs(acc[keep]), mag[keep]
    cap = len(keys) + 16
    d_out, a_out, n = ctx.asci_candidates(W, _capi.ASCISettings.defaults(), ctx.to_device(core),
                                          len(core), E, ctx.to_device(C), norb, cap)
    dets = d_out.numpy()[:n]
    sc = a_out.numpy()[:n]
    gk, gi = np.unique(_as_void(dets), return_index=True)
    assert len(gk) == n, "candidates are unique"
    assert np.array_equal(gk, keys), "candidate determinants differ"
    err = np.max(np.abs(sc[gi] - ref) / mag)
    assert err <= 1e-12, err


@pytest.mark.parametrize("tol", [np.inf, 0.0])
@pytest.mark.parametrize("frac", [0.1, 0.45])
def test_asci_search_pruning_limits(ctx, h2o, h2o_oracle, tol, frac):
    """pair_size_max pruning at the two tolerances whose outcome the loop's
    definition fixes (the oracle side is pinned in
    test_oracle_asci_pruning.py): +inf erases every pair at each event,
    0 erases none and only regroups the sums."""
    announce(f"asci_search pruning limits tol={tol} pair_size_max={frac} x pairs")
    norb, _, T, V = h2o
    ctx.set_integrals(norb, T, V)
    hf = D.canonical_hf_determinant(1, 5, 5)
    E0 = h2o_oracle.matrix_element(1, hf, hf)
    s = O.ASCISettings.defaults()
    E1, w, X = h2o_oracle.asci_iter(1, s, 100, E0, hf, np.array([1.0]))
    E2, w, X = h2o_oracle.asci_iter(1, s, 800, E1, w, X)
    order = np.argsort(-np.abs(X), kind="stable")
    core, C = w[order[:100]], X[order[:100]]
    npairs = len(h2o_oracle.asci_contributions(1, s, core, E2, C)[1])
    got, ref, ref_rv = run_both(ctx, h2o_oracle, 1, norb, core, C, E2, 3000,
                                pair_size_max=int(frac * npairs), rv_prune_tol=tol)
    check_topk(got, ref, ref_rv, 100)


@pytest.mark.parametrize("case", ["w1_pruned", "w1_singles", "w2"])
def test_asci_key_forms_agree(ctx, h2o, h2o_oracle, monkeypatch, case):
    """The search sorts either the full determinant strings or dense keys
    (rank of the beta string, rank of the alpha string among all strings one
    or two excitations from a core string); both orders are bitset_less, so
    the selected determinants, in order, and the exported candidates with
    their scores are identical. MHIP_ASCI_KEYS forces one form."""
    announce(f"asci_search dense vs full-string keys: {case}")
    kw = {}
    if case == "w2":
        W, norb = 2, 64
        T, V = seeded_integrals(norb, 31)
        orc = O.Ham(T, V)
        core = selected_ci_dets(norb, 20, 20, 30, seed=4, T=T)
        C = np.random.default_rng(5).standard_normal(30)
        C /= np.linalg.norm(C)
        E = orc.matrix_element(2, core[0], core[0]) - 1.0
        ndets_max = 20000
    else:
        W = 1
        norb, _, T, V = h2o
        hf = D.canonical_hf_determinant(1, 5, 5)
        E0 = h2o_oracle.matrix_element(1, hf, hf)
        s = O.ASCISettings.defaults()
        E1, w, X = h2o_oracle.asci_iter(1, s, 100, E0, hf, np.array([1.0]))
        E, w, X = h2o_oracle.asci_iter(1, s, 800, E1, w, X)
        order = np.argsort(-np.abs(X), kind="stable")
        core, C = w[order[:100]], X[order[:100]]
        ndets_max = 3000
        if case == "w1_singles":
            kw["just_singles"] = 1
        else:
            ctx.set_integrals(norb, T, V)
            # pair_size_max at a quarter of the raw pairs: prune events replayed
            out = ctx.empty((ndets_max, W), np.uint64)
            ctx.asci_search(W, _capi.ASCISettings.defaults(), ndets_max, ctx.to_device(core),
                            len(core), E, ctx.to_device(C), norb, out)
            npairs = ctx.timings()["asci_npairs"]
            kw["pair_size_max"] = npairs // 4
    ctx.set_integrals(norb, T, V)
    s_gpu = _capi.ASCISettings.defaults(**kw)
    dcore, dC = ctx.to_device(core), ctx.to_device(C)
    res = {}
    for keys in ("rank", "bits"):
        monkeypatch.setenv("MHIP_ASCI_KEYS", keys)
        out = ctx.empty((ndets_max, W), np.uint64)
        n = ctx.asci_search(W, s_gpu, ndets_max, dcore, len(core), E, dC, norb, out)
        sel = out.numpy()[:n].copy()
        cap = ctx.timings()["asci_npairs"] + 16
        d_out, a_out, m = ctx.asci_candidates(W, s_gpu, dcore, len(core), E, dC, norb, cap)
        res[keys] = (sel, d_out.numpy()[:m].copy(), a_out.numpy()[:m].copy())
    for a, b in zip(res["rank"], res["bits"]):
        assert np.array_equal(a, b)
    assert len(res["rank"][0]) == ndets_max
