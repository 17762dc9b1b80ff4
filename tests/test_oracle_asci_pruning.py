"""Pins the oracle's pair_size_max pruning branch (determinant_search.hpp:
120-137) by what the loop's definition fixes, independently of how it is
coded:

- rv_prune_tol = +inf: every prune event erases all pairs gathered so far
  (stage 1 keeps only |rv| > tol), so the pairs that reach the sort are those
  of the core determinants after the last event. The search must then equal,
  bit for bit, an unpruned search whose coefficients are zero for the core
  determinants up to that event: their pairs carry rv = +-0, add nothing to
  any accumulated sum and leave with the rv < 0 filter (:542-544), while the
  full core list still excludes and closes the result (:533-536, :653).
- rv_prune_tol = 0: stage 1 erases nothing (no pair has rv exactly 0), so
  each event runs stage 2 (sort and accumulate) when the list stays above
  pair_size_max; accumulating in stages gives the unpruned sums up to their
  summation order, hence the same top-k and scores to rounding.

The event positions are derived here from per-core-determinant pair counts,
not taken from the oracle's loop.
"""
import numpy as np
import pytest

from macis_amd import dets as D
from oracle import oracle as O


@pytest.fixture(scope="module")
def h2o_core(h2o, h2o_oracle):
    norb, _, T, V = h2o
    hf = D.canonical_hf_determinant(1, 5, 5)
    E0 = h2o_oracle.matrix_element(1, hf, hf)
    s = O.ASCISettings.defaults()
    E1, w, X = h2o_oracle.asci_iter(1, s, 100, E0, hf, np.array([1.0]))
    E2, w, X = h2o_oracle.asci_iter(1, s, 400, E1, w, X)
    order = np.argsort(-np.abs(X), kind="stable")
    core, C = w[order[:40]], X[order[:40]]
    counts = [len(h2o_oracle.asci_contributions(1, s, core[i:i + 1], E2, C[i:i + 1])[1])
              for i in range(len(core))]
    return core, C, E2, np.array(counts)


def last_event_erase_all(counts, pmax):
    """Index of the last core determinant after which the list exceeds
    pmax, when each such event empties the list (-1 if none)."""
    S, last = 0, -1
    for d, c in enumerate(counts):
        S += c
        if S > pmax:
            S, last = 0, d
    return last


@pytest.mark.parametrize("frac", [0.05, 0.2, 0.45, 0.8])
def test_pruning_with_infinite_tolerance_drops_early_core_dets(h2o_oracle, h2o_core, frac):
    core, C, E2, counts = h2o_core
    pmax = int(frac * counts.sum())
    last = last_event_erase_all(counts, pmax)
    assert last >= 0
    s_pruned = O.ASCISettings.defaults(pair_size_max=pmax, rv_prune_tol=np.inf)
    got, got_rv = h2o_oracle.asci_search(1, s_pruned, 2000, core, E2, C)
    Cz = C.copy()
    Cz[: last + 1] = 0.0
    ref, ref_rv = h2o_oracle.asci_search(1, O.ASCISettings.defaults(), 2000, core, E2, Cz)
    assert np.array_equal(got, ref)
    assert np.array_equal(got_rv.view(np.uint64), ref_rv.view(np.uint64))
    # a mistake in the event placement changes the result
    if last + 1 < len(core):
        Cz2 = C.copy()
        Cz2[: last + 2] = 0.0
        alt, _ = h2o_oracle.asci_search(1, O.ASCISettings.defaults(), 2000, core, E2, Cz2)
        assert not np.array_equal(alt, got)


@pytest.mark.parametrize("frac", [0.1, 0.3, 0.6])
def test_pruning_with_zero_tolerance_only_regroups_sums(h2o_oracle, h2o_core, frac):
    core, C, E2, counts = h2o_core
    pmax = int(frac * counts.sum())
    s_pruned = O.ASCISettings.defaults(pair_size_max=pmax, rv_prune_tol=0.0)
    got, got_rv = h2o_oracle.asci_search(1, s_pruned, 2000, core, E2, C)
    ref, ref_rv = h2o_oracle.asci_search(1, O.ASCISettings.defaults(), 2000, core, E2, C)
    assert len(got) == len(ref)
    assert np.array_equal(got[-len(core):], core)
    g = {tuple(r): v for r, v in zip(got[: len(got_rv)].tolist(), got_rv)}
    r = {tuple(x): v for x, v in zip(ref[: len(ref_rv)].tolist(), ref_rv)}
    kth = np.abs(ref_rv).min()
    for k in set(g) ^ set(r):  # only ties at the k-th score may swap
        v = g.get(k, r.get(k))
        assert abs(abs(v) - kth) <= 1e-12 * kth
    for k in set(g) & set(r):
        assert abs(g[k] - r[k]) <= 1e-12 * abs(r[k])
