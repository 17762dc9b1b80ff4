"""Seeded synthetic integrals and determinant lists for the benchmark
workloads (CAS(10e,10o), CAS(12e,12o), the 64-orbital selected-CI lists and
the CAS-CI batch).

Integral recipe (one fixed spec so every run sees identical inputs):
  * RNG: numpy Generator(PCG64(seed)), standard_normal draws.
  * h_pq = h_qp = z * exp(-|p-q|/3) for p <= q, drawn row by row (p outer,
    q >= p inner); then h_pp += -2*(n - p) with 0-based p.
  * For k = 0 .. 3n-1: L^k symmetric, L^k_pq = L^k_qp = 0.3 * z (0.3 is the
    standard deviation), drawn row by row over p <= q.
  * V[pq, rs] = sum_k L^k_pq * L^k_rs accumulated in k order as elementwise
    outer products, so the 8-fold symmetry holds bit for bit and V is
    positive semidefinite as a pq x rs matrix.
Arrays are flat column-major: T(p,q) = T[p+n*q], V(p,q,r,s) =
V[p + n*q + n^2*r + n^3*s].
"""
from __future__ import annotations

import math

import numpy as np


def seeded_integrals(n: int, seed: int):
    rng = np.random.Generator(np.random.PCG64(seed))
    iu = np.triu_indices(n)
    T = np.zeros((n, n))
    z = rng.standard_normal(len(iu[0]))
    T[iu] = z * np.exp(-np.abs(iu[0] - iu[1]) / 3.0)
    T = np.triu(T) + np.triu(T, 1).T
    T[np.arange(n), np.arange(n)] += -2.0 * (n - np.arange(n))
    M = np.zeros((n * n, n * n))
    for _ in range(3 * n):
        L = np.zeros((n, n))
        L[iu] = 0.3 * rng.standard_normal(len(iu[0]))
        L = np.triu(L) + np.triu(L, 1).T
        l = L.ravel(order="F")
        M += np.multiply.outer(l, l)
    return np.ascontiguousarray(T.ravel(order="F")), np.ascontiguousarray(M.ravel(order="F"))


def _gumbel_choose(rng, logw, k):
    """Per row, choose k[i] distinct columns with probability proportional to
    exp(logw) (Gumbel top-k). Returns a boolean (rows, cols) mask."""
    g = logw[None, :] - np.log(-np.log(rng.random((len(k), len(logw)))))
    mask = np.zeros(g.shape, dtype=bool)
    rows = np.arange(len(k))
    for j in range(int(k.max(initial=0))):  # k[i] largest, one argmax at a time
        idx = np.argmax(g, axis=1)
        take = k > j
        mask[rows[take], idx[take]] = True
        g[rows, idx] = -np.inf
    return mask


def selected_ci_dets(n: int, nalpha: int, nbeta: int, ndets: int, seed: int,
                     T=None, weights=(0.1, 0.5, 0.2, 0.2), batch: int = 400_000,
                     energies=None):
    """Seeded selected-CI determinant list for wfn_t<128> (W = 2).

    The HF determinant plus random 1-4-fold excitations of it: the level L
    is drawn with probabilities `weights`; L_alpha ~ Binomial(L, 1/2) clipped
    to what each spin allows, L_beta = L - L_alpha; holes and particles are
    drawn without replacement with probability proportional to
    exp(-|e_p - e_F|), e_p = h_pp (one-body diagonal, or p when T is None) and
    e_F halfway between the highest occupied and lowest virtual e_p of each
    spin; `energies` (length n) overrides e_p. With the seeded h_pp (spacing
    about 2 Ha from the diagonal offset) only a few orbitals around e_F carry
    weight and a few 1e4 determinants are reachable, so lists of millions use
    `energies` with a finer spacing. Unique dets are collected until `ndets`
    (HF included), sorted by
    bitset_less (beta word most significant, then alpha) and truncated.
    Returns an (ndets, 2) uint64 array [alpha, beta].
    """
    assert n <= 64
    rng = np.random.Generator(np.random.PCG64(seed))
    if energies is not None:
        e = np.asarray(energies, dtype=float).copy()
    else:
        e = np.arange(n, dtype=float) if T is None else np.asarray(T).reshape(n, n, order="F").diagonal().copy()
    order = np.argsort(e, kind="stable")
    def spin_setup(ne):
        occ = order[:ne]
        vir = order[ne:]
        eF = 0.5 * (e[occ].max() + e[vir].min()) if 0 < ne < n else 0.0
        return occ, vir, -np.abs(e[occ] - eF), -np.abs(e[vir] - eF)
    oa, va, lwoa, lwva = spin_setup(nalpha)
    ob, vb, lwob, lwvb = spin_setup(nbeta)
    hf_a = np.uint64(sum(1 << int(i) for i in oa))
    hf_b = np.uint64(sum(1 << int(i) for i in ob))
    bits_oa = (np.uint64(1) << oa.astype(np.uint64))
    bits_va = (np.uint64(1) << va.astype(np.uint64))
    bits_ob = (np.uint64(1) << ob.astype(np.uint64))
    bits_vb = (np.uint64(1) << vb.astype(np.uint64))
    maxa = min(nalpha, n - nalpha)
    maxb = min(nbeta, n - nbeta)
    seen = {(int(hf_a), int(hf_b))}
    a_all = [np.array([hf_a])]
    b_all = [np.array([hf_b])]
    total = 1
    while total < ndets:
        M = int(min(batch, max(1000, 2 * (ndets - total))))
        L = rng.choice(4, size=M, p=np.asarray(weights) / np.sum(weights)) + 1
        la = rng.binomial(L, 0.5)
        la = np.minimum(la, maxa)
        lb = np.minimum(L - la, maxb)
        la = np.minimum(L - lb, maxa)
        ha = _gumbel_choose(rng, lwoa, la)
        pa = _gumbel_choose(rng, lwva, la)
        hb = _gumbel_choose(rng, lwob, lb)
        pb = _gumbel_choose(rng, lwvb, lb)
        xa = np.bitwise_or.reduce(np.where(ha, bits_oa, np.uint64(0)), axis=1) | \
            np.bitwise_or.reduce(np.where(pa, bits_va, np.uint64(0)), axis=1)
        xb = np.bitwise_or.reduce(np.where(hb, bits_ob, np.uint64(0)), axis=1) | \
            np.bitwise_or.reduce(np.where(pb, bits_vb, np.uint64(0)), axis=1)
        A = hf_a ^ xa
        B = hf_b ^ xb
        key = np.stack([B, A], axis=1)
        _, first = np.unique(key, axis=0, return_index=True)
        first.sort()
        newa, newb = [], []
        for i in first:
            t = (int(A[i]), int(B[i]))
            if t not in seen:
                seen.add(t)
                newa.append(A[i])
                newb.append(B[i])
                total += 1
                if total >= ndets:
                    break
        a_all.append(np.array(newa, dtype=np.uint64))
        b_all.append(np.array(newb, dtype=np.uint64))
    a = np.concatenate(a_all)[:ndets]
    b = np.concatenate(b_all)[:ndets]
    srt = np.lexsort((a, b))
    return np.ascontiguousarray(np.stack([a[srt], b[srt]], axis=1))


def cas_batch_instances(count: int = 256, seed: int = 2024, sizes_from=(6, 8, 10)):
    """The batched CAS-CI workload: `count` instances of CAS(n e, n o) with n
    drawn uniformly (seeded) from `sizes_from`, n/2 alpha + n/2 beta
    electrons, the complete determinant list of generate_hilbert_space
    (wfn_t<64> words) and integrals seeded_integrals(n, 1000 + i) for
    instance i. Returns (norbs, det_lists, Ts, Vs)."""
    from . import dets as D

    rng = np.random.default_rng(seed)
    sizes = rng.choice(list(sizes_from), size=count)
    cache, norbs, dl, Ts, Vs = {}, [], [], [], []
    for i, n in enumerate(sizes):
        n = int(n)
        T, V = seeded_integrals(n, 1000 + i)
        if n not in cache:
            cache[n] = D.generate_hilbert_space(1, n, n // 2, n // 2)[:, 0]
        norbs.append(n)
        dl.append(cache[n])
        Ts.append(T)
        Vs.append(V)
    return norbs, dl, Ts, Vs


def perturbed_diagonal_guess(D, delta: float = 1e-3, seed: int = 0):
    """Davidson guess that reaches the lowest eigenvalue of any spin sector:
    the unit vector at the first minimum of D (diagonal_guess,
    davidson.hpp:61-68) plus `delta` times a seeded uniform unit vector,
    normalised. With equal alpha and beta counts a closed-shell unit vector
    is even under the alpha <-> beta exchange that commutes with H, and the
    diagonal preconditioner keeps that parity, so from the plain diagonal
    guess Davidson converges to the lowest even state; when the ground state
    is odd, only round-off can lead it there. The seeded component gives
    every sector a weight of order delta, far above any residual tolerance.
    """
    D = np.asarray(D, dtype=np.float64)
    rng = np.random.Generator(np.random.PCG64(seed))
    r = rng.uniform(-1.0, 1.0, size=D.size)
    X = delta * r / np.linalg.norm(r)
    X[int(np.argmin(D))] += 1.0
    return X / np.linalg.norm(X)
