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
