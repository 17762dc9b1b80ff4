"""Determinant-list helpers (restating include/macis/sd_operations.hpp).

Layout: W = 1 words per det for wfn_t<64> (alpha bits 0-31, beta bits
32-63); W = 2 for wfn_t<128> (word 0 alpha, word 1 beta).
"""
from __future__ import annotations

import itertools

import numpy as np


def generate_combs(nbits: int, nset: int) -> np.ndarray:
    """generate_combs (sd_operations.hpp:440-458).

    std::prev_permutation over the 0/1 vector v (v[i] = bit i) walks v in
    decreasing lexicographic order, which is the increasing lexicographic
    order of the occupied-index tuples, i.e. itertools.combinations order.
    """
    return np.array([sum(1 << i for i in c) for c in itertools.combinations(range(nbits), nset)],
                    dtype=np.uint64)


def generate_hilbert_space(W: int, norb: int, nalpha: int, nbeta: int) -> np.ndarray:
    """generate_hilbert_space (sd_operations.hpp:460-476): alpha outer, beta
    inner. Returns an (ndets, W) uint64 array."""
    a = generate_combs(norb, nalpha)
    b = generate_combs(norb, nbeta)
    if W == 1:
        d = (a[:, None] | (b[None, :] << np.uint64(32))).reshape(-1, 1)
    else:
        d = np.empty((len(a) * len(b), 2), dtype=np.uint64)
        d[:, 0] = np.repeat(a, len(b))
        d[:, 1] = np.tile(b, len(a))
    return np.ascontiguousarray(d)


def canonical_hf_determinant(W: int, nalpha: int, nbeta: int) -> np.ndarray:
    """canonical_hf_determinant (sd_operations.hpp:31-37)."""
    a = (1 << nalpha) - 1
    b = (1 << nbeta) - 1
    if W == 1:
        return np.array([[a | (b << 32)]], dtype=np.uint64)
    return np.array([[a, b]], dtype=np.uint64)


def split_spins(dets: np.ndarray, W: int):
    """(alpha, beta) uint64 arrays of a det list."""
    d = np.asarray(dets, dtype=np.uint64).reshape(-1, W)
    if W == 1:
        return d[:, 0] & np.uint64(0xFFFFFFFF), d[:, 0] >> np.uint64(32)
    return d[:, 0].copy(), d[:, 1].copy()


def to_canonical_string(det, W: int) -> str:
    """to_canonical_string (sd_operations.hpp:596-613): one character per
    orbital of the spin-string width (32 W), '2', 'u', 'd' or '0'."""
    d = np.asarray(det, dtype=np.uint64).reshape(-1)
    if W == 1:
        a, b, width = int(d[0]) & 0xFFFFFFFF, int(d[0]) >> 32, 32
    else:
        a, b, width = int(d[0]), int(d[1]), 64
    out = []
    for i in range(width):
        x, y = (a >> i) & 1, (b >> i) & 1
        out.append("2" if x and y else "u" if x else "d" if y else "0")
    return "".join(out)


def from_canonical_string(s: str, W: int) -> np.ndarray:
    """from_canonical_string (sd_operations.hpp:615-630) -> one det (W words);
    orbitals past the string are empty."""
    width = 32 * W
    if len(s) > width:
        raise ValueError("canonical string longer than the spin-string width")
    a = sum(1 << i for i, c in enumerate(s) if c in "2u")
    b = sum(1 << i for i, c in enumerate(s) if c in "2d")
    if W == 1:
        return np.array([a | (b << 32)], dtype=np.uint64)
    return np.array([a, b], dtype=np.uint64)
