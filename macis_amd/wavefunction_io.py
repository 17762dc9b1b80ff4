"""ASCII wavefunction files (include/macis/wavefunction_io.hpp): a header
"nstates norb nalpha nbeta", then "coefficient canonical-string" per line,
the coefficient as %30.16e."""
from __future__ import annotations

import numpy as np

from . import dets as D


def read_wavefunction(fname, W: int):
    """read_wavefunction -> (dets (n, W) uint64, coefficients)."""
    with open(fname) as f:
        head = f.readline().split()
        int(head[0]), int(head[1]), int(head[2]), int(head[3])
        states, coeffs = [], []
        for line in f:
            tok = line.split()
            if len(tok) < 2:
                continue
            coeffs.append(float(tok[0]))
            states.append(D.from_canonical_string(tok[1], W))
    return np.array(states, dtype=np.uint64).reshape(-1, W), np.array(coeffs)


def write_wavefunction(fname, norb: int, dets, coeffs, W: int):
    """write_wavefunction: nalpha / nbeta from the first determinant."""
    dets = np.asarray(dets, dtype=np.uint64).reshape(-1, W)
    coeffs = np.asarray(coeffs, dtype=np.float64)
    if len(dets) != len(coeffs):
        raise RuntimeError("Invalid Wave Function Dimensions")
    if not len(dets):
        return
    a, b = D.split_spins(dets[:1], W)
    with open(fname, "w") as f:
        f.write(f"{len(dets)} {norb} {bin(int(a[0])).count('1')} {bin(int(b[0])).count('1')}\n")
        for d, c in zip(dets, coeffs):
            f.write(f"{c:30.16e} {D.to_canonical_string(d, W)} \n")
