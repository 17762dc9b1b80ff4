"""FCIDUMP reader (restates src/macis/fcidump.cxx:78-182).

Lines with five tokens are integrals: either "p q r s value" or
"value p q r s" (fcidump.cxx:36-62). Indices are 1-based; all-zero indices
give the core energy, p,q nonzero with r=s=0 a one-body term, four nonzero a
two-body term that is unpacked with 8-fold symmetry (fcidump.cxx:141-174).
Arrays are returned flat in column-major order, T(p,q) = T[p + n*q],
V(p,q,r,s) = V[p + n*q + n^2*r + n^3*s].
"""
from __future__ import annotations

import numpy as np


def _lines(fname):
    with open(fname) as f:
        for line in f:
            tok = line.split()
            if len(tok) != 5:
                continue
            idx_first = any(c.isalpha() or c == "." for c in tok[-1])
            int_first = any(c.isalpha() or c == "." for c in tok[0])
            if idx_first and int_first:
                raise RuntimeError("Invalid FCIDUMP Line")
            if idx_first:
                p, q, r, s = (int(t) for t in tok[:4])
                val = float(tok[4])
            else:
                p, q, r, s = (int(t) for t in tok[1:])
                val = float(tok[0])
            if min(p, q, r, s) < 0:
                raise RuntimeError("Invalid Orb Idx")
            yield p, q, r, s, val


def read_fcidump(fname):
    """Return (norb, E_core, T, V) with T, V flat column-major float64."""
    rows = list(_lines(fname))
    norb = max(max(p, q, r, s) for p, q, r, s, _ in rows)
    n = norb
    T = np.zeros(n * n)
    V = np.zeros(n**4)
    E_core = 0.0
    core_seen = False
    for p, q, r, s, val in rows:
        if not (p or q or r or s):
            if not core_seen:  # read_fcidump_core returns the first core line
                E_core = val
                core_seen = True
        elif p and q and r and s:
            p, q, r, s = p - 1, q - 1, r - 1, s - 1
            for a, b, c, d in ((p, q, r, s), (p, q, s, r), (q, p, r, s), (q, p, s, r),
                               (r, s, p, q), (s, r, p, q), (r, s, q, p), (s, r, q, p)):
                V[a + n * b + n * n * c + n**3 * d] = val
        else:
            p, q = p - 1, q - 1
            T[p + n * q] = val
            T[q + n * p] = val
    return norb, E_core, T, V


def read_rdms_binary(fname, norb):
    """read_rdms_binary (fcidump.cxx:209-251): int32 norb, then the 1-RDM
    (norb^2) and 2-RDM (norb^4) column-major doubles -> (D1, D2) flat."""
    with open(fname, "rb") as f:
        n = int(np.frombuffer(f.read(4), dtype=np.int32)[0])
        if n != norb:
            raise RuntimeError(f"NORB in RDM file doesn't match {norb} {n}")
        d1 = np.frombuffer(f.read(8 * n * n), dtype=np.float64).copy()
        d2 = np.frombuffer(f.read(8 * n ** 4), dtype=np.float64).copy()
    return d1, d2


def write_rdms_binary(fname, norb, D1, D2):
    """write_rdms_binary (fcidump.cxx:253-280), flat column-major inputs."""
    with open(fname, "wb") as f:
        f.write(np.int32(norb).tobytes())
        f.write(np.ascontiguousarray(D1, dtype=np.float64).reshape(-1)[: norb * norb].tobytes())
        f.write(np.ascontiguousarray(D2, dtype=np.float64).reshape(-1)[: norb ** 4].tobytes())
