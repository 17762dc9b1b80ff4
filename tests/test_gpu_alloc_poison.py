"""Reads of device memory that no kernel has written. The allocator hands
released blocks out again, so a buffer can hold whatever its last owner left
there, NaN and Inf bit patterns included. MHIP_ALLOC_POISON=1 fills every
block it hands out with 0xff bytes (NaN doubles, -1 indices), which turns
such a read into a wrong result on every run. Under it, the batched Davidson
on the sliced copy must converge to the same energies (its SpMV forms
AV = H v + 0 * AV, gespmbv's form (spmbv.hpp:71), which reads AV), and a
sliced product into a zeroed output must equal the CSR kernel's."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from macis_amd import _capi
from macis_amd import dets as D
from macis_amd.synthetic import seeded_integrals
from oracle import oracle as O

from conftest import announce

pytestmark = pytest.mark.gpu
EPS = np.finfo(float).eps


def test_batch_davidson_under_poisoned_allocations(ctx, monkeypatch):
    announce("batched Davidson, sliced SpMV, with every allocation poisoned")
    sizes = [6, 8, 6]
    norbs, dl, Ts, Vs = [], [], [], []
    for i, n in enumerate(sizes):
        T, V = seeded_integrals(n, 300 + i)
        norbs.append(n)
        dl.append(D.generate_hilbert_space(1, n, n // 2, n // 2)[:, 0])
        Ts.append(T)
        Vs.append(V)
    ref = []
    for n, d, T, V in zip(norbs, dl, Ts, Vs):
        H = O.Ham(T, V).csr(1, d.reshape(-1, 1), thresh=EPS)
        A = sp.csr_matrix((H[2], H[1], H[0]), shape=(len(d), len(d)))
        ref.append(sla.eigsh(A, k=1, which="SA")[0][0])
    monkeypatch.setenv("MHIP_ALLOC_POISON", "1")
    B = _capi.Batch(ctx, norbs, dl, Ts, Vs)
    B.build(EPS)
    assert B.spmv_layout(1) == 1
    E, res, _ = B.davidson(max_m=20, max_restarts=200, tol=1e-8)
    assert np.all(np.isfinite(E))
    assert np.max(np.abs(np.asarray(E) - np.asarray(ref))) < 1e-8


def test_sliced_spmv_under_poisoned_allocations(ctx, monkeypatch):
    announce("sliced SpMV with every allocation poisoned")
    monkeypatch.setenv("MHIP_ALLOC_POISON", "1")
    rng = np.random.default_rng(8)
    m, L = 1000, 10
    rp = np.arange(0, m * L + 1, L, dtype=np.int64)
    ci = np.concatenate([np.sort(rng.choice(m, size=L, replace=False)) for _ in range(m)])
    nz = rng.standard_normal(m * L)
    A = ctx.csr_from_host(m, m, rp, ci.astype(np.int32), nz)
    x = ctx.to_device(rng.standard_normal(m))
    y0 = ctx.empty(m, np.float64).zero()
    ctx.spmv(A, 1, 1.0, x, m, 0.0, y0, m)  # CSR kernel
    assert ctx.prepare_spmv(A, always=True)
    y1 = ctx.empty(m, np.float64).zero()
    ctx.spmv(A, 1, 1.0, x, m, 0.0, y1, m)
    a, b = y0.numpy(), y1.numpy()
    assert np.all(np.isfinite(b))
    assert np.max(np.abs(a - b)) <= 1e-13 * np.abs(a).max()
