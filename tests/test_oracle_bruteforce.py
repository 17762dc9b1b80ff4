"""Pin the oracle where the reference fixtures leave gaps: brute-force
second quantisation, closed forms and dense eigensolves on small spaces."""
import itertools
import math

import numpy as np
import pytest

from macis_amd import dets as D
from macis_amd.synthetic import seeded_integrals
from oracle import oracle as O


# ---- literals from the reference tests --------------------------------
def test_generate_combs_literal():
    # tests/sd_operations.cxx:217-222
    assert D.generate_combs(4, 2).tolist() == [0x3, 0x5, 0x9, 0x6, 0xA, 0xC]
    o = O.hilbert_space(1, 4, 2, 0)[:, 0].tolist()
    assert o == [0x3, 0x5, 0x9, 0x6, 0xA, 0xC]


def test_generate_hilbert_space_literal():
    # tests/sd_operations.cxx:224-237
    ref = [0x300000003, 0x500000003, 0x900000003, 0x600000003, 0xA00000003,
           0xC00000003, 0x300000005, 0x500000005, 0x900000005, 0x600000005,
           0xA00000005, 0xC00000005, 0x300000009, 0x500000009, 0x900000009,
           0x600000009, 0xA00000009, 0xC00000009, 0x300000006, 0x500000006,
           0x900000006, 0x600000006, 0xA00000006, 0xC00000006, 0x30000000A,
           0x50000000A, 0x90000000A, 0x60000000A, 0xA0000000A, 0xC0000000A,
           0x30000000C, 0x50000000C, 0x90000000C, 0x60000000C, 0xA0000000C,
           0xC0000000C]
    assert O.hilbert_space(1, 4, 2, 2)[:, 0].tolist() == ref
    assert D.generate_hilbert_space(1, 4, 2, 2)[:, 0].tolist() == ref
    w2 = D.generate_hilbert_space(2, 4, 2, 2)
    assert np.array_equal(w2[:, 0] | (w2[:, 1] << np.uint64(32)), np.array(ref, np.uint64))
    assert np.array_equal(O.hilbert_space(2, 4, 2, 2), w2)


def test_cisd_space_matches_sets():
    # tests/sd_operations.cxx:201-214 (CISD == HF + singles + doubles)
    for W in (1, 2):
        hf = D.canonical_hf_determinant(W, 4, 4)
        cisd = O.cisd_space(W, 8, hf)
        a, b = D.split_spins(cisd, W)
        ha, hb = D.split_spins(hf, W)
        exc = np.array([bin(int(x)).count("1") for x in (a ^ ha[0])]) // 2 + \
            np.array([bin(int(x)).count("1") for x in (b ^ hb[0])]) // 2
        assert exc.max() == 2
        # every det with excitation level <= 2 appears exactly once
        fci = D.generate_hilbert_space(W, 8, 4, 4)
        fa, fb = D.split_spins(fci, W)
        fexc = np.array([bin(int(x)).count("1") for x in (fa ^ ha[0])]) // 2 + \
            np.array([bin(int(x)).count("1") for x in (fb ^ hb[0])]) // 2
        assert len(cisd) == int((fexc <= 2).sum())
        assert len({tuple(r) for r in cisd.tolist()}) == len(cisd)
        assert tuple(cisd[0]) == tuple(hf[0])


# ---- second quantisation ------------------------------------------------
def _apply(op, p, det):
    """a_p (op=0) or a+_p (op=1) on a spin-orbital bitmask; (sign, det)."""
    occ = (det >> p) & 1
    if occ == op:
        return 0, det
    sign = -1 if bin(det & ((1 << p) - 1)).count("1") % 2 else 1
    return sign, det ^ (1 << p)


def _sq_hamiltonian(n, T, V, dets):
    """<bra|H|ket> from H = sum h_pq a+_p a_q + 1/2 sum (pq|rs) a+_p a+_r a_s a_q
    over spin orbitals (alpha p -> bit p, beta p -> bit 32+p)."""
    Tm = T.reshape(n, n, order="F")
    Vm = V.reshape(n, n, n, n, order="F")
    index = {d: i for i, d in enumerate(dets)}
    H = np.zeros((len(dets), len(dets)))
    so = [(p, s) for s in (0, 32) for p in range(n)]
    for j, ket in enumerate(dets):
        for (p, sp), (q, sq) in itertools.product(so, so):
            if sp != sq:
                continue
            s1, d1 = _apply(0, q + sq, ket)
            if not s1:
                continue
            s2, d2 = _apply(1, p + sp, d1)
            if s2 and d2 in index:
                H[index[d2], j] += s1 * s2 * Tm[p, q]
        for (p, sp), (q, sq), (r, sr), (s, ss) in itertools.product(so, so, so, so):
            if sp != sq or sr != ss:
                continue
            c1, d1 = _apply(0, q + sq, ket)
            if not c1:
                continue
            c2, d2 = _apply(0, s + ss, d1)
            if not c2:
                continue
            c3, d3 = _apply(1, r + sr, d2)
            if not c3:
                continue
            c4, d4 = _apply(1, p + sp, d3)
            if c4 and d4 in index:
                H[index[d4], j] += 0.5 * c1 * c2 * c3 * c4 * Vm[p, q, r, s]
    return H


@pytest.mark.parametrize("n,na,nb,seed", [(4, 2, 2, 7), (4, 3, 1, 11), (5, 2, 3, 5)])
def test_matrix_elements_vs_second_quantisation(n, na, nb, seed):
    T, V = seeded_integrals(n, seed)
    dets = D.generate_hilbert_space(1, n, na, nb)[:, 0].tolist()
    Hsq = _sq_hamiltonian(n, T, V, dets)
    Ho = O.Ham(T, V)
    Hor = np.array([[Ho.matrix_element(1, np.array([b], np.uint64), np.array([k], np.uint64))
                     for k in dets] for b in dets])
    assert np.allclose(Hor, Hsq, rtol=0, atol=1e-12)
    # the CSR build carries the same elements
    rp, ci, nz = Ho.csr(1, np.array(dets, np.uint64).reshape(-1, 1), thresh=0.0)
    dense = np.zeros_like(Hor)
    for i in range(len(dets)):
        dense[i, ci[rp[i]:rp[i + 1]]] = nz[rp[i]:rp[i + 1]]
    nzmask = np.abs(Hor) > 0
    assert np.array_equal(dense[nzmask], Hor[nzmask])


def test_sign_primitives_vs_second_quantisation():
    # single_excitation_sign / doubles_sign_indices (sd_operations.hpp:497-551)
    rng = np.random.default_rng(3)
    nb = 10
    for _ in range(300):
        occ = rng.choice(nb, 5, replace=False)
        ket = sum(1 << int(x) for x in occ)
        o1, o2 = rng.choice(occ, 2, replace=False)
        virt = [x for x in range(nb) if not (ket >> x) & 1]
        v1, v2 = rng.choice(virt, 2, replace=False)
        # single: <bra| a+_v a_o |ket>
        bra = ket ^ (1 << int(o1)) ^ (1 << int(v1))
        c1, d = _apply(0, int(o1), ket)
        c2, d = _apply(1, int(v1), d)
        assert d == bra
        assert O.single_excitation_sign(ket, int(v1), int(o1)) == c1 * c2
        # double via two sequential singles on the updated ket
        bra2 = ket ^ (1 << int(o1)) ^ (1 << int(v1)) ^ (1 << int(o2)) ^ (1 << int(v2))
        ex = bra2 ^ ket
        O1, V1, O2, V2, sgn = O.doubles_sign_indices(bra2, ket, ex)
        lo = sorted([int(o1), int(o2)])
        lv = sorted([int(v1), int(v2)])
        assert (O1, O2) == (lo[0], lo[1]) and (V1, V2) == (lv[0], lv[1])
        c1, d = _apply(0, O1, ket)
        c2, d = _apply(1, V1, d)
        c3, d = _apply(0, O2, d)
        c4, d = _apply(1, V2, d)
        assert d == bra2 and sgn == c1 * c2 * c3 * c4


def test_syev_lower_vs_numpy():
    rng = np.random.default_rng(0)
    for n in (1, 2, 5, 20, 21):
        A = rng.standard_normal((n, n))
        A = A + A.T
        W, Vec = O.syev_lower(np.tril(A))
        assert np.allclose(W, np.linalg.eigvalsh(A), atol=1e-12)
        assert np.allclose(A @ Vec, Vec * W, atol=1e-11)


@pytest.mark.parametrize("n,ne", [(6, 3), (8, 4)])
def test_fci_davidson_vs_dense(n, ne):
    T, V = seeded_integrals(n, 100 + n)
    Ho = O.Ham(T, V)
    dets = D.generate_hilbert_space(1, n, ne, ne)
    rp, ci, nz = Ho.csr(1, dets, thresh=np.finfo(float).eps)
    N = len(dets)
    # FCI row count closed form: 1 + 2 no nv + 2 C(no,2) C(nv,2) + (no nv)^2
    no, nv = ne, n - ne
    per_row = 1 + 2 * no * nv + 2 * math.comb(no, 2) * math.comb(nv, 2) + (no * nv) ** 2
    assert np.all(np.diff(rp) == per_row)
    dense = np.zeros((N, N))
    for i in range(N):
        dense[i, ci[rp[i]:rp[i + 1]]] = nz[rp[i]:rp[i + 1]]
    assert np.array_equal(dense, dense.T)
    E_dense = np.linalg.eigvalsh(dense)[0]
    # the synthetic integrals are strongly correlated: the reference
    # (no-restart) Davidson with max_m 20 stops before 1e-8, the restarted
    # variant converges
    with pytest.raises(RuntimeError, match="Did Not Converge"):
        Ho.selected_ci_diag(1, dets)
    E, X, nnz = Ho.selected_ci_diag(1, dets, max_restarts=50)
    assert E == pytest.approx(E_dense, abs=1e-10)


@pytest.mark.parametrize("max_m", [2, 3, 4, 6])
def test_restarted_davidson_small_subspace_reaches_ground_state(max_m):
    """The restarted variant with tiny subspaces (one-column restart at
    max_m 2, previous-Ritz-vector restart with the second-Ritz-vector
    fallback at 3, regular cycles at 4 and 6) converges to the dense ground
    state, and its vector is an eigenvector (residual from the definition)."""
    n = 6
    T, V = seeded_integrals(n, 3)
    dets = D.generate_hilbert_space(1, n, 3, 3)
    rp, ci, nz = O.Ham(T, V).csr(1, dets, thresh=1e-16)
    N = len(dets)
    dense = np.zeros((N, N))
    for i in range(N):
        dense[i, ci[rp[i]:rp[i + 1]]] = nz[rp[i]:rp[i + 1]]
    w = np.linalg.eigvalsh(dense)
    Dg = np.diag(dense).copy()
    X0 = np.zeros(N)
    X0[np.argmin(Dg)] = 1.0
    it, E, X = O.davidson(N, max_m, rp, ci, nz, Dg, 1e-8, X0, max_restarts=20000)
    assert E == pytest.approx(w[0], abs=1e-9)
    assert np.linalg.norm(X) == pytest.approx(1.0, abs=1e-10)
    assert np.linalg.norm(dense @ X - E * X) < 1e-8


def test_oracle_parallel_and_one_pass_rows_equal_serial():
    """The OpenMP row-parallel double loop and the one-pass variant give the
    serial loop's CSR rows exactly; the threaded gespmbv gives the row dots
    of a plain loop."""
    n = 8
    T, V = seeded_integrals(n, 5)
    Ho = O.Ham(T, V)
    dets = D.generate_hilbert_space(1, n, 4, 4)
    c0, col0, val0 = Ho.csr_rows(1, dets, 100, 700, thresh=1e-16)
    c1, col1, val1 = Ho.csr_rows_parallel(1, dets, 100, 700, thresh=1e-16)
    assert np.array_equal(c0, c1) and np.array_equal(col0, col1) and np.array_equal(val0, val1)
    col2 = np.zeros(len(col0) + 5, dtype=np.int32)
    val2 = np.zeros(len(col0) + 5)
    c2 = Ho.csr_rows_into(1, dets, 100, 700, col2, val2, thresh=1e-16)
    assert np.array_equal(c0, c2) and np.array_equal(col0, col2[:len(col0)])
    assert np.array_equal(val0, val2[:len(col0)])
    c3 = Ho.csr_rows_parallel(1, dets, 100, 700, thresh=1e-16, col=col2, val=val2)
    assert np.array_equal(c0, c3) and np.array_equal(val0, val2[:len(col0)])
    assert O.num_threads() >= 1
    rp, ci, nz = Ho.csr(1, dets[:600], thresh=1e-16)
    x = np.random.default_rng(1).standard_normal(600)
    y = np.zeros(600)
    O.gespmbv(1, 1.0, 600, rp, ci, nz, 0, x, 600, 0.0, y, 600)
    ref = np.array([sum(nz[j] * x[ci[j]] for j in range(rp[i], rp[i + 1])) for i in range(600)])
    assert np.array_equal(y, ref)


def test_synthetic_integrals_symmetry():
    n = 6
    T, V = seeded_integrals(n, 1)
    Vm = V.reshape(n, n, n, n, order="F")
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.array_equal(Vm, Vm.transpose(perm))
    Tm = T.reshape(n, n, order="F")
    assert np.array_equal(Tm, Tm.T)
    ev = np.linalg.eigvalsh(V.reshape(n * n, n * n))
    assert ev.min() > -1e-10
    T2, V2 = seeded_integrals(n, 1)
    assert np.array_equal(T, T2) and np.array_equal(V, V2)


def test_asci_search_scores_by_brute_force():
    """Rescore every candidate of a small search from its definition:
    rv(D) = sum_i C_i <D|H|D_i> / (E0 - H_DD) with the exact diagonal; the
    fast diagonals equal the exact ones, so accumulated rv must match."""
    n, ne = 6, 3
    T, V = seeded_integrals(n, 42)
    Ho = O.Ham(T, V)
    dets = D.generate_hilbert_space(1, n, ne, ne)
    E, X, _ = Ho.selected_ci_diag(1, dets, max_restarts=50)
    order = np.argsort(-np.abs(X), kind="stable")
    core = dets[order[:5]]
    Cc = X[order[:5]]
    s = O.ASCISettings.defaults(h_el_tol=0.0)
    st, rv = Ho.asci_contributions(1, s, core, E, Cc)
    acc = {}
    for k, r in zip(st[:, 0].tolist(), rv.tolist()):
        acc[k] = acc.get(k, 0.0) + r
    for k, r in list(acc.items())[::7]:
        ka = np.array([k], np.uint64)
        hdd = Ho.matrix_element(1, ka, ka)
        ref = sum(Cc[i] * Ho.matrix_element(1, ka, core[i]) for i in range(5)) / (E - hdd)
        # the excitation generator may also reach other core dets; exclude
        # candidates that are themselves core dets
        if k in set(core[:, 0].tolist()):
            continue
        assert r == pytest.approx(ref, rel=1e-10, abs=1e-14)
