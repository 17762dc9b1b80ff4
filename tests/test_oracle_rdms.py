"""Pin the oracle's form_rdms (double_loop.hpp:116-165, util/rdms.hpp:15-132):
the reference's own water CAS(8e,8o) RDM fixture, brute-force second
quantisation on small spaces, the HF identities and the trace of the
reference's RDMS test (tests/double_loop.cxx:180-223), the energy identity
tr(D1 T) + tr(D2 V) = C^T H C, and the 1-RDM-only branch order."""
import itertools
import os

import numpy as np
import pytest

from macis_amd import dets as D
from macis_amd.fcidump import read_fcidump
from macis_amd.synthetic import seeded_integrals
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def read_rdms_binary(fname):
    """fcidump.cxx:209-251: int32 norb, then the 1-RDM (norb^2) and the 2-RDM
    (norb^4), column-major doubles."""
    b = open(fname, "rb").read()
    n = int(np.frombuffer(b[:4], dtype=np.int32)[0])
    f = np.frombuffer(b[4:], dtype=np.float64)
    assert f.size == n * n + n ** 4
    return n, f[: n * n].reshape((n, n), order="F"), f[n * n:].reshape((n,) * 4, order="F")


def cas_active_integrals(T, V, norb, ninact, nact):
    """Active-space one-body operator with the inactive (doubly occupied)
    orbitals folded in: T_pq + sum_i [2 (pq|ii) - (pi|iq)]; V restricted."""
    Tm = T.reshape(norb, norb, order="F")
    Vm = V.reshape((norb,) * 4, order="F")
    act = np.arange(ninact, ninact + nact)
    Ta = Tm[np.ix_(act, act)].copy()
    for i in range(ninact):
        Ta += 2.0 * Vm[np.ix_(act, act, [i], [i])][:, :, 0, 0] - Vm[np.ix_(act, [i], [i], act)][:, 0, 0, :]
    Va = np.ascontiguousarray(Vm[np.ix_(act, act, act, act)])
    return Ta.reshape(-1, order="F"), Va.reshape(-1, order="F")


def water_cas88():
    norb, _, T, V = read_fcidump(os.path.join(GOLDEN, "h2o.ccpvdz.fci.dat"))
    return cas_active_integrals(T, V, norb, 1, 8)


def test_rdms_water_cas88_fixture():
    # h2o.ccpvdz.cas88.rdms.bin: the reference's RDMs of the CAS(8e,8o)
    # ground state (1 inactive orbital), read by tests/mcscf.cxx:69
    n, F1, F2 = read_rdms_binary(os.path.join(GOLDEN, "h2o.ccpvdz.cas88.rdms.bin"))
    assert n == 8
    Ta, Va = water_cas88()
    dets = D.generate_hilbert_space(1, 8, 4, 4)
    E, X = O.Ham(Ta, Va).selected_ci_diag(1, dets)[:2]
    d1, d2 = O.form_rdms(8, 1, dets, X)
    assert np.abs(d1 - F1).max() < 1e-10
    assert np.abs(d2 - F2).max() < 1e-10
    # the CAS energy from the RDMs
    assert float(np.dot(d1.reshape(-1, order="F"), Ta) + np.dot(d2.reshape(-1, order="F"), Va)) == \
        pytest.approx(E, abs=1e-9)


# ---- second quantisation ------------------------------------------------
def _apply(op, p, det):
    occ = (det >> p) & 1
    if occ == op:
        return 0, det
    sign = -1 if bin(det & ((1 << p) - 1)).count("1") % 2 else 1
    return sign, det ^ (1 << p)


def _sq_rdms(n, dets, C):
    """D1(p,q) = sum_s <a+_ps a_qs>, D2(p,q,r,s) = 1/2 sum_st <a+_ps a+_rt a_st a_qs>
    with alpha p -> bit p, beta p -> bit 32 + p (wfn_t<64> layout)."""
    index = {d: i for i, d in enumerate(dets)}
    D1 = np.zeros((n, n))
    D2 = np.zeros((n,) * 4)
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
                D1[p, q] += s1 * s2 * C[index[d2]] * C[j]
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
                D2[p, q, r, s] += 0.5 * c1 * c2 * c3 * c4 * C[index[d4]] * C[j]
    return D1, D2


def _sym8(D2):
    """Projection onto the 8-fold symmetry of real integrals (pq|rs)."""
    t = (D2 + D2.transpose(1, 0, 2, 3) + D2.transpose(0, 1, 3, 2) + D2.transpose(1, 0, 3, 2))
    return (t + t.transpose(2, 3, 0, 1)) / 8.0


@pytest.mark.parametrize("n,na,nb,seed", [(4, 2, 2, 3), (4, 3, 1, 4), (5, 2, 3, 9)])
def test_rdms_vs_second_quantisation(n, na, nb, seed):
    # D1 entry by entry. D2: the reference books the same-spin exchange of the
    # diagonal branch at (p,q,p,q) (util/rdms.hpp:87) where the operator
    # definition has (p,q,q,p); both meet the same integral under the 8-fold
    # symmetry, so D2 is compared on that subspace (the water fixture above
    # pins the reference's entry-wise placement).
    dets = D.generate_hilbert_space(1, n, na, nb)
    C = np.random.default_rng(seed).standard_normal(len(dets))
    d1, d2 = O.form_rdms(n, 1, dets, C)
    S1, S2 = _sq_rdms(n, dets[:, 0].tolist(), C)
    assert np.abs(d1 - S1).max() < 1e-12
    assert np.abs(_sym8(d2) - _sym8(S2)).max() < 1e-12
    # every other entry agrees exactly as stored
    p, q, r, s = np.indices(d2.shape)
    exch = (p == r) & (q == s) | (p == s) & (q == r)
    assert np.abs((d2 - S2)[~exch]).max() < 1e-12


def test_rdms_hf_identities():
    # tests/double_loop.cxx:199-223 (norb 34, nocc 5, wfn_t<128>)
    n, nocc = 34, 5
    hf = D.canonical_hf_determinant(2, nocc, nocc)
    d1, d2 = O.form_rdms(n, 2, hf, np.array([1.0]))
    t = d2.copy()
    for i, j, k, l in itertools.product(range(nocc), repeat=4):
        t[i, j, l, k] -= 0.5 * d1[i, j] * d1[k, l]
        t[i, j, l, k] += 0.25 * d1[i, l] * d1[k, j]
    assert np.abs(t).sum() < 1e-15
    e = d1.copy()
    for i in range(nocc):
        e[i, i] -= 2.0
    assert np.abs(e).sum() < 1e-15


@pytest.mark.parametrize("W,n,na,nb", [(1, 6, 3, 3), (2, 7, 3, 2)])
def test_rdms_energy_identity(W, n, na, nb):
    # tr(D1 T) + tr(D2 V) = C^T H C for any C (H from the pinned CSR build,
    # every connected pair kept)
    T, V = seeded_integrals(n, 40 + n)
    dets = D.generate_hilbert_space(W, n, na, nb)
    C = np.random.default_rng(n).standard_normal(len(dets))
    d1, d2 = O.form_rdms(n, W, dets, C)
    rp, ci, nz = O.Ham(T, V).csr(W, dets, thresh=-1.0)
    HC = np.array([np.dot(nz[rp[i]:rp[i + 1]], C[ci[rp[i]:rp[i + 1]]]) for i in range(len(dets))])
    E = float(np.dot(d1.reshape(-1, order="F"), T) + np.dot(d2.reshape(-1, order="F"), V))
    assert E == pytest.approx(float(C @ HC), rel=1e-12)
    # traces: N electrons, N(N-1)/2 pairs
    N = na + nb
    c2 = float(C @ C)
    assert np.trace(d1) == pytest.approx(N * c2, rel=1e-12)
    assert np.einsum("ppqq->", d2) == pytest.approx(0.5 * N * (N - 1) * c2, rel=1e-12)


def test_rdms_ordm_only_branch_order():
    # Without a 2-RDM the dispatch of util/rdms.hpp:103-130 sends a same-spin
    # double (4 flipped bits in one spin) to the diagonal branch: with two
    # determinants that differ by an alpha double and C = (1, 1), every one of
    # the four ordered pairs adds the bra's occupations to the diagonal.
    d0 = D.canonical_hf_determinant(1, 2, 2)
    d1 = np.array([[int(d0[0, 0]) ^ 0b1111]], dtype=np.uint64)  # alpha 0,1 -> 2,3
    dets = np.concatenate([d0, d1])
    o, t = O.form_rdms(6, 1, dets, np.array([1.0, 1.0]), want_trdm=False)
    assert t is None
    occ = lambda d: [((int(d) >> p) & 1) + ((int(d) >> (32 + p)) & 1) for p in range(6)]
    assert np.array_equal(np.diag(o), 2.0 * (np.array(occ(d0[0, 0])) + np.array(occ(d1[0, 0]))))
    assert np.count_nonzero(o - np.diag(np.diag(o))) == 0
    # with the 2-RDM the same pairs go to rdm_contributions_4 instead
    o2, _ = O.form_rdms(6, 1, dets, np.array([1.0, 1.0]))
    assert np.array_equal(np.diag(o2), np.array(occ(d0[0, 0])) + np.array(occ(d1[0, 0])))


# ---- natural-orbital rotation of asci_grow (grow.hpp:73-114) ----------
def test_index_transforms_vs_einsum():
    # transform.cxx:22-115 against the defining sums
    n = 5
    rng = np.random.default_rng(12)
    U = np.linalg.qr(rng.standard_normal((n, n)))[0]
    X2 = rng.standard_normal(n * n)
    X4 = rng.standard_normal(n ** 4)
    ref2 = np.einsum("ip,ij,jq->pq", U, X2.reshape(n, n, order="F"), U)
    ref4 = np.einsum("ijkl,ip,jq,kr,ls->pqrs", X4.reshape((n,) * 4, order="F"), U, U, U, U)
    assert np.abs(O.two_index_transform(X2, U) - ref2.reshape(-1, order="F")).max() < 1e-13
    assert np.abs(O.four_index_transform(X4, U) - ref4.reshape(-1, order="F")).max() < 1e-12


def test_two_index_transform_literals():
    # tests/transform.cxx: identity and permutation of a 4 x 4 matrix
    A = np.array([1, 5, 9, 13, 2, 6, 10, 14, 3, 7, 11, 15, 4, 8, 12, 16], dtype=float)
    P = np.array([1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 1, 0, 0], dtype=float).reshape(4, 4, order="F")
    assert np.array_equal(O.two_index_transform(A, np.eye(4)), A)
    assert np.array_equal(O.two_index_transform(A, P),
                          np.array([1, 9, 13, 5, 3, 11, 15, 7, 4, 12, 16, 8, 2, 10, 14, 6], dtype=float))


def test_natural_orbital_rotation_fci_invariants():
    # In a complete space the rotation changes no energy, and the ground
    # state's 1-RDM in the rotated orbitals is diagonal with the occupations
    # in descending order.
    n = 6
    T, V = seeded_integrals(n, 61)
    dets = D.generate_hilbert_space(1, n, 3, 3)
    H = O.Ham(T, V)
    E, X = H.selected_ci_diag(1, dets, max_m=60)[:2]
    occ = H.natural_orbital_rotation(1, dets, X)
    assert np.all(np.diff(occ) <= 1e-12) and occ.sum() == pytest.approx(6.0, abs=1e-10)
    Tr, Vr = H.array(0), H.array(1)
    Hr = O.Ham(Tr, Vr)
    Er, Xr = Hr.selected_ci_diag(1, dets, max_m=60)[:2]
    assert Er == pytest.approx(E, abs=1e-9)
    # (with the 2-RDM: the 1-RDM-only call has the reference's branch quirk)
    d1, _ = O.form_rdms(n, 1, dets, Xr)
    assert np.abs(d1 - np.diag(occ)).max() < 1e-7
    # T was rotated as U^T T U with U the eigenvectors of -D1
    d0, _ = O.form_rdms(n, 1, dets, X)
    w, U = np.linalg.eigh(-d0)
    Tm = T.reshape(n, n, order="F")
    Tref = U.T @ Tm @ U
    # columns are fixed up to sign: compare |entries| and the diagonal
    assert np.abs(np.abs(Tr.reshape(n, n, order="F")) - np.abs(Tref)).max() < 1e-8
    assert np.abs(np.diag(Tr.reshape(n, n, order="F")) - np.diag(Tref)).max() < 1e-8


def test_grow_with_rotation_lowers_the_energy():
    # H2O: growing to 2,000 determinants with rotations from 100 on ends
    # below the unrotated grow (selected CI in natural orbitals converges
    # faster); both keep the requested size.
    norb, _, T, V = read_fcidump(os.path.join(GOLDEN, "h2o.ccpvdz.fci.dat"))
    s = O.ASCISettings.defaults()
    s.ntdets_max = 2000
    hf = D.canonical_hf_determinant(1, 5, 5)
    H0 = O.Ham(T, V)
    E0 = H0.matrix_element(1, hf, hf)
    Ep, wp, _ = H0.asci_grow(1, s, E0, hf, np.array([1.0]))
    Hr = O.Ham(T, V)
    Er, wr, Xr = Hr.asci_grow_rot(1, s, E0, hf, np.array([1.0]), rot_size_start=100)
    assert len(wp) == len(wr) == 2000
    assert Er < Ep - 1e-3
    assert np.linalg.norm(Xr) == pytest.approx(1.0, abs=1e-10)
