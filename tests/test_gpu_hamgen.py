"""GPU parity of the Hamiltonian build (K1) against the CPU oracle.

The oracle restates DoubleLoopHamiltonianGenerator::make_csr_hamiltonian_block_
(double_loop.hpp:30-98); the reference test csr_hamiltonian.cxx:18-71 pins
the H2O CISD row pointers. Pattern (rowptr, colind) must be identical and
values bit-identical (the device evaluates the same sums in the same order
without FMA contraction; the required bound is 1e-12 relative).
"""
import os

import numpy as np
import pytest

from macis_amd import dets as D
from macis_amd.synthetic import seeded_integrals, selected_ci_dets
from oracle import oracle as O

from conftest import GOLDEN, announce

pytestmark = pytest.mark.gpu
EPS = np.finfo(float).eps


def assert_same_csr(gpu, ref):
    rp, ci, nz = gpu
    rpo, cio, nzo = ref
    assert np.array_equal(rp, rpo), "rowptr differs"
    assert np.array_equal(ci, cio), "colind differs"
    rel = np.abs(nz - nzo) / np.maximum(np.abs(nzo), 1e-300)
    assert rel.max(initial=0.0) <= 1e-12
    assert np.array_equal(nz, nzo), f"{np.count_nonzero(nz != nzo)} values not bit-identical"


def test_h2o_cisd_matches_fixture_and_oracle(ctx, h2o, h2o_oracle, h2o_cisd):
    announce("hamgen H2O CISD 12636 dets, thresh 1e-16")
    norb, _, T, V = h2o
    ctx.set_integrals(norb, T, V)
    d = ctx.to_device(h2o_cisd)
    H = ctx.ham_build(1, d, len(h2o_cisd), thresh=1e-16)
    ref_rp = np.fromfile(os.path.join(GOLDEN, "h2o.ccpvdz.cisd.rowptr.bin"), dtype=np.int32)
    rp32, ci32, nz32 = H.to_host_i32()
    assert np.array_equal(rp32, ref_rp)
    assert H.nnz == 3517816
    assert_same_csr(H.to_host(), h2o_oracle.csr(1, h2o_cisd, thresh=1e-16))
    t = ctx.timings()
    assert t["ham_nnz"] == 3517816 and t["ham_nnz_struct"] >= t["ham_nnz"]
    assert t["ham_compact_ms"] > 0  # symmetry zeros are dropped by compaction


@pytest.mark.parametrize("W", [1, 2])
@pytest.mark.parametrize("n,na,nb,seed", [(6, 3, 3, 1), (8, 4, 4, 2), (8, 5, 3, 3), (10, 3, 2, 4)])
def test_fci_synthetic(ctx, W, n, na, nb, seed):
    announce(f"hamgen FCI n={n} na={na} nb={nb} W={W}")
    T, V = seeded_integrals(n, seed)
    ctx.set_integrals(n, T, V)
    dets = D.generate_hilbert_space(W, n, na, nb)
    H = ctx.ham_build(W, ctx.to_device(dets), len(dets), thresh=EPS)
    assert_same_csr(H.to_host(), O.Ham(T, V).csr(W, dets, thresh=EPS))


@pytest.mark.parametrize("thresh", [0.0, 1e-3, 0.05, 1e3])
def test_threshold_compaction(ctx, thresh):
    announce(f"hamgen threshold {thresh}")
    n = 8
    T, V = seeded_integrals(n, 9)
    ctx.set_integrals(n, T, V)
    dets = D.generate_hilbert_space(1, n, 4, 3)
    H = ctx.ham_build(1, ctx.to_device(dets), len(dets), thresh=thresh)
    assert_same_csr(H.to_host(), O.Ham(T, V).csr(1, dets, thresh=thresh))


def test_rectangular_blocks_zero_dets_and_empty(ctx):
    announce("hamgen rectangular / zero dets / empty ranges")
    n = 8
    T, V = seeded_integrals(n, 5)
    ctx.set_integrals(n, T, V)
    orc = O.Ham(T, V)
    dets = D.generate_hilbert_space(1, n, 4, 4)
    rng = np.random.default_rng(0)
    bra = dets[rng.permutation(len(dets))[:777]].copy()
    ket = dets[::-1].copy()
    bra[5] = 0   # zero bra -> empty row (double_loop.hpp:50)
    ket[17] = 0  # zero ket -> never a column (double_loop.hpp:62)
    H = ctx.ham_build(1, ctx.to_device(bra), len(bra), ctx.to_device(ket), len(ket), thresh=EPS)
    assert (H.m, H.n) == (len(bra), len(ket))
    assert_same_csr(H.to_host(), orc.csr(1, bra, ket, thresh=EPS))
    # empty ranges give an empty matrix (csr_hamiltonian.hpp:30-35)
    d = ctx.to_device(bra)
    E = ctx.ham_build(1, d, 0, d, len(bra), thresh=EPS)
    assert (E.m, E.n, E.nnz) == (0, len(bra), 0)
    E = ctx.ham_build(1, d, 10, d, 0, thresh=EPS)
    assert (E.m, E.n, E.nnz) == (10, 0, 0)
    assert np.array_equal(E.to_host()[0], np.zeros(11, np.int64))


def test_selected_ci_128bit(ctx):
    announce("hamgen 64 orbitals 20a20b, 6000 sorted dets, W=2")
    n = 64
    T, V = seeded_integrals(n, 77)
    ctx.set_integrals(n, T, V)
    dets = selected_ci_dets(n, 20, 20, 6000, seed=3, T=T)
    H = ctx.ham_build(2, ctx.to_device(dets), len(dets), thresh=EPS)
    assert_same_csr(H.to_host(), O.Ham(T, V).csr(2, dets, thresh=EPS))


def test_diagonal_entry_point(ctx):
    announce("ham_diagonal")
    n = 10
    T, V = seeded_integrals(n, 8)
    ctx.set_integrals(n, T, V)
    orc = O.Ham(T, V)
    dets = D.generate_hilbert_space(1, n, 5, 4)[:500]
    out = ctx.empty(len(dets), np.float64)
    ctx.ham_diagonal(1, ctx.to_device(dets), len(dets), out)
    ref = np.array([orc.matrix_element(1, d, d) for d in dets])
    assert np.array_equal(out.numpy(), ref)


def test_cas10_sampled_rows(ctx):
    announce("hamgen CAS(10e,10o) 63504 dets, sampled rows vs oracle")
    n = 10
    T, V = seeded_integrals(n, 10)
    ctx.set_integrals(n, T, V)
    dets = D.generate_hilbert_space(1, n, 5, 5)
    H = ctx.ham_build(1, ctx.to_device(dets), len(dets), thresh=EPS)
    assert H.nnz == 55629504
    rp, ci, nz = H.to_host()
    assert np.all(np.diff(rp) == 876)
    orc = O.Ham(T, V)
    for r0 in (0, 12345, 31752, 63504 - 40):
        cnt, col, val = orc.csr_rows(1, dets, r0, r0 + 40, thresh=EPS)
        assert np.array_equal(np.diff(rp[r0:r0 + 41]), cnt)
        assert np.array_equal(ci[rp[r0]:rp[r0 + 40]], col)
        assert np.array_equal(nz[rp[r0]:rp[r0 + 40]], val)
    # pattern symmetry and ascending columns
    assert np.all(np.diff(ci.reshape(-1, 876), axis=1) > 0)
    rows = np.repeat(np.arange(len(dets), dtype=np.int64), 876)
    key = rows * len(dets) + ci
    keyT = ci.astype(np.int64) * len(dets) + rows
    assert np.array_equal(np.sort(key), np.sort(keyT))


def test_cas12_full_size_properties(ctx):
    announce("hamgen CAS(12e,12o) 853776 dets full size")
    import ctypes

    from macis_amd import _capi

    n = 12
    T, V = seeded_integrals(n, 12)
    ctx.set_integrals(n, T, V)
    dets = D.generate_hilbert_space(1, n, 6, 6)
    N = len(dets)
    H = ctx.ham_build(1, ctx.to_device(dets), N, thresh=EPS)
    assert H.nnz == 1553018544  # 1819 per row: no element at or below eps
    d_rp, d_ci, d_nz = H.device_arrays()
    L = _capi.lib()
    rp = np.empty(N + 1, np.int64)
    L.mhip_memcpy_d2h(ctx.h, rp.ctypes.data, d_rp, rp.nbytes)
    assert np.all(np.diff(rp) == 1819)
    orc = O.Ham(T, V)
    for r0 in (0, 400000, N - 16):
        cnt, col, val = orc.csr_rows(1, dets, r0, r0 + 16, thresh=EPS)
        lo, hi = int(rp[r0]), int(rp[r0 + 16])
        gc = np.empty(hi - lo, np.int32)
        gv = np.empty(hi - lo, np.float64)
        L.mhip_memcpy_d2h(ctx.h, gc.ctypes.data, d_ci + 4 * lo, gc.nbytes)
        L.mhip_memcpy_d2h(ctx.h, gv.ctypes.data, d_nz + 8 * lo, gv.nbytes)
        assert np.array_equal(gc, col)
        assert np.array_equal(gv, val)


def _bitset_less_sorted(dets, W):
    """The list in bitset_less order (bitset_operations.hpp:289-307): the
    beta string is the most significant part."""
    if W == 1:
        key = [(int(x[0]) >> 32, int(x[0]) & 0xffffffff) for x in dets]
    else:
        key = [(int(x[1]), int(x[0])) for x in dets]
    return dets[sorted(range(len(dets)), key=lambda i: key[i])]


@pytest.mark.parametrize("W,n,na,nb,order", [(1, 8, 4, 4, "generated"), (1, 8, 4, 3, "sorted"),
                                             (2, 10, 3, 2, "generated"), (2, 9, 4, 4, "sorted"),
                                             (1, 10, 5, 5, "generated"), (1, 7, 0, 3, "generated")])
def test_build_paths_agree(ctx, W, n, na, nb, order):
    """The product-space passes and the tiled double loop give the same CSR
    (pattern and bits) on complete spaces in both determinant orders."""
    from macis_amd import _capi

    announce(f"hamgen paths n={n} na={na} nb={nb} W={W} {order}")
    T, V = seeded_integrals(n, 40 + n)
    ctx.set_integrals(n, T, V)
    dets = D.generate_hilbert_space(W, n, na, nb)
    if order == "sorted":
        dets = _bitset_less_sorted(dets, W)
    dd = ctx.to_device(dets)
    out = {}
    for path in (_capi.Context.BUILD_TILED, _capi.Context.BUILD_PRODUCT, _capi.Context.BUILD_AUTO):
        ctx.set_build_path(path)
        H = ctx.ham_build(W, dd, len(dets), thresh=EPS)
        out[path] = (H.to_host(), ctx.timings()["ham_path"])
    ctx.set_build_path(_capi.Context.BUILD_AUTO)
    assert out[_capi.Context.BUILD_TILED][1] == _capi.Context.BUILD_TILED
    assert out[_capi.Context.BUILD_PRODUCT][1] == _capi.Context.BUILD_PRODUCT
    assert out[_capi.Context.BUILD_AUTO][1] == _capi.Context.BUILD_PRODUCT
    ref = out[_capi.Context.BUILD_TILED][0]
    for path in (_capi.Context.BUILD_PRODUCT, _capi.Context.BUILD_AUTO):
        assert_same_csr(out[path][0], ref)
    if len(dets) <= 5000:
        assert_same_csr(ref, O.Ham(T, V).csr(W, dets, thresh=EPS))


def test_product_path_only_for_product_lists(ctx, h2o, h2o_cisd):
    """A CISD list (not a product of spin strings), a list with a repeated
    determinant and a rectangular block take the tiled passes; forcing the
    product passes on them is an error."""
    from macis_amd import _capi

    announce("hamgen product-path recognition")
    norb, _, T, V = h2o
    ctx.set_integrals(norb, T, V)
    d = ctx.to_device(h2o_cisd)
    ctx.ham_build(1, d, len(h2o_cisd), thresh=1e-16)
    assert ctx.timings()["ham_path"] == _capi.Context.BUILD_TILED
    ctx.set_build_path(_capi.Context.BUILD_PRODUCT)
    try:
        with pytest.raises(_capi.MhipError):
            ctx.ham_build(1, d, len(h2o_cisd), thresh=1e-16)
        n = 6
        T6, V6 = seeded_integrals(n, 3)
        ctx.set_integrals(n, T6, V6)
        dets = D.generate_hilbert_space(1, n, 3, 3)
        rep = np.concatenate([dets, dets[:20]])  # 420 = 21 runs of 20: repeats
        with pytest.raises(_capi.MhipError):
            ctx.ham_build(1, ctx.to_device(rep), len(rep), thresh=EPS)
        ctx.set_build_path(_capi.Context.BUILD_AUTO)
        H = ctx.ham_build(1, ctx.to_device(rep), len(rep), thresh=EPS)
        assert ctx.timings()["ham_path"] == _capi.Context.BUILD_TILED
        assert_same_csr(H.to_host(), O.Ham(T6, V6).csr(1, rep, thresh=EPS))
    finally:
        ctx.set_build_path(_capi.Context.BUILD_AUTO)


def test_selected_ci_64_orbitals_properties(ctx):
    """Selected-CI list at ASCI scale (64 orbitals, 20a20b, wfn_t<128>),
    200,000 determinants: sampled rows equal the oracle's double loop, the
    pattern is symmetric with ascending columns, and the lowest root of the
    restarted Davidson is an eigenpair (Rayleigh quotient, residual)."""
    announce("hamgen 64 orbitals 20a20b, 200000 selected dets, W=2")
    n = 64
    T, V = seeded_integrals(n, 64)
    ctx.set_integrals(n, T, V)
    dets = selected_ci_dets(n, 20, 20, 200_000, seed=64, energies=0.25 * np.arange(n))
    N = len(dets)
    dd = ctx.to_device(dets)
    H = ctx.ham_build(2, dd, N, thresh=EPS)
    rp, ci, nz = H.to_host()
    orc = O.Ham(T, V)
    for r0 in (0, 77_777, N - 8):
        cnt, col, val = orc.csr_rows(2, dets, r0, r0 + 8, thresh=EPS)
        assert np.array_equal(np.diff(rp[r0:r0 + 9]), cnt)
        assert np.array_equal(ci[rp[r0]:rp[r0 + 8]], col)
        assert np.array_equal(nz[rp[r0]:rp[r0 + 8]], val)
    rows = np.repeat(np.arange(N, dtype=np.int64), np.diff(rp))
    assert np.all((np.diff(ci) > 0) | (np.diff(rows) > 0))
    key = rows * N + ci
    keyT = ci.astype(np.int64) * N + rows
    assert np.array_equal(np.sort(key), np.sort(keyT))
    dX = ctx.empty(N, np.float64).zero()
    E, nnz, _ = ctx.selected_ci_diag(2, dd, N, dX, max_restarts=200)
    assert nnz == H.nnz
    x = dX.numpy()
    assert abs(np.linalg.norm(x) - 1.0) < 1e-12
    Hx = np.zeros(N)
    np.add.at(Hx, rows, nz * x[ci])
    assert abs(x @ Hx - E) <= 1e-12 * abs(E)
    # the device residual is below 1e-8; the host product adds rounding of
    # order 1e-16 |H| |x| per row
    assert np.linalg.norm(Hx - E * x) <= 1e-8 + 1e-14 * abs(E) * np.sqrt(N)


@pytest.mark.parametrize("W", [1, 2])
@pytest.mark.parametrize("pool", ["0", "1024"])
def test_tiled_build_fallback_paths(ctx, pool, W, monkeypatch):
    # MHIP_STAGE_POOL_CHUNKS=0: no staging (count + fill walks); 1024 chunks
    # (131,072 entries) overflow on this list and the build redoes count +
    # fill. Both must give the staged build's CSR exactly, for one-word
    # (W = 1, alpha | beta << 32, the same bitset_less order) and two-word
    # determinants.
    announce(f"hamgen tiled fallback pool={pool} W={W}")
    n = 16
    T, V = seeded_integrals(n, 77)
    ctx.set_integrals(n, T, V)
    dets = selected_ci_dets(n, 5, 5, 4000, seed=77)
    if W == 1:
        dets = (dets[:, 0] | (dets[:, 1] << np.uint64(32))).reshape(-1, 1)
    dd = ctx.to_device(dets)
    ref = _build_with(ctx, ctx.BUILD_TILED, W, dd, len(dets))
    assert ref[0][-1] > 131072
    monkeypatch.setenv("MHIP_STAGE_POOL_CHUNKS", pool)
    got = _build_with(ctx, ctx.BUILD_TILED, W, dd, len(dets))
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    assert_same_csr(got, O.Ham(T, V).csr(W, dets, thresh=EPS))


@pytest.mark.parametrize("W", [1, 2])
@pytest.mark.parametrize("chunk,pool", [("4096", None), ("12288", None), ("4096", "1024")])
def test_tiled_scan_ket_chunks(ctx, W, chunk, pool, monkeypatch):
    """A long list's staged scan runs one launch per ket chunk, each row's
    hits appended chunk after chunk and its state stored between launches.
    MHIP_SCAN_KET_CHUNK forces chunks of 4,096 / 12,288 kets on a 20,000-det
    list (5 / 2 launches, the last chunk partial); with a 1,024-chunk pool
    the staging overflows during the chunked launches and the build redoes
    count + fill. The CSR equals the unchunked build's and the oracle's."""
    announce(f"hamgen ket-chunked scan W={W} chunk={chunk} pool={pool}")
    n = 16
    T, V = seeded_integrals(n, 79)
    ctx.set_integrals(n, T, V)
    dets = selected_ci_dets(n, 5, 5, 20000, seed=79, energies=0.25 * np.arange(n))
    if W == 1:
        dets = (dets[:, 0] | (dets[:, 1] << np.uint64(32))).reshape(-1, 1)
    dd = ctx.to_device(dets)
    ref = _build_with(ctx, ctx.BUILD_TILED, W, dd, len(dets))
    monkeypatch.setenv("MHIP_SCAN_KET_CHUNK", chunk)
    if pool:
        monkeypatch.setenv("MHIP_STAGE_POOL_CHUNKS", pool)
    got = _build_with(ctx, ctx.BUILD_TILED, W, dd, len(dets))
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    assert_same_csr(got, O.Ham(T, V).csr(W, dets, thresh=EPS))


def test_product_fill_non_uniform_inner_strings(ctx):
    """A product list whose inner strings are not a complete space (every
    third beta string dropped) has uneven inner neighbour counts; the
    product path still recognises it, and its CSR equals the oracle's and
    the tiled loop's."""
    from macis_amd import _capi

    announce("product fill, incomplete inner strings")
    n = 8
    T, V = seeded_integrals(n, 88)
    ctx.set_integrals(n, T, V)
    a = D.generate_combs(n, 4)
    b = D.generate_combs(n, 3)
    b = b[np.arange(len(b)) % 3 != 1]
    dets = np.array([[int(x) | (int(y) << 32)] for x in a for y in b], dtype=np.uint64)
    dd = ctx.to_device(dets)
    ctx.set_build_path(_capi.Context.BUILD_PRODUCT)
    try:
        prod = ctx.ham_build(1, dd, len(dets), thresh=EPS).to_host()
        assert ctx.timings()["ham_path"] == _capi.Context.BUILD_PRODUCT
        ctx.set_build_path(_capi.Context.BUILD_TILED)
        tiled = ctx.ham_build(1, dd, len(dets), thresh=EPS).to_host()
    finally:
        ctx.set_build_path(_capi.Context.BUILD_AUTO)
    assert_same_csr(prod, tiled)
    assert_same_csr(prod, O.Ham(T, V).csr(1, dets, thresh=EPS))


def _build_with(ctx, path, W, dd, n):
    ctx.set_build_path(path)
    try:
        H = ctx.ham_build(W, dd, n, thresh=EPS)
        got = H.to_host()
        assert ctx.timings()["ham_path"] == path
    finally:
        ctx.set_build_path(ctx.BUILD_AUTO)
    return got


def _one_word(dets):
    return (dets[:, 0] | (dets[:, 1] << np.uint64(32))).reshape(-1, 1)


@pytest.mark.parametrize("case", ["sci64", "sci16_w1", "shuffled_w1", "duplicates", "cas_w1"])
def test_spin_group_build_matches_tiled(ctx, case):
    """The spin-string group scan (rows scan the determinants that share
    their alpha or beta string and those of the single-excitation neighbour
    strings; columns sorted per row afterwards) gives the tiled double
    loop's CSR bit for bit: selected-CI lists at 64 and 16 orbitals (W = 2
    and W = 1), a list in random order, a list with repeated determinants
    and a complete space in generation order; the small lists also against
    the oracle."""
    rng = np.random.default_rng(5)
    if case == "sci64":
        n, W = 64, 2
        dets = selected_ci_dets(n, 20, 20, 12000, seed=11, energies=0.25 * np.arange(n))
    elif case == "sci16_w1":
        n, W = 16, 1
        dets = _one_word(selected_ci_dets(n, 5, 5, 4000, seed=12, energies=0.25 * np.arange(n)))
    elif case == "shuffled_w1":
        n, W = 16, 1
        dets = _one_word(selected_ci_dets(n, 6, 5, 6000, seed=13, energies=0.25 * np.arange(n)))
        dets = dets[rng.permutation(len(dets))]
    elif case == "duplicates":
        n, W = 32, 2
        dets = selected_ci_dets(n, 8, 8, 3000, seed=14, energies=0.25 * np.arange(n))
        dets = np.concatenate([dets, dets[rng.choice(len(dets), 40, replace=False)]])
        dets = dets[rng.permutation(len(dets))]
    else:
        n, W = 10, 1
        dets = D.generate_hilbert_space(1, n, 4, 5)
    announce(f"hamgen spin-group scan vs tiled: {case} ({len(dets)} dets, W={W})")
    T, V = seeded_integrals(n, 90 + n)
    ctx.set_integrals(n, T, V)
    dd = ctx.to_device(np.ascontiguousarray(dets))
    tiled = _build_with(ctx, ctx.BUILD_TILED, W, dd, len(dets))
    grp = _build_with(ctx, ctx.BUILD_SPIN_GROUPS, W, dd, len(dets))
    for a, b in zip(grp, tiled):
        assert np.array_equal(a, b)
    if len(dets) <= 6000:
        assert_same_csr(grp, O.Ham(T, V).csr(W, dets, thresh=EPS))


@pytest.mark.parametrize("thresh", [1e-3, 0.05])
def test_spin_group_threshold_compaction(ctx, thresh):
    """Elements at or below the threshold are dropped after the group scan
    by the shared compaction, as after the tiled loop: same CSR."""
    announce(f"hamgen spin-group scan with threshold {thresh}")
    n = 32
    T, V = seeded_integrals(n, 95)
    ctx.set_integrals(n, T, V)
    dets = selected_ci_dets(n, 8, 8, 3000, seed=16, energies=0.25 * np.arange(n))
    dd = ctx.to_device(dets)
    out = {}
    for path in (ctx.BUILD_TILED, ctx.BUILD_SPIN_GROUPS):
        ctx.set_build_path(path)
        try:
            out[path] = ctx.ham_build(2, dd, len(dets), thresh=thresh).to_host()
            assert ctx.timings()["ham_path"] == path
        finally:
            ctx.set_build_path(ctx.BUILD_AUTO)
    for a, b in zip(out[ctx.BUILD_SPIN_GROUPS], out[ctx.BUILD_TILED]):
        assert np.array_equal(a, b)
    full = O.Ham(T, V).csr(2, dets, thresh=EPS)
    assert len(out[ctx.BUILD_TILED][1]) < len(full[1])
    assert_same_csr(out[ctx.BUILD_SPIN_GROUPS], O.Ham(T, V).csr(2, dets, thresh=thresh))


@pytest.mark.parametrize("piece", ["50000", "333333"])
def test_spin_group_sort_in_pieces(ctx, piece, monkeypatch):
    """The per-row column sort runs in pieces of whole rows when the CSR
    holds more entries than one sort call indexes (2^31 - 1);
    MHIP_ROW_SORT_PIECE lowers the piece so a short list takes that path.
    The CSR equals the one-piece sort's."""
    announce(f"hamgen spin-group row sort in pieces of {piece}")
    n = 64
    T, V = seeded_integrals(n, 93)
    ctx.set_integrals(n, T, V)
    dets = selected_ci_dets(n, 20, 20, 20000, seed=15, energies=0.25 * np.arange(n))
    dd = ctx.to_device(dets)
    ref = _build_with(ctx, ctx.BUILD_SPIN_GROUPS, 2, dd, len(dets))
    assert ref[0][-1] > 3 * int(piece)
    monkeypatch.setenv("MHIP_ROW_SORT_PIECE", piece)
    got = _build_with(ctx, ctx.BUILD_SPIN_GROUPS, 2, dd, len(dets))
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_spin_group_refuses_varying_counts(ctx):
    """A list whose alpha electron count varies is not grouped: the forced
    path is an error, AUTO on a long such list takes the tiled loop, and a
    rectangular block cannot take the group scan."""
    from macis_amd import _capi

    announce("hamgen spin-group scan: varying electron counts, rectangular block")
    n = 12
    T, V = seeded_integrals(n, 94)
    ctx.set_integrals(n, T, V)
    a = D.generate_hilbert_space(1, n, 6, 6)
    b = D.generate_hilbert_space(1, n, 5, 6)
    dets = np.concatenate([a[:30000], b[:8000]])
    dd = ctx.to_device(dets)
    ctx.set_build_path(ctx.BUILD_SPIN_GROUPS)
    try:
        with pytest.raises(_capi.MhipError, match="electron counts"):
            ctx.ham_build(1, dd, len(dets), thresh=EPS)
    finally:
        ctx.set_build_path(ctx.BUILD_AUTO)
    H = ctx.ham_build(1, dd, len(dets), thresh=EPS)
    assert ctx.timings()["ham_path"] == ctx.BUILD_TILED
    del H
    ctx.set_build_path(ctx.BUILD_SPIN_GROUPS)
    try:
        with pytest.raises(_capi.MhipError, match="square"):
            ctx.ham_build(1, dd, 100, dd, 200, thresh=EPS)
    finally:
        ctx.set_build_path(ctx.BUILD_AUTO)
