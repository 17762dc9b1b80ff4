"""GPU parity of SpMV (K2), the MFMA tall-skinny GEMM (K3) and Davidson
(K3-K5) against the oracle and the reference's known answers
(test_spmbv.cxx:14-53, davidson.cxx:19-70)."""
import numpy as np
import pytest

from macis_amd import dets as D
from macis_amd.synthetic import seeded_integrals
from oracle import oracle as O

from conftest import announce

pytestmark = pytest.mark.gpu
EPS = np.finfo(float).eps


LAYOUTS = ["csr", "sliced", "panels", "xtiles"]


def _layout(ctx, A, layout):
    """Attach the requested product layout to A (the CSR kernel: nothing)."""
    if layout == "sliced":
        assert ctx.prepare_spmv(A, always=True)
    elif layout == "panels":
        assert ctx.prepare_spmv(A, panels=True)
    elif layout == "xtiles":
        assert ctx.prepare_spmv(A, xtiles=True)


@pytest.mark.parametrize("indexing", [0, 1])
@pytest.mark.parametrize("layout", LAYOUTS)
def test_spmv_known_answer(ctx, indexing, layout):
    announce(f"spmv 6x6 known answer indexing={indexing} layout={layout}")
    i = indexing
    rowptr = np.array([i, i + 2, i + 5, i + 8, i + 11, i + 14, i + 16])
    colind = np.array([0, 1, 0, 1, 2, 1, 2, 3, 2, 3, 4, 3, 4, 5, 4, 5]) + i
    nzval = np.array([2, 1, 1, 2, 1, 1, 2, 1, 1, 2, 1, 1, 2, 1, 1, 2], dtype=float)
    A = ctx.csr_from_host(6, 6, rowptr, colind, nzval, indexing=i)
    _layout(ctx, A, layout)
    x = ctx.to_device(np.arange(1, 7, dtype=float))
    y = ctx.to_device(np.ones(6))
    ctx.spmv(A, 1, 1.0, x, 6, 0.0, y, 6)
    assert np.array_equal(y.numpy(), [4, 8, 12, 16, 20, 17])
    Dg = ctx.empty(6, np.float64)
    ctx.extract_diagonal(A, Dg)
    assert np.array_equal(Dg.numpy(), np.full(6, 2.0))


@pytest.mark.parametrize("layout", LAYOUTS)
def test_spmv_block_alpha_beta(ctx, h2o, h2o_oracle, h2o_cisd, layout):
    announce(f"spmv H2O CISD K=2 alpha/beta layout={layout}")
    norb, _, T, V = h2o
    ctx.set_integrals(norb, T, V)
    N = len(h2o_cisd)
    H = ctx.ham_build(1, ctx.to_device(h2o_cisd), N, thresh=1e-16)
    _layout(ctx, H, layout)
    rp, ci, nz = h2o_oracle.csr(1, h2o_cisd, thresh=1e-16)
    rng = np.random.default_rng(1)
    LD = N + 3
    Vh = rng.standard_normal(2 * LD)
    AVh = rng.standard_normal(2 * LD)
    dV, dAV = ctx.to_device(Vh), ctx.to_device(AVh)
    ctx.spmv(H, 2, 0.7, dV, LD, -1.3, dAV, LD)
    ref = AVh.copy()
    O.gespmbv(2, 0.7, N, rp, ci, nz, 0, Vh, LD, -1.3, ref, LD)
    got = dAV.numpy()
    scale = np.abs(nz).max() * np.abs(Vh).max() * 4000
    assert np.max(np.abs(got - ref)) <= 1e-13 * scale
    assert np.array_equal(got[N:LD], ref[N:LD])  # padding untouched
    Dg = ctx.empty(N, np.float64)
    ctx.extract_diagonal(H, Dg)
    assert np.array_equal(Dg.numpy(), O.extract_diagonal(N, rp, ci, nz))


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("indexing", [0, 1])
def test_spmv_ragged_rows(ctx, indexing, layout):
    """Rows of every length class around the kernel's 16-byte quads (0-9
    nonzeros at all start offsets mod 4, long rows, an empty last row; in
    the sliced layout slices mixing empty, short and 4096-long rows and a
    partial last slice; in the panel split rows whose runs in a panel are
    empty or a few entries long) against the oracle's gespmbv
    (spmbv.hpp:37-73)."""
    announce(f"spmv ragged rows indexing={indexing} layout={layout}")
    rng = np.random.default_rng(7 + indexing)
    lens = np.concatenate([np.tile(np.arange(10), 8), rng.integers(0, 700, 300),
                           [1819, 4096, 3, 0]])
    m = n = len(lens)
    rowptr = np.zeros(m + 1, dtype=np.int64)
    rowptr[1:] = np.cumsum(lens)
    nnz = int(rowptr[-1])
    colind = np.concatenate([np.sort(rng.choice(n, size=L, replace=L > n))
                             for L in lens]).astype(np.int32)
    nzval = rng.standard_normal(nnz)
    A = ctx.csr_from_host(m, n, rowptr + indexing, colind + indexing, nzval, indexing=indexing)
    _layout(ctx, A, layout)
    x = rng.standard_normal(n)
    y0 = rng.standard_normal(m)
    dx, dy = ctx.to_device(x), ctx.to_device(y0)
    ctx.spmv(A, 1, 0.9, dx, n, 0.4, dy, m)
    ref = y0.copy()
    O.gespmbv(1, 0.9, m, rowptr + indexing, colind + indexing, nzval, indexing, x, n, 0.4, ref, m)
    got = dy.numpy()
    tol = 1e-14 * (np.abs(nzval).max() * np.abs(x).max() * lens.max() + 1)
    assert np.max(np.abs(got - ref)) <= tol


@pytest.mark.parametrize("indexing", [0, 1])
def test_spmv_xtiles_many_panels(ctx, indexing):
    """The x-tile layout over 61,440 columns (5 panels of 12,288 staged in
    LDS, the last one partial) and 4,100 rows (3 row blocks of 2,048, the
    last partial): ragged rows, empty rows, rows with no entry in some
    panels, a 30,000-entry row whose runs span several 4,096-entry pieces
    and rows that begin at a piece boundary; alpha/beta and 0/1-based
    indices against the oracle's gespmbv, and the same bits call to call."""
    announce(f"spmv x tiles, many panels, indexing={indexing}")
    rng = np.random.default_rng(21 + indexing)
    m, n = 4100, 61440
    lens = rng.integers(0, 400, m)
    lens[::7] = 0
    lens[1000] = 30000
    lens[2047:2050] = [4096, 1, 4095]
    rowptr = np.zeros(m + 1, dtype=np.int64)
    rowptr[1:] = np.cumsum(lens)
    nnz = int(rowptr[-1])

    def cols(i, L):
        if i % 5 == 1:  # only the first and last panels
            pool = np.concatenate([np.arange(0, 12288), np.arange(49152, n)])
            return np.sort(rng.choice(pool, size=L, replace=False))
        return np.sort(rng.choice(n, size=L, replace=False))

    colind = np.concatenate([cols(i, L) for i, L in enumerate(lens)]).astype(np.int32)
    nzval = rng.standard_normal(nnz)
    A = ctx.csr_from_host(m, n, rowptr + indexing, colind + indexing, nzval, indexing=indexing)
    from macis_amd._capi import MhipError

    with pytest.raises(MhipError, match="square"):  # m != n: no x tiles
        ctx.prepare_spmv(A, xtiles=True)
    # a square matrix: the same rows padded with empty rows up to n
    rp2 = np.concatenate([rowptr, np.full(n - m, rowptr[-1], dtype=np.int64)])
    B = ctx.csr_from_host(n, n, rp2 + indexing, colind + indexing, nzval, indexing=indexing)
    assert ctx.prepare_spmv(B, xtiles=True)
    x = rng.standard_normal(n)
    y0 = rng.standard_normal(n)
    dx = ctx.to_device(x)
    dy1, dy2 = ctx.to_device(y0), ctx.to_device(y0)
    ctx.spmv(B, 1, 0.7, dx, n, -0.3, dy1, n)
    ctx.spmv(B, 1, 0.7, dx, n, -0.3, dy2, n)
    got = dy1.numpy()
    assert np.array_equal(got, dy2.numpy())
    ref = y0.copy()
    O.gespmbv(1, 0.7, n, rp2 + indexing, colind + indexing, nzval, indexing, x, n, -0.3, ref, n)
    tol = 1e-14 * (np.abs(nzval).max() * np.abs(x).max() * lens.max() + 1)
    assert np.max(np.abs(got - ref)) <= tol


def test_spmv_sliced_copy_rules(ctx):
    """The sliced copy is made for square matrices only; the automatic policy
    refuses one whose padding would exceed a quarter of the nonzeros (one
    long row among short ones) and accepts equal rows; the forced copy of
    the padded matrix still multiplies correctly; products are bit-identical
    call to call."""
    announce("spmv sliced copy rules")
    rng = np.random.default_rng(3)
    rowptr = np.array([0, 2, 3], dtype=np.int64)
    colind = np.array([0, 2, 1], dtype=np.int32)
    R = ctx.csr_from_host(2, 3, rowptr, colind, np.array([1.0, 2.0, 3.0]))
    assert not ctx.prepare_spmv(R) and not ctx.prepare_spmv(R, always=True)

    def make(lens):
        m = len(lens)
        rp = np.zeros(m + 1, dtype=np.int64)
        rp[1:] = np.cumsum(lens)
        ci = np.concatenate([np.sort(rng.choice(m, size=L, replace=False)) for L in lens])
        return m, rp, ci.astype(np.int32), rng.standard_normal(int(rp[-1]))

    m, rp, ci, nz = make(np.full(1000, 10))  # 10,000 nonzeros, 10,240 stored
    A = ctx.csr_from_host(m, m, rp, ci, nz)
    assert ctx.prepare_spmv(A) and ctx.prepare_spmv(A)
    lens = np.full(2500, 4)
    lens[1700] = 2000  # one long row: its slice stores 64 x 2000 entries
    m, rp, ci, nz = make(lens)
    B = ctx.csr_from_host(m, m, rp, ci, nz)
    assert not ctx.prepare_spmv(B)
    x = rng.standard_normal(m)
    dx = ctx.to_device(x)
    # outputs start at zero: beta = 0 still reads y, as gespmbv does
    y0, y1, y2 = (ctx.empty(m, np.float64).zero() for _ in range(3))
    ctx.spmv(B, 1, 1.0, dx, m, 0.0, y0, m)  # CSR kernel
    assert ctx.prepare_spmv(B, always=True)
    ctx.spmv(B, 1, 1.0, dx, m, 0.0, y1, m)
    ctx.spmv(B, 1, 1.0, dx, m, 0.0, y2, m)
    assert np.array_equal(y1.numpy(), y2.numpy())
    ref = np.zeros(m)
    O.gespmbv(1, 1.0, m, rp, ci, nz, 0, x, m, 0.0, ref, m)
    tol = 1e-14 * np.abs(nz).max() * np.abs(x).max() * 2000
    assert np.max(np.abs(y1.numpy() - ref)) <= tol
    assert np.max(np.abs(y0.numpy() - ref)) <= tol


def test_spmv_sliced_narrow_and_wide_slices(ctx):
    """The sliced copy stores a slice's columns as 16-bit offsets from
    per-slot bases when every slot spans fewer than 65,536 columns and as
    32-bit columns otherwise. Banded rows of a 131,072-column matrix (all
    slices narrow) against the same matrix with a far zero-valued entry
    appended to the shortest row of every other slice (most of those slices
    wide; the slot counts and summation order unchanged): the products
    agree bit for bit and with the oracle's gespmbv; slots that span
    exactly 65,535 (narrow) and 65,536 (wide) columns are present."""
    announce("spmv sliced narrow / wide slices")
    rng = np.random.default_rng(11)
    m = n = 131072
    lens = rng.integers(0, 40, m)

    def row_cols(i, L):
        lo, hi = max(0, i - 20000), min(n, i + 20000)
        return np.unique(rng.integers(lo, hi, L))

    cols = [row_cols(i, L) for i, L in enumerate(lens)]
    vals = [rng.standard_normal(len(c)) for c in cols]
    # boundary spans: in slice 100, slot 0 spans 65,535 columns; in slice
    # 300, 65,536 (row 64 s + 1 starts at the slot's smallest column + span)
    for s, span in ((100, 65535), (300, 65536)):
        r0, r1 = 64 * s, 64 * s + 1
        c0 = 1000
        cols[r0] = np.array([c0, c0 + 5])
        cols[r1] = np.array([c0 + span, c0 + span + 3])
        vals[r0] = rng.standard_normal(2)
        vals[r1] = rng.standard_normal(2)
        for r in range(64 * s + 2, 64 * s + 64):
            cols[r] = np.array([c0 + 10 + r % 50, c0 + 100 + r % 50])
            vals[r] = rng.standard_normal(2)

    def csr(cols, vals):
        lens = np.array([len(c) for c in cols])
        rp = np.zeros(m + 1, dtype=np.int64)
        rp[1:] = np.cumsum(lens)
        return rp, np.concatenate(cols).astype(np.int32), np.concatenate(vals)

    rp, ci, nz = csr(cols, vals)
    cols_w, vals_w = list(cols), list(vals)
    for s in range(0, m // 64, 2):
        # the slice's shortest row, so its slot count (and every row's
        # summation order) stays; the new entry replaces a zero padding term
        ls = [len(cols[r]) for r in range(64 * s, 64 * s + 64)]
        r = 64 * s + int(np.argmin(ls))
        if min(ls) < max(ls) and (len(cols[r]) == 0 or cols[r][-1] < n - 1):
            cols_w[r] = np.append(cols[r], n - 1)
            vals_w[r] = np.append(vals[r], 0.0)
    rpw, ciw, nzw = csr(cols_w, vals_w)
    A = ctx.csr_from_host(m, n, rp, ci, nz)
    B = ctx.csr_from_host(m, n, rpw, ciw, nzw)
    assert ctx.prepare_spmv(A, always=True) and ctx.prepare_spmv(B, always=True)
    x = rng.standard_normal(n)
    dx = ctx.to_device(x)
    ya, yb = ctx.empty(m, np.float64).zero(), ctx.empty(m, np.float64).zero()
    ctx.spmv(A, 1, 1.0, dx, n, 0.0, ya, m)
    ctx.spmv(B, 1, 1.0, dx, n, 0.0, yb, m)
    ref = np.zeros(m)
    O.gespmbv(1, 1.0, m, rp, ci, nz, 0, x, n, 0.0, ref, m)
    got = ya.numpy()
    assert np.array_equal(got, yb.numpy())
    tol = 1e-14 * np.abs(nz).max() * np.abs(x).max() * 64
    assert np.max(np.abs(got - ref)) <= tol
    for s in (100, 300):  # the boundary rows, exactly as the oracle sums two terms
        for r in (64 * s, 64 * s + 1):
            assert abs(got[r] - ref[r]) <= 1e-15 * (abs(ref[r]) + 1)


@pytest.mark.parametrize("k,m", [(1, 1), (5, 1), (16, 16), (17, 17), (21, 21), (3, 29)])
def test_gemm_tn_mfma(ctx, k, m):
    announce(f"gemm_tn k={k} m={m}")
    rng = np.random.default_rng(k * 100 + m)
    N = 10007
    A = rng.standard_normal((k, N))  # column-major N x k == row-major k x N
    B = rng.standard_normal((m, N))
    Cd = ctx.empty(k * m, np.float64)
    ctx.gemm_tn(N, k, ctx.to_device(A), N, m, ctx.to_device(B), N, Cd, k)
    C = Cd.numpy().reshape(m, k).T
    ref = A @ B.T
    assert np.allclose(C, ref, rtol=0, atol=1e-11 * np.sqrt(N))


def test_davidson_h2o_cisd(ctx, h2o, h2o_oracle, h2o_cisd):
    announce("davidson H2O CISD max_m 15 tol 1e-8 (reference semantics)")
    norb, E_core, T, V = h2o
    ctx.set_integrals(norb, T, V)
    N = len(h2o_cisd)
    H = ctx.ham_build(1, ctx.to_device(h2o_cisd), N, thresh=1e-16)
    Dg = ctx.empty(N, np.float64)
    ctx.extract_diagonal(H, Dg)
    Dh = Dg.numpy()
    X0 = np.zeros(N)
    X0[np.argmin(Dh)] = 1.0
    X = ctx.to_device(X0)
    it, E0 = ctx.davidson(H, 15, 1e-8, X, Dg)
    # davidson.cxx:62 and the oracle
    assert E0 + E_core == pytest.approx(-76.23197835987, abs=1e-9)
    rp, ci, nz = h2o_oracle.csr(1, h2o_cisd, thresh=1e-16)
    ito, Eo, Xo = O.davidson(N, 15, rp, ci, nz, Dh, 1e-8, X0)
    assert E0 == pytest.approx(Eo, abs=1e-10)
    Xh = X.numpy()
    assert np.linalg.norm(Xh) == pytest.approx(1.0, abs=1e-12)
    assert abs(abs(Xh @ Xo) - 1.0) < 1e-9
    AX = np.zeros(N)
    O.gespmbv(1, 1.0, N, rp, ci, nz, 0, Xh, N, 0.0, AX, N)
    assert Xh @ AX == pytest.approx(E0, abs=1e-9)
    t = ctx.timings()
    # one product per iteration plus the guess's; the product of the next
    # correction is queued before the convergence test reads the norm
    assert it + 1 <= t["dav_spmv_calls"] <= it + 2 and t["dav_restarts"] == 0


def test_davidson_not_converged(ctx, h2o, h2o_cisd):
    announce("davidson non-convergence error path")
    from macis_amd import _capi

    norb, _, T, V = h2o
    ctx.set_integrals(norb, T, V)
    N = len(h2o_cisd)
    H = ctx.ham_build(1, ctx.to_device(h2o_cisd), N, thresh=1e-16)
    X = ctx.empty(N, np.float64).zero()
    Dg = ctx.empty(N, np.float64)
    ctx.extract_diagonal(H, Dg)
    X.upload(np.eye(1, N, int(np.argmin(Dg.numpy())))[0])
    with pytest.raises(_capi.MhipError, match="Did Not Converge") as e:
        ctx.davidson(H, 3, 1e-8, X, Dg)
    assert e.value.code == _capi.MHIP_ERR_NOT_CONVERGED


@pytest.mark.parametrize("n,ne,seed", [(8, 4, 108), (10, 5, 1)])
def test_selected_ci_diag_restarted_vs_oracle(ctx, n, ne, seed):
    announce(f"selected_ci_diag restarted CAS({2*ne},{n})")
    T, V = seeded_integrals(n, seed)
    ctx.set_integrals(n, T, V)
    dets = D.generate_hilbert_space(1, n, ne, ne)
    N = len(dets)
    X = ctx.empty(N, np.float64).zero()
    E, nnz, _ = ctx.selected_ci_diag(1, ctx.to_device(dets), N, X, max_restarts=100)
    Eo, Xo, nnzo = O.Ham(T, V).selected_ci_diag(1, dets, max_restarts=100)
    assert nnz == nnzo
    assert E == pytest.approx(Eo, abs=1e-8)
    assert abs(abs(X.numpy() @ Xo) - 1.0) < 1e-6
    t = ctx.timings()
    assert t["dav_residual"] < 1e-8
    # the restart forms A*[X, Xp] from the stored products: one SpMV per
    # iteration plus the initial one (and at most one queued ahead of the
    # convergence test), none at the restarts
    assert t["dav_restarts"] > 0
    assert t["dav_iters"] + 1 <= t["dav_spmv_calls"] <= t["dav_iters"] + 2
    assert t["dav_sliced"] == 1  # SpMVs on the sliced copy of the Hamiltonian


@pytest.mark.parametrize("max_m", [2, 3, 4, 6])
def test_davidson_restarted_small_subspace(ctx, max_m):
    """Tiny subspaces exercise every restart form (one column at max_m 2,
    the second-Ritz-vector fallback at 3): the result is an eigenpair of the
    dense matrix among its two lowest (the closed-shell guess may keep the
    iteration in one spin-flip sector), with the oracle's ground state when
    the oracle and the device land in the same sector."""
    announce(f"restarted davidson max_m {max_m}")
    n = 6
    T, V = seeded_integrals(n, 3)
    ctx.set_integrals(n, T, V)
    dets = D.generate_hilbert_space(1, n, 3, 3)
    N = len(dets)
    rp, ci, nz = O.Ham(T, V).csr(1, dets, thresh=1e-16)
    dense = np.zeros((N, N))
    for i in range(N):
        dense[i, ci[rp[i]:rp[i + 1]]] = nz[rp[i]:rp[i + 1]]
    w = np.linalg.eigvalsh(dense)
    H = ctx.ham_build(1, ctx.to_device(dets), N, thresh=1e-16)
    Dg = ctx.empty(N, np.float64)
    ctx.extract_diagonal(H, Dg)
    X0 = np.zeros(N)
    X0[np.argmin(Dg.numpy())] = 1.0
    X = ctx.to_device(X0)
    it, E = ctx.davidson(H, max_m, 1e-8, X, Dg, max_restarts=20000)
    assert np.min(np.abs(w[:2] - E)) < 1e-9
    Xh = X.numpy()
    assert np.linalg.norm(dense @ Xh - E * Xh) < 1e-7
    t = ctx.timings()
    assert it + 1 <= t["dav_spmv_calls"] <= it + 2
    _, Eo, _ = O.davidson(N, max_m, rp, ci, nz, Dg.numpy(), 1e-8, X0, max_restarts=20000)
    assert Eo == pytest.approx(w[0], abs=1e-9)


def test_selected_ci_diag_reuses_kept_storage(ctx):
    """The context keeps the internal Hamiltonian's storage between
    selected_ci_diag calls; reusing it (same list, then a smaller one that
    fits) and releasing it give bit-identical results."""
    from macis_amd import dets as D
    from macis_amd.synthetic import seeded_integrals

    announce("selected_ci_diag with kept Hamiltonian storage")
    n = 8
    T, V = seeded_integrals(n, 21)
    ctx.set_integrals(n, T, V)
    big = D.generate_hilbert_space(1, n, 4, 4)
    small = D.generate_hilbert_space(1, n, 4, 3)  # 3920 of 4900 dets: reuses the kept buffers
    out = []
    for dets in (big, big, small, big):
        dX = ctx.empty(len(dets), np.float64).zero()
        E, nnz, _ = ctx.selected_ci_diag(1, ctx.to_device(dets), len(dets), dX, max_restarts=200)
        out.append((E, nnz, dX.numpy()))
    ctx.release_cached()
    dX = ctx.empty(len(small), np.float64).zero()
    E, nnz, _ = ctx.selected_ci_diag(1, ctx.to_device(small), len(small), dX, max_restarts=200)
    assert out[0][0] == out[1][0] == out[3][0] and np.array_equal(out[0][2], out[1][2])
    assert np.array_equal(out[0][2], out[3][2])
    assert E == out[2][0] and nnz == out[2][1] and np.array_equal(dX.numpy(), out[2][2])


def _cas_csr(ctx, n, ne, seed):
    T, V = seeded_integrals(n, seed)
    ctx.set_integrals(n, T, V)
    dets = D.generate_hilbert_space(1, n, ne, ne)
    M = len(dets)
    H = ctx.ham_build(1, ctx.to_device(dets), M, thresh=1e-16)
    Dg = ctx.empty(M, np.float64)
    ctx.extract_diagonal(H, Dg)
    d = Dg.numpy()
    g = np.zeros(M)
    g[np.argmin(d)] = 1.0
    return H, Dg, d, g, O.Ham(T, V).csr(1, dets, thresh=1e-16)


def test_davidson_wide_subspace(ctx):
    """Subspaces wider than one 32-column projection block (davidson.hpp
    sizes V and AV for any max_m, :145-147). The synthetic CAS(8e,8o) of seed
    108 needs 73 iterations without restarts: with max_m 100 the reference
    algorithm converges on the device as in the oracle (subspaces up to 74
    columns, three column blocks), with max_m 70 both report non-convergence,
    and the restarted variant at max_m 35 agrees with the oracle's."""
    from macis_amd import _capi

    announce("davidson wide subspace, CAS(8,8) max_m 100 / 70 / restarted 35")
    H, Dg, d, g, (rp, ci, nz) = _cas_csr(ctx, 8, 4, 108)
    M = len(d)
    X = ctx.to_device(g)
    it, E = ctx.davidson(H, 100, 1e-8, X, Dg)
    ito, Eo, Xo = O.davidson(M, 100, rp, ci, nz, d, 1e-8, g)
    assert ito == 73 and abs(it - ito) <= 1
    assert E == pytest.approx(Eo, abs=1e-10)
    assert abs(abs(X.numpy() @ Xo) - 1.0) < 1e-8
    with pytest.raises(_capi.MhipError, match="Did Not Converge"):
        ctx.davidson(H, 70, 1e-8, ctx.to_device(g), Dg)
    Xr = ctx.to_device(g)
    itr, Er = ctx.davidson(H, 35, 1e-9, Xr, Dg, max_restarts=50)
    assert ctx.timings()["dav_restarts"] >= 1
    itro, Ero, _ = O.davidson(M, 35, rp, ci, nz, d, 1e-9, g, max_restarts=50)
    assert abs(itr - itro) <= 1
    assert Er == pytest.approx(Ero, abs=1e-10)


def test_davidson_subspace_limit(ctx, h2o, h2o_cisd):
    from macis_amd import _capi

    announce("davidson max_m above the 1024 limit is refused")
    norb, _, T, V = h2o
    ctx.set_integrals(norb, T, V)
    N = len(h2o_cisd)
    H = ctx.ham_build(1, ctx.to_device(h2o_cisd), N, thresh=1e-16)
    X = ctx.to_device(np.eye(1, N, 0)[0])
    with pytest.raises(_capi.MhipError, match=r"\[1, 1024\]"):
        ctx.davidson(H, 1025, 1e-8, X)
