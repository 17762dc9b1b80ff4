"""Batched CAS-CI on the GPU: block-diagonal build and lockstep Davidson
against per-instance oracle solves (energies within 1e-8 Ha), and the full
256-instance CAS(6e,6o)/(8e,8o)/(10e,10o) workload: every instance's CSR
block against the oracle's double loop, every batched energy against the
per-instance selected_ci_diag solve from the same guess, and the energies of
the CAS(6,6) and CAS(8,8) instances against an independent lowest
eigenvalue."""
import numpy as np
import pytest

from macis_amd import _capi
from macis_amd import dets as D
from macis_amd.synthetic import seeded_integrals
from oracle import oracle as O

from conftest import announce

pytestmark = pytest.mark.gpu
EPS = np.finfo(float).eps


def make_instances(sizes, seed0):
    norbs, dl, Ts, Vs = [], [], [], []
    for i, n in enumerate(sizes):
        T, V = seeded_integrals(n, seed0 + i)
        norbs.append(n)
        dl.append(D.generate_hilbert_space(1, n, n // 2, n // 2)[:, 0])
        Ts.append(T)
        Vs.append(V)
    return norbs, dl, Ts, Vs


def test_batch_build_matches_per_instance_oracle(ctx):
    announce("batch build: 8 instances CAS(4..10), blocks vs oracle")
    sizes = [6, 4, 8, 6, 10, 4, 8, 6]
    norbs, dl, Ts, Vs = make_instances(sizes, 500)
    B = _capi.Batch(ctx, norbs, dl, Ts, Vs)
    B.build(EPS)
    ndets, nnz = B.info()
    assert ndets == sum(len(d) for d in dl)
    ref_nnz = 0
    for i, (n, d, T, V) in enumerate(zip(norbs, dl, Ts, Vs)):
        rp, ci, nz = B.instance_csr(i)
        ref = O.Ham(T, V).csr(1, d.reshape(-1, 1), thresh=EPS)
        assert np.array_equal(rp, ref[0]), i
        assert np.array_equal(ci, ref[1]), i
        assert np.array_equal(nz.view(np.uint64), ref[2].view(np.uint64)), i
        ref_nnz += len(nz)
    assert nnz == ref_nnz


@pytest.mark.parametrize("use_graphs,layout", [(True, None), (False, None), (True, 0),
                                               (True, 1)])
def test_batch_davidson_energies(ctx, use_graphs, layout):
    """Energies against the per-instance oracle solve with the autotuner's
    SpMV layout (None) and with each layout forced (0: tuned CSR schedule,
    1: sliced copy); graphs hold one restart cycle each."""
    announce(f"batch davidson use_graphs={use_graphs} layout={layout}")
    sizes = [6, 8, 4, 6, 8, 6, 4, 8, 6, 6]
    norbs, dl, Ts, Vs = make_instances(sizes, 900)
    B = _capi.Batch(ctx, norbs, dl, Ts, Vs)
    B.build(EPS)
    B.tune()
    assert B.spmv_layout() in (0, 1)  # the tuner decided
    if layout is not None:
        assert B.spmv_layout(layout) == layout
    E, res, ms = B.davidson(max_m=20, max_restarts=200, tol=1e-8, use_graphs=use_graphs)
    assert np.all(res < 1e-8)
    for i, (n, d, T, V) in enumerate(zip(norbs, dl, Ts, Vs)):
        Eo, _, _ = O.Ham(T, V).selected_ci_diag(1, d.reshape(-1, 1), max_restarts=200)
        assert E[i] == pytest.approx(Eo, abs=1e-8), i
    # replaying the captured graphs gives the same answer as the first run
    E2, res2, _ = B.davidson(max_m=20, max_restarts=200, tol=1e-8, use_graphs=use_graphs)
    assert np.allclose(E2, E, rtol=0, atol=1e-10)


def test_batch_reference_semantics_not_converged(ctx):
    announce("batch davidson without restarts reports non-convergence")
    norbs, dl, Ts, Vs = make_instances([8, 8], 77)
    B = _capi.Batch(ctx, norbs, dl, Ts, Vs)
    B.build(EPS)
    with pytest.raises(_capi.MhipError, match="Did Not Converge"):
        B.davidson(max_m=5, max_restarts=0, tol=1e-12)


def test_batch_guess_and_vectors(ctx):
    """A caller-supplied guess replaces the diagonal guess; the returned
    Ritz vectors are normalised eigenvectors of each instance's block."""
    from macis_amd.synthetic import perturbed_diagonal_guess

    announce("batch davidson with a caller guess, Ritz vectors")
    sizes = [6, 8, 6]
    norbs, dl, Ts, Vs = make_instances(sizes, 300)
    B = _capi.Batch(ctx, norbs, dl, Ts, Vs)
    B.build(EPS)
    Dg = B.diagonal()
    X0 = np.concatenate([perturbed_diagonal_guess(Dg[B.offsets[i]:B.offsets[i + 1]], seed=i)
                         for i in range(len(sizes))])
    B.set_guess(X0)
    E, res, _ = B.davidson(max_m=20, max_restarts=200, tol=1e-8)
    X = B.vectors()
    for i in range(len(sizes)):
        rp, ci, nz = B.instance_csr(i)
        x = X[B.offsets[i]:B.offsets[i + 1]]
        assert np.array_equal(Dg[B.offsets[i]:B.offsets[i + 1]],
                              O.extract_diagonal(len(x), rp, ci, nz))
        Hx = O.gespmbv(1, 1.0, len(x), rp, ci, nz, 0, x, len(x), 0.0, np.zeros(len(x)), len(x))
        assert abs(np.linalg.norm(x) - 1) < 1e-12
        assert np.linalg.norm(Hx - E[i] * x) < 1e-8 + 1e-13 * abs(E[i])
        # the same guess through the per-instance path gives the same energy
        ctx.set_integrals(norbs[i], Ts[i], Vs[i])
        dX = ctx.to_device(X0[B.offsets[i]:B.offsets[i + 1]])
        Ep, _, _ = ctx.selected_ci_diag(1, ctx.to_device(dl[i].reshape(-1, 1)), len(x), dX,
                                        max_restarts=200)
        assert abs(Ep - E[i]) <= 1e-8
    B.set_guess(None)
    E2, _, _ = B.davidson(max_m=20, max_restarts=200, tol=1e-8)
    assert np.all(np.isfinite(E2))


@pytest.fixture(scope="module")
def c4(ctx):
    """The 256-instance workload of the benchmark (seed 2024)."""
    from macis_amd.synthetic import cas_batch_instances

    announce("C4: 256 CAS instances, batched build")
    norbs, dl, Ts, Vs = cas_batch_instances(256, 2024)
    B = _capi.Batch(ctx, norbs, dl, Ts, Vs)
    B.build(EPS)
    return norbs, dl, Ts, Vs, B


def test_c4_every_instance_block_matches_oracle(c4):
    import time

    norbs, dl, Ts, Vs, B = c4
    t0 = time.perf_counter()
    total = 0
    for i in range(len(norbs)):
        rp, ci, nz = B.instance_csr(i)
        d = dl[i].reshape(-1, 1)
        bad, first = O.Ham(Ts[i], Vs[i]).csr_verify_rows(1, d, 0, len(d), rp, ci, nz,
                                                         thresh=EPS)
        assert bad == 0, (i, norbs[i], first)
        total += len(nz)
    assert total == B.info()[1]
    print(f"[c4] 256 blocks, {total} nonzeros verified in {time.perf_counter() - t0:.1f} s",
          flush=True)


def test_c4_all_energies_match_per_instance_solves(ctx, c4):
    """Every one of the 256 batched energies within 1e-8 Ha of the
    per-instance selected_ci_diag solve started from the same guess; the
    CAS(6,6) and CAS(8,8) energies equal the lowest eigenvalue of their
    block (dense / Lanczos on the host copy); the CAS(10,10) energies are
    reproduced by the batch from a second, independent guess."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla

    from macis_amd.synthetic import perturbed_diagonal_guess

    norbs, dl, Ts, Vs, B = c4
    ninst = len(norbs)
    Dg = B.diagonal()
    seg = [slice(int(B.offsets[i]), int(B.offsets[i + 1])) for i in range(ninst)]
    X0 = np.concatenate([perturbed_diagonal_guess(Dg[seg[i]], seed=i) for i in range(ninst)])
    B.set_guess(X0)
    announce("C4 batched davidson, 256 instances")
    E, res, ms = B.davidson(max_m=20, max_restarts=200, tol=1e-8, use_graphs=True)
    assert np.all(res < 1e-8)
    announce("C4 per-instance selected_ci_diag, 256 solves")
    Ep = np.zeros(ninst)
    for i in range(ninst):
        ctx.set_integrals(norbs[i], Ts[i], Vs[i])
        d = dl[i].reshape(-1, 1)
        Ep[i], _, _ = ctx.selected_ci_diag(1, ctx.to_device(d), len(d),
                                           ctx.to_device(X0[seg[i]]), max_restarts=200)
    dE = np.abs(E - Ep)
    assert np.all(dE <= 1e-8), (int(np.argmax(dE)), float(dE.max()))
    announce("C4 lowest eigenvalues of the CAS(6,6) and CAS(8,8) blocks")
    for i in range(ninst):
        if norbs[i] == 10:
            continue
        rp, ci, nz = B.instance_csr(i)
        m = len(rp) - 1
        H = sp.csr_matrix((nz, ci, rp), shape=(m, m))
        if m <= 500:
            lo = np.linalg.eigvalsh(H.toarray())[0]
        else:
            lo = sla.eigsh(H, k=1, which="SA", tol=1e-13, v0=X0[seg[i]],
                           return_eigenvectors=False)[0]
        assert abs(E[i] - lo) <= 1e-8, (i, norbs[i], E[i], lo)
    announce("C4 second guess for the CAS(10,10) instances")
    X1 = np.concatenate([perturbed_diagonal_guess(Dg[seg[i]], seed=10_000 + i)
                         for i in range(ninst)])
    B.set_guess(X1)
    E1, res1, _ = B.davidson(max_m=20, max_restarts=200, tol=1e-8, use_graphs=True)
    B.set_guess(None)
    assert np.all(np.abs(E1 - E) <= 1e-8), float(np.abs(E1 - E).max())


def test_batch_128bit_instances(ctx):
    """wfn_t<128> instances (mhip_batch_create_w, words_per_det 2): complete
    CAS(6,6) / CAS(8,8) spaces and selected lists over 36 and 40 orbitals
    (norb > 32). Every block equals the oracle's CSR and every energy the
    per-instance selected_ci_diag solve from the same guess."""
    from macis_amd.synthetic import perturbed_diagonal_guess, selected_ci_dets

    announce("batch of wfn_t<128> instances, norb up to 40")
    norbs, dl, Ts, Vs = [], [], [], []
    for i, (n, kind) in enumerate([(6, "cas"), (36, "sel"), (8, "cas"), (40, "sel"), (36, "sel")]):
        T, V = seeded_integrals(n, 700 + i)
        if kind == "cas":
            d = D.generate_hilbert_space(2, n, n // 2, n // 2)
        else:
            d = selected_ci_dets(n, 4, 4, 1500 + 300 * i, seed=i, energies=0.5 * np.arange(n))
        norbs.append(n)
        dl.append(d)
        Ts.append(T)
        Vs.append(V)
    B = _capi.Batch(ctx, norbs, dl, Ts, Vs, W=2)
    B.build(EPS)
    for i in range(len(norbs)):
        rp, ci, nz = B.instance_csr(i)
        ref = O.Ham(Ts[i], Vs[i]).csr(2, dl[i], thresh=EPS)
        assert np.array_equal(rp, ref[0]) and np.array_equal(ci, ref[1]), i
        assert np.array_equal(nz.view(np.uint64), ref[2].view(np.uint64)), i
    Dg = B.diagonal()
    seg = [slice(int(B.offsets[i]), int(B.offsets[i + 1])) for i in range(len(norbs))]
    X0 = np.concatenate([perturbed_diagonal_guess(Dg[seg[i]], seed=i) for i in range(len(norbs))])
    B.set_guess(X0)
    E, res, _ = B.davidson(max_m=20, max_restarts=200, tol=1e-8)
    assert np.all(res < 1e-8)
    for i in range(len(norbs)):
        ctx.set_integrals(norbs[i], Ts[i], Vs[i])
        Ep, _, _ = ctx.selected_ci_diag(2, ctx.to_device(dl[i]), len(dl[i]),
                                        ctx.to_device(X0[seg[i]]), max_restarts=200)
        assert abs(Ep - E[i]) <= 1e-8, (i, E[i], Ep)
    # the one-word entry point refuses norb > 32
    with pytest.raises(_capi.MhipError, match="norb <= 32"):
        _capi.Batch(ctx, [36], [dl[1][:, 0]], [Ts[1]], [Vs[1]], W=1)


def test_batch_product_and_tiled_builds_agree(ctx):
    """The batched build walks the instances' spin-string neighbour tables
    (complete spaces; instances with the same list share the tables) and
    must give the block-diagonal CSR of the tiled per-instance double loop
    bit for bit; a batch with one list that is not a product space falls
    back to the tiled walk."""
    announce("batched build: product-space path vs tiled path")
    norbs, dl, Ts, Vs = make_instances([6, 8, 4, 6, 8, 4, 6], 321)
    B = _capi.Batch(ctx, norbs, dl, Ts, Vs)
    try:
        ctx.set_build_path(ctx.BUILD_TILED)
        B.build(EPS)
        assert ctx.timings()["ham_path"] == ctx.BUILD_TILED
        tiled = [B.instance_csr(i) for i in range(len(norbs))]
        ctx.set_build_path(ctx.BUILD_AUTO)
        B.build(EPS)
        assert ctx.timings()["ham_path"] == ctx.BUILD_PRODUCT
        for i in range(len(norbs)):
            got = B.instance_csr(i)
            assert np.array_equal(got[0], tiled[i][0]), i
            assert np.array_equal(got[1], tiled[i][1]), i
            assert np.array_equal(got[2].view(np.uint64), tiled[i][2].view(np.uint64)), i
        # a list that is not a complete product of spin strings (a truncated
        # complete space) keeps the whole batch on the tiled walk
        dl2 = list(dl)
        dl2[1] = dl2[1][: len(dl2[1]) - 3]
        B2 = _capi.Batch(ctx, norbs, dl2, Ts, Vs)
        B2.build(EPS)
        assert ctx.timings()["ham_path"] == ctx.BUILD_TILED
        rp, ci, nz = B2.instance_csr(1)
        ref = O.Ham(Ts[1], Vs[1]).csr(1, dl2[1].reshape(-1, 1), thresh=EPS)
        assert np.array_equal(rp, ref[0]) and np.array_equal(ci, ref[1])
    finally:
        ctx.set_build_path(ctx.BUILD_AUTO)
