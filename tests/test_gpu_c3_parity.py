"""Full-size parity of the 64-orbital selected-CI Hamiltonian (wfn_t<128>,
20 alpha + 20 beta electrons, 2,000,000 determinants in bitset_less order,
the benchmark's seeded list).

Every row of the device CSR (about 2.8e9 nonzeros) is checked against the
oracle's restatement of the reference double loop (double_loop.hpp:45-89):
the same columns in ascending order and bitwise equal values. The matrix is
then checked for symmetry through the device SpMV (x.Hy == y.Hx), and the
restarted Davidson's lowest root (davidson.hpp:128-225) for being an
eigenpair of the verified host matrix.
"""
import time

import numpy as np
import pytest

from macis_amd.synthetic import seeded_integrals, selected_ci_dets
from oracle import oracle as O

from conftest import announce

pytestmark = pytest.mark.gpu
EPS = np.finfo(float).eps
N_ORB, N_DETS, SEED = 64, 2_000_000, 64
BLOCKS = 8


@pytest.fixture(scope="module")
def c3(ctx):
    announce("C3 list generation, 2,000,000 dets")
    t0 = time.perf_counter()
    dets = selected_ci_dets(N_ORB, 20, 20, N_DETS, seed=SEED, energies=0.25 * np.arange(N_ORB))
    print(f"[c3] generated in {time.perf_counter() - t0:.1f} s", flush=True)
    T, V = seeded_integrals(N_ORB, SEED)
    announce("C3 tiled build + host copy")
    ctx.set_integrals(N_ORB, T, V)
    dd = ctx.to_device(dets)
    H = ctx.ham_build(2, dd, N_DETS, thresh=EPS)
    t0 = time.perf_counter()
    csr = H.to_host_native()
    print(f"[c3] nnz {H.nnz}, copied in {time.perf_counter() - t0:.1f} s", flush=True)
    yield T, V, dets, dd, H, csr
    del H
    ctx.release_cached()


@pytest.mark.parametrize("block", range(BLOCKS))
def test_c3_rows_match_oracle(c3, block):
    T, V, dets, _, _, (rp, ci, nz) = c3
    r0, r1 = block * N_DETS // BLOCKS, (block + 1) * N_DETS // BLOCKS
    t0 = time.perf_counter()
    bad, first = O.Ham(T, V).csr_verify_rows(2, dets, r0, r1, rp, ci, nz, thresh=EPS)
    print(f"[c3] rows {r0}..{r1}: {bad} differ ({time.perf_counter() - t0:.1f} s, "
          f"{O.num_threads()} threads)", flush=True)
    assert bad == 0, f"first differing row {first}"


def test_c3_symmetric_and_lowest_root_is_eigenpair(ctx, c3):
    T, V, dets, dd, H, (rp, ci, nz) = c3
    N = N_DETS
    announce("C3 symmetry through the device SpMV")
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal(N), rng.standard_normal(N)
    dx, dy = ctx.to_device(x), ctx.to_device(y)
    dHx, dHy = ctx.empty(N, np.float64).zero(), ctx.empty(N, np.float64).zero()
    ctx.spmv(H, 1, 1.0, dx, N, 0.0, dHx, N)
    ctx.spmv(H, 1, 1.0, dy, N, 0.0, dHy, N)
    a, b = float(x @ dHy.numpy()), float(y @ dHx.numpy())
    scale = float(np.abs(nz).max()) * np.sqrt(len(nz))
    assert abs(a - b) <= 1e-12 * scale, (a, b)
    announce("C3 SpMV on the column-panel split vs the host product")
    ref = np.zeros(N)
    O.gespmbv(1, 1.0, N, rp, ci, nz, 0, x, N, 0.0, ref, N)
    assert ctx.prepare_spmv(H, panels=True)
    ctx.spmv(H, 1, 1.0, dx, N, 0.0, dHx, N)
    lens = np.diff(rp)
    tol = 1e-13 * np.abs(nz).max() * np.abs(x).max() * lens.max()
    assert np.max(np.abs(dHx.numpy() - ref)) <= tol
    announce("C3 SpMV on the x-tile layout vs the host product")
    H2 = ctx.ham_build(2, dd, N, thresh=EPS)
    assert ctx.prepare_spmv(H2, xtiles=True)
    ctx.spmv(H2, 1, 1.0, dx, N, 0.0, dHx, N)
    assert np.max(np.abs(dHx.numpy() - ref)) <= tol
    del H2
    announce("C3 restarted Davidson, residual on the host matrix")
    ctx.set_integrals(N_ORB, T, V)
    dX = ctx.empty(N, np.float64).zero()
    E, nnz, _ = ctx.selected_ci_diag(2, dd, N, dX, max_restarts=200)
    assert nnz == len(nz)
    assert ctx.timings()["dav_sliced"] == 2  # the Davidson multiplied on column panels
    X = dX.numpy()
    assert abs(np.linalg.norm(X) - 1.0) < 1e-12
    HX = np.zeros(N)
    O.gespmbv(1, 1.0, N, rp, ci, nz, 0, X, N, 0.0, HX, N)
    assert abs(X @ HX - E) <= 1e-12 * abs(E)
    assert np.linalg.norm(HX - E * X) <= 1e-8 + 1e-14 * abs(E) * np.sqrt(N)
    print(f"[c3] E0 {E:.12f}, residual {np.linalg.norm(HX - E * X):.2e}", flush=True)
