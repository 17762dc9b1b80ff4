"""Benchmark of the MI355X determinant-CI hot path.

Workload (one GPU): CAS(12e,12o) FCI -- 12 spatial orbitals, 6 alpha + 6
beta electrons, all 853,776 determinants as wfn_t<64>, seeded synthetic FP64
integrals (macis_amd.synthetic.seeded_integrals(12, 12)). Determinants and
integrals are resident in HBM before the timed region.

One step = the Hamiltonian build (make_csr_hamiltonian: count, scan, fill,
elements; 1.553e9 nonzeros, 18.6 GB CSR) followed by a Davidson solve of
the lowest root to a residual of 1e-8 (subspace 20, the restarted variant:
the reference's no-restart Davidson cannot reach 1e-8 on these strongly
correlated integrals). `value` is the Hamiltonian-generation rate in nnz/s;
the SpMV HBM rate, the roofline of the SpMV kernel (HIP events on the
kernel's stream) and the CPU oracle baseline are reported beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0  # dense BF16 MFMA
# HBM bytes per spmv_sell_kernel dispatch on this workload from two rocprofv3
# --pmc runs (FETCH_SIZE and WRITE_SIZE, each scaled by what it reads for a
# 1 GiB device copy), summarised in profiles/cas12_fci_build_davidson.md:
# the sliced copy streams 16-bit column offsets (15.64 GB per product).
SPMV_TRAFFIC_GB = 16.14
N_ORB, N_ALPHA, N_BETA, SEED = 12, 6, 6, 12


def spmv_bytes(nnz, n, sliced):
    """Algorithmic bytes of one SpMV: values + int32 columns + the row
    structure (int64 row pointers of the CSR, or int64 offset + int32 length
    per 64-row slice of the sliced copy; padding entries not counted) + x read
    + y write (y is read when beta != 0; Davidson passes 0)."""
    rows = 12 * ((n + 63) // 64) if sliced else (n + 1) * 8
    return nnz * (8 + 4) + rows + 2 * n * 8


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(T, V, dets, budget_s=12.0):
    """Oracle restatement of the reference's serial double loop
    (double_loop.hpp:45-91), one pass per row as the reference's push_back
    loop, on a contiguous sample of bra rows x all kets."""
    from oracle import oracle as O

    orc = O.Ham(T, V)
    eps = np.finfo(float).eps
    rows, nnz, t_used = 0, 0, 0.0
    start = len(dets) // 3
    chunk = 16
    col = np.zeros(chunk * 4096, dtype=np.int32)  # >= 16 rows of 1,819 entries
    val = np.zeros(chunk * 4096)
    while t_used < budget_s and rows < 20000:
        t0 = time.perf_counter()
        cnt = orc.csr_rows_into(1, dets, start + rows, start + rows + chunk, col, val, thresh=eps)
        t_used += time.perf_counter() - t0
        rows += chunk
        nnz += int(cnt.sum())
    return {
        "value": nnz / t_used,
        "unit": "nnz/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{rows} contiguous bra rows x {len(dets)} kets of CAS(12e,12o), "
                  f"serial double loop in one pass (the reference's loop has no threading), "
                  f"{t_used:.1f} s; nnz/s extrapolates to the full build",
    }


def cpu_rdm_baseline(dets, X, budget_s=5.0):
    """The reference's serial form_rdms (the oracle's restatement of
    double_loop.hpp:116-165) for the first bra rows of the list against all
    kets, with the ground state of the timed solve, sized to about budget_s;
    the full-list time is extrapolated per bra row."""
    from oracle import oracle as O

    n = N_ORB
    rows = 32
    t0 = time.perf_counter()
    O.form_rdms(n, 1, dets[:rows], X, ket=dets)
    dt = time.perf_counter() - t0
    rows = int(min(len(dets), max(rows, rows * budget_s / max(dt, 1e-6))))
    t0 = time.perf_counter()
    O.form_rdms(n, 1, dets[:rows], X, ket=dets)
    dt = time.perf_counter() - t0
    full_ms = dt / rows * len(dets) * 1e3
    return {"value": full_ms, "unit": "ms (full list, extrapolated)", "cores": 1,
            "kind": "port",
            "sample": f"{rows} bra rows x {len(dets)} kets of CAS(12e,12o), serial double loop "
                      f"with rdm_contributions, {dt:.1f} s"}


def cpu_fair_baseline(T, V, dets, budget_s=6.0):
    """The same double loop with rows spread over the host's OpenMP threads
    (count pass + fill pass), not in the reference: the multi-core CPU
    figure beside the serial one."""
    from oracle import oracle as O

    orc = O.Ham(T, V)
    eps = np.finfo(float).eps
    threads = O.num_threads()
    chunk = 32 * threads
    col = np.zeros(chunk * 4096, dtype=np.int32)
    val = np.zeros(chunk * 4096)
    rows, nnz, t_used = 0, 0, 0.0
    start = len(dets) // 3
    while t_used < budget_s and start + rows + chunk <= len(dets):
        t0 = time.perf_counter()
        cnt = orc.csr_rows_parallel(1, dets, start + rows, start + rows + chunk, thresh=eps,
                                    col=col, val=val)
        t_used += time.perf_counter() - t0
        rows += chunk
        nnz += int(cnt.sum())
    return {
        "value": nnz / t_used, "unit": "nnz/s", "cores": threads, "kind": "port",
        "sample": f"{rows} contiguous bra rows x {len(dets)} kets, OpenMP row-parallel double "
                  f"loop (count + fill), {t_used:.1f} s",
    }


def cpu_spmv_baseline(ctx, T, V, d_dets, N, reps=3):
    """The reference's row-parallel gespmbv (spmbv.hpp:37-73, omp parallel
    for at :54) restated in the oracle, on the full CAS(12e,12o) CSR (built
    on the GPU, copied to host memory); then the oracle's Davidson
    (davidson.hpp:128-225, restarted variant, the benchmark's settings) on
    the same host matrix, whose energy the GPU energy is checked against."""
    from oracle import oracle as O

    ctx.set_integrals(N_ORB, T, V)  # the ASCI section switched the context to H2O
    H = ctx.ham_build(1, d_dets, N, thresh=np.finfo(float).eps)
    nnz = H.nnz
    rp, ci, nz = H.to_host_native()
    del H
    x = np.random.default_rng(0).standard_normal(N)
    y = np.zeros(N)
    O.gespmbv(1, 1.0, N, rp, ci, nz, 0, x, N, 0.0, y, N)
    t0 = time.perf_counter()
    for _ in range(reps):
        O.gespmbv(1, 1.0, N, rp, ci, nz, 0, x, N, 0.0, y, N)
    t = (time.perf_counter() - t0) / reps
    gbs = spmv_bytes(nnz, N, False) / t / 1e9
    Dg = O.extract_diagonal(N, rp, ci, nz)
    X0 = np.zeros(N)
    X0[int(np.argmin(Dg))] = 1.0
    t0 = time.perf_counter()
    it, Eo, _ = O.davidson(N, 20, rp, ci, nz, Dg, 1e-8, X0, max_restarts=200)
    dav_s = time.perf_counter() - t0
    return {
        "value": gbs, "unit": "GB/s", "ms_per_spmv": t * 1e3, "cores": O.num_threads(),
        "kind": "port",
        "sample": f"full CAS(12e,12o) SpMV ({nnz} nonzeros, CSR in host memory), {reps} "
                  f"products after one warm-up",
        "oracle_E0": Eo, "oracle_davidson_iterations": it, "oracle_davidson_s": dav_s,
    }


def rdm_timing(ctx, T, V, d_dets, N, nnz, d_X, E):
    """Spin-summed 1- and 2-RDMs of the CAS(12e,12o) ground state left by the
    last timed solve (form_rdms on the device, connected pairs = the
    Hamiltonian's pattern), timed after the timed region; the energy from
    the RDMs checks them."""
    n = N_ORB
    d1 = ctx.empty(n * n, np.float64).zero()
    d2 = ctx.empty(n ** 4, np.float64).zero()
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.form_rdms(1, d_dets, N, d_X, d1, d2)
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    o, t = d1.numpy(), d2.numpy()
    E_rdm = float(o @ T + t @ V)
    return {"ms": ms, "pairs": nnz, "pairs_per_s": nnz / (ms * 1e-3),
            "E_from_rdms": E_rdm, "dE": E_rdm - E,
            "trace_D1": float(np.trace(o.reshape(n, n, order="F"))),
            "note": "includes the structure pass that lists the pairs; FP64 atomics"}


def selected_ci_section(C, ctx, cpu=True, keep=None):
    """The 64-orbital 20a20b selected-CI configuration: 2,000,000 determinants
    from the seeded generator (HF + 1-4-fold excitations, level weights
    0.1/0.5/0.2/0.2, holes/particles ~ exp(-|e_p - e_F|) with e_p = 0.25 p,
    sorted by bitset_less), seeded integrals n = 64. Tiled device build
    (warm), one CSR SpMV, then selected_ci_diag to 1e-8 (restarted)."""
    from macis_amd.synthetic import seeded_integrals, selected_ci_dets

    n, nd = 64, 2_000_000
    t0 = time.perf_counter()
    T, V = seeded_integrals(n, 64)
    dets = selected_ci_dets(n, 20, 20, nd, seed=64, energies=0.25 * np.arange(n))
    setup_s = time.perf_counter() - t0
    ctx.set_integrals(n, T, V)
    dd = ctx.to_device(dets)
    eps = np.finfo(float).eps
    # warm-up build (pool, scan storage); its matrix serves the 8-panel split
    H8 = ctx.ham_build(2, dd, nd, thresh=eps)
    ctx.synchronize()
    t0 = time.perf_counter()
    H = ctx.ham_build(2, dd, nd, thresh=eps)
    build_ms = (time.perf_counter() - t0) * 1e3
    tm = ctx.timings()
    nnz = H.nnz
    rowptr = np.zeros(nd + 1, dtype=np.int64)
    C.check(C.lib().mhip_csr_copy_to_host_i64(ctx.h, H.h, rowptr.ctypes.data, None, None), "rp")
    x = ctx.to_device(np.random.default_rng(0).standard_normal(nd))
    y = ctx.empty(nd, np.float64).zero()
    ctx.spmv(H, 1, 1.0, x, nd, 0.0, y, nd)
    ctx.synchronize()
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.spmv(H, 1, 1.0, x, nd, 0.0, y, nd)
    ctx.synchronize()
    csr_spmv_ms = (time.perf_counter() - t0) * 1e3 / reps
    def layout_ms(A, **kw):
        """(prepare ms, ms per product) of a layout attached to A"""
        t0 = time.perf_counter()
        assert ctx.prepare_spmv(A, **kw)
        ctx.synchronize()
        prep = (time.perf_counter() - t0) * 1e3
        ctx.spmv(A, 1, 1.0, x, nd, 0.0, y, nd)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.spmv(A, 1, 1.0, x, nd, 0.0, y, nd)
        ctx.synchronize()
        return prep, (time.perf_counter() - t0) * 1e3 / reps

    # the layout the Davidson uses for this matrix: the 8-panel split (one
    # panel per XCD); beside it the panel-major copy with x tiles in LDS
    panel_prep_ms, spmv_ms = layout_ms(H8, panels=True)
    xtile_prep_ms, xtile_spmv_ms = layout_ms(H, xtiles=True)
    del H, H8, x, y
    dX = ctx.empty(nd, np.float64).zero()
    t0 = time.perf_counter()
    E, _, _ = ctx.selected_ci_diag(2, dd, nd, dX, max_restarts=200)
    diag_s = time.perf_counter() - t0
    td = ctx.timings()
    if keep is not None:  # the list and its ground state, for the c5 section
        keep.update(dets=dets, X=dX.numpy(), E0=E)
    cpu_leg = None
    if cpu:
        # the reference's serial double loop (oracle restatement) on
        # contiguous bra rows from a third of the way in x all kets; its
        # buffers take the exact row counts of the device CSR (equal by the
        # parity tests)
        from oracle import oracle as O

        orc = O.Ham(T, V)
        start, chunk, rows, nnz_c, t_used = nd // 3, 16, 0, 0, 0.0
        while t_used < 8.0 and rows < 20000:
            r0, r1 = start + rows, start + rows + chunk
            need = int(rowptr[r1] - rowptr[r0])
            col = np.zeros(max(need, 1), dtype=np.int32)
            val = np.zeros(max(need, 1))
            t1 = time.perf_counter()
            cnt = orc.csr_rows_into(2, dets, r0, r1, col, val, thresh=eps)
            t_used += time.perf_counter() - t1
            assert int(cnt.sum()) == need
            rows += chunk
            nnz_c += need
        cpu_leg = {"value": nnz_c / t_used, "unit": "nnz/s", "cores": 1, "kind": "port",
                   "sample": f"{rows} contiguous bra rows x {nd} kets, serial double loop, "
                             f"{t_used:.1f} s",
                   "gpu_speedup": (nnz / (build_ms * 1e-3)) / (nnz_c / t_used)}
        del orc
    return {
        "workload": "64 orbitals, 20a20b, 2,000,000 selected determinants (wfn_t<128>)",
        "nnz": nnz, "build_ms": build_ms, "nnz_per_s": nnz / (build_ms * 1e-3),
        "build_scan_ms": tm["ham_count_ms"], "build_gather_ms": tm["ham_fill_ms"],
        "build_elements_ms": tm["ham_elem_ms"],
        "spmv_ms": spmv_ms, "spmv_GBs": spmv_bytes(nnz, nd, False) / (spmv_ms * 1e-3) / 1e9,
        "spmv_layout": "column panels (8 nnz-balanced panels, one per XCD)",
        "panel_split_ms": panel_prep_ms,
        "xtile_spmv_ms": xtile_spmv_ms,
        "xtile_spmv_GBs": spmv_bytes(nnz, nd, False) / (xtile_spmv_ms * 1e-3) / 1e9,
        "xtile_prepare_ms": xtile_prep_ms,
        "csr_spmv_ms": csr_spmv_ms,
        "csr_spmv_GBs": spmv_bytes(nnz, nd, False) / (csr_spmv_ms * 1e-3) / 1e9,
        "davidson_layout": td["dav_sliced"],
        "selected_ci_diag_s": diag_s, "E0": E, "davidson_iterations": td["dav_iters"],
        "davidson_restarts": td["dav_restarts"], "setup_s_host": setup_s,
        "cpu_baseline": cpu_leg,
        "note": "build = host wall time of mhip_ham_build (second call); spmv on the "
                "column-panel split the Davidson attaches to this matrix (xtile_* the "
                "panel-major copy with x in LDS, csr_spmv_* the plain CSR kernel; GB/s of the "
                "CSR's algorithmic bytes); solve includes its own build",
    }


def batch_section(C, ctx):
    """256 independent CAS-CI instances, CAS(6e,6o) / CAS(8e,8o) / CAS(10e,10o)
    drawn uniformly (seeded), each with its own seeded integrals: one batched
    build, the autotuner, the batched restarted Davidson (hipGraph replay),
    then the same instances solved one after another with selected_ci_diag
    (the reference calls it once per instance). Both solves start from the
    same guess per instance: the diagonal guess plus a seeded 1e-3 component
    (synthetic.perturbed_diagonal_guess), so both reach the lowest
    eigenvalue of every spin sector."""
    from macis_amd.synthetic import cas_batch_instances, perturbed_diagonal_guess

    norbs, dl, Ts, Vs = cas_batch_instances(256, 2024)
    sizes = np.array(norbs)
    B = C.Batch(ctx, norbs, dl, Ts, Vs)
    eps = np.finfo(float).eps
    B.build(eps)  # warm-up
    ctx.synchronize()
    t0 = time.perf_counter()
    B.build(eps)
    ctx.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    nd, nnz = B.info()
    Dg = B.diagonal()
    X0 = np.concatenate([perturbed_diagonal_guess(Dg[B.offsets[i]:B.offsets[i + 1]], seed=i)
                         for i in range(len(norbs))])
    B.set_guess(X0)
    lanes = B.tune()
    layout = B.spmv_layout()
    B.davidson(max_m=20, max_restarts=200, tol=1e-8, use_graphs=True)  # graphs recorded
    t0 = time.perf_counter()
    E, res, _ = B.davidson(max_m=20, max_restarts=200, tol=1e-8, use_graphs=True)
    dav_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    Ep = np.zeros(len(sizes))
    for i in range(len(sizes)):
        ctx.set_integrals(norbs[i], Ts[i], Vs[i])
        d = dl[i].reshape(-1, 1)
        dX = ctx.to_device(X0[B.offsets[i]:B.offsets[i + 1]])
        Ep[i], _, _ = ctx.selected_ci_diag(1, ctx.to_device(d), len(d), dX, max_restarts=200)
    seq_ms = (time.perf_counter() - t0) * 1e3
    dE = np.abs(E - Ep)
    return {
        "workload": "256 CAS-CI instances, CAS(6e,6o)/(8e,8o)/(10e,10o) "
                    f"({int((sizes == 6).sum())}/{int((sizes == 8).sum())}/"
                    f"{int((sizes == 10).sum())}), own seeded integrals",
        "ndets": int(nd), "nnz": int(nnz), "build_ms": build_ms,
        "build_nnz_per_s": nnz / (build_ms * 1e-3), "davidson_ms": dav_ms,
        "spmv_layout": "sliced copy" if layout == 1 else "CSR, autotuned lanes per row",
        "tuned_lanes_per_bucket": [int(v) for v in lanes[:3]],
        "max_residual": float(res.max()), "sequential_solves_ms": seq_ms,
        "speedup_vs_sequential": seq_ms / (build_ms + dav_ms),
        "instances_within_1e-8": int((dE <= 1e-8).sum()),
        "max_abs_dE": float(dE.max()),
        "guess": "diagonal guess + seeded 1e-3 component, same vector for both solves",
        "note": "sequential = selected_ci_diag per instance on the same GPU (build + solve)",
    }


def mlp_throughput(C, ctx):
    """Forward (prediction) and training-step throughput of the pre-screen
    network on random 128-bit occupation strings."""
    W, n, nb = 2, 1 << 20, 65536
    rng = np.random.default_rng(5)
    d = ctx.to_device(rng.integers(0, 2**63, size=(n, W), dtype=np.uint64))
    t = ctx.to_device(rng.standard_normal(nb).astype(np.float32))
    out = ctx.empty(n, np.float32)
    net = C.MLP(ctx, seed=1)
    net.predict(W, d, n, out)
    net.train_step(W, d, t, nb, 0.0, want_loss=False)
    ctx.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        net.predict(W, d, n, out)
    ctx.synchronize()
    fwd_s = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        net.train_step(W, d, t, nb, 0.0, want_loss=False)
    ctx.synchronize()
    trn_s = (time.perf_counter() - t0) / reps
    fwd_flop = 2 * (128 * 512 + 512 * 512 + 512)          # per candidate
    bwd_flop = 2 * (2 * 512 * 512 + 128 * 512 + 2 * 512)  # dW2, dA1, dW1, head
    f_tf = fwd_flop * n / fwd_s / 1e12
    t_tf = (fwd_flop + bwd_flop) * nb / trn_s / 1e12
    return {
        "forward_rows": n, "forward_ms": fwd_s * 1e3, "forward_TFLOPs": f_tf,
        "train_batch": nb, "train_step_ms": trn_s * 1e3, "train_TFLOPs": t_tf,
        "mfma_util_forward": f_tf / BF16_PEAK_TFLOPS, "mfma_util_train": t_tf / BF16_PEAK_TFLOPS,
        "note": "wall time per call incl. the head and launches (inference forms the first "
                "layer's inputs from the determinant bits; training featurises, its steps "
                "queued back to back, loss not read); peak = dense BF16 MFMA",
    }


def _blocks_per_core(core, W):
    """Number of excitation blocks of each core determinant."""
    out = []
    for d in core:
        if W == 1:
            a, b = bin(int(d[0]) & 0xFFFFFFFF).count("1"), bin(int(d[0]) >> 32).count("1")
        else:
            a, b = bin(int(d[0])).count("1"), bin(int(d[1])).count("1")
        out.append(a + b + a * (a - 1) // 2 + b * (b - 1) // 2 + a * b)
    return np.array(out)


def _train_blocks(C, ctx, W, Xt, yt, epochs, batch, seed=3):
    """Network trained on standardised block targets; returns (net, mu, sd, s)."""
    mu, sd = float(yt.mean()), float(yt.std())
    yn = ((yt - mu) / sd).astype(np.float32)
    net = C.MLP(ctx, seed=seed)
    rng = np.random.default_rng(0)
    dd, dy = ctx.empty((len(yn), W), np.uint64), ctx.empty(len(yn), np.float32)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(epochs):
        p = rng.permutation(len(yn))
        dd.upload(Xt[p])
        dy.upload(yn[p])
        for b0 in range(0, len(yn) - batch + 1, batch):
            net.train_step(W, dd.ptr + 8 * W * b0, dy.ptr + 4 * b0, batch, 1e-3, want_loss=False)
    ctx.synchronize()
    return net, mu, sd, time.perf_counter() - t0


def _block_training_data(ctx, W, s, cores, E0, norb):
    Xs, ys = [], []
    for cc in cores:
        d_bd, d_t, nb = ctx.asci_blocks(W, s, ctx.to_device(cc), len(cc), E0, norb)
        t = d_t.numpy()[:nb]
        f = np.isfinite(t)
        Xs.append(d_bd.numpy()[:nb][f])
        ys.append(t[f])
    return np.concatenate(Xs), np.concatenate(ys).astype(np.float32)


def _sort_dets(d):
    return d[np.lexsort(tuple(d[:, k] for k in range(d.shape[1])))]


def asci_step_comparison(C, ctx):
    """One ASCI refinement step on H2O cc-pVDZ (core 100, 6400 dets): plain
    asci_search vs the learned block screen trained on the blocks of the
    three earlier plain steps' cores (HF -> 100 -> 800 -> 6400), and vs the
    screen that ranks blocks by |C_core| alone; re-diagonalised on the GPU."""
    from macis_amd import dets as D
    from macis_amd.fcidump import read_fcidump

    norb, _, T, V = read_fcidump(os.path.join(ROOT, "tests", "golden", "h2o.ccpvdz.fci.dat"))
    ctx.set_integrals(norb, T, V)
    W, s = 1, C.ASCISettings.defaults()

    def core_of(wfn, X):
        o = np.argsort(-np.abs(X), kind="stable")[:100]
        return wfn[o], X[o]

    def diag(dets):
        dets = _sort_dets(dets)
        dX = ctx.empty(len(dets), np.float64).zero()
        E, _, _ = ctx.selected_ci_diag(W, ctx.to_device(dets), len(dets), dX)
        return E, dX.numpy()

    hf = D.canonical_hf_determinant(W, 5, 5)
    dE = ctx.empty(1, np.float64)
    ctx.ham_diagonal(W, ctx.to_device(hf), 1, dE)
    E0, wfn, X = float(dE.numpy()[0]), hf, np.array([1.0])
    cores = []
    for nd in (100, 800, 6400):
        core, Cc = core_of(wfn, X)
        cores.append((core, E0))
        out = ctx.empty((nd, W), np.uint64)
        n = ctx.asci_search(W, s, nd, ctx.to_device(core), len(core), E0, ctx.to_device(Cc),
                            norb, out)
        wfn = out.numpy()[:n]
        E0, X = diag(wfn)
    Xs, ys = [], []
    for core, e in cores:
        a, b = _block_training_data(ctx, W, s, [core], e, norb)
        Xs.append(a)
        ys.append(b)
    Xt, yt = np.concatenate(Xs), np.concatenate(ys)
    net, mu, sd, train_s = _train_blocks(C, ctx, W, Xt, yt, 20, 1024)
    nd = len(wfn)
    core, Cc = core_of(wfn, X)
    dcore, dC = ctx.to_device(core), ctx.to_device(Cc)
    coeff_scores = ctx.to_device(np.repeat(np.log(np.abs(Cc)), _blocks_per_core(core, W))
                                 .astype(np.float32))

    def run(kind, keep):
        out = ctx.empty((nd, W), np.uint64)
        ts = []
        for _ in range(5):
            if kind == "plain":
                n = ctx.asci_search(W, s, nd, dcore, len(core), E0, dC, norb, out)
            elif kind == "nn":
                n = ctx.asci_search_screened(W, s, net, keep, nd, dcore, len(core), E0, dC, norb,
                                             out, scale=sd, shift=mu)
            else:
                n = ctx.asci_search_blocks(W, s, coeff_scores, keep, nd, dcore, len(core), E0,
                                           dC, norb, out)
            ts.append(ctx.timings()["asci_total_ms"])
        return out.numpy()[:n], sorted(ts)[2], ctx.timings()

    plain, t_plain, tm_p = run("plain", 1.0)
    E_plain = diag(plain)[0]
    pset = set(plain[:, 0].tolist())
    res = {}
    for kind in ("nn", "coeff"):
        for keep in (0.05, 0.2, 0.3):
            got, t, tm = run(kind, keep)
            res[f"{kind}_keep{keep}"] = {
                "search_ms": t, "speedup": t_plain / t,
                "recall": len(pset & set(got[:, 0].tolist())) / len(plain),
                "dE_mHa": (diag(got)[0] - E_plain) * 1e3,
                "raw_candidates": tm["asci_npairs"],
                "blocks_kept": tm["asci_nblocks_kept"], "blocks": tm["asci_nblocks"],
            }
    return {
        "system": "H2O cc-pVDZ (24 orbitals, 5a5b), core 100, 6400 dets (refinement step)",
        "training_blocks": int(len(yt)), "train_s": train_s,
        "plain_search_ms": t_plain, "plain_raw_candidates": tm_p["asci_npairs"],
        "screens": res,
        "note": "nn: blocks ranked by log|C| + network prediction; coeff: by log|C| alone",
    }


def c5_section(C, ctx, sel):
    """The learned block screen at the 64-orbital scale: the 2,000,000-det
    selected-CI list and its ground state from the selected_ci section; the
    core is its 1,000 largest-|C| determinants (1.14e9 raw candidates, the
    config's 'about 1e9'), the search keeps 2,000,000 determinants. Training:
    the blocks of the cores of two earlier, smaller steps (100 and 300
    determinants). Plain search with the reference's pair_size_max pruning
    replay (default settings) and without it (pair_size_max above the raw
    count: the same output here); screened searches at several keep
    fractions, the |C|-only block ranking and the perfect block ranking (the
    true block targets); recall against the plain top-k and the variational
    energy difference after re-diagonalising each 2e6-det set."""
    dets, X, E, n = sel["dets"], sel["X"], sel["E0"], 64
    W, nd, ncore = 2, 2_000_000, 1000
    s = C.ASCISettings.defaults()
    snp = C.ASCISettings.defaults(pair_size_max=2**62)
    order = np.argsort(-np.abs(X), kind="stable")
    Xt, yt = _block_training_data(ctx, W, s, [dets[order[:100]], dets[order[:300]]], E, n)
    net, mu, sd, train_s = _train_blocks(C, ctx, W, Xt, yt, 20, 4096)
    core, Cc = dets[order[:ncore]], X[order[:ncore]]
    dcore, dC = ctx.to_device(core), ctx.to_device(Cc)
    nper = _blocks_per_core(core, W)
    lc = np.repeat(np.log(np.abs(Cc)), nper).astype(np.float32)
    _, d_t, nblk = ctx.asci_blocks(W, s, dcore, ncore, E, n)
    true_scores = ctx.to_device((d_t.numpy()[:nblk] + lc).astype(np.float32))
    coeff_scores = ctx.to_device(lc)

    def run(kind, keep=1.0, reps=3):
        out = ctx.empty((nd, W), np.uint64)
        ts = []
        for _ in range(reps):
            if kind == "plain":
                m = ctx.asci_search(W, s, nd, dcore, ncore, E, dC, n, out)
            elif kind == "plain_nopruning":
                m = ctx.asci_search(W, snp, nd, dcore, ncore, E, dC, n, out)
            elif kind == "nn":
                m = ctx.asci_search_screened(W, s, net, keep, nd, dcore, ncore, E, dC, n, out,
                                             scale=sd, shift=mu)
            else:
                m = ctx.asci_search_blocks(W, s, true_scores if kind == "perfect" else coeff_scores,
                                           keep, nd, dcore, ncore, E, dC, n, out)
            ts.append(ctx.timings()["asci_total_ms"])
        return out.numpy()[:m], sorted(ts)[len(ts) // 2], ctx.timings()

    def energy(d):
        d = _sort_dets(d)
        dX = ctx.empty(len(d), np.float64).zero()
        e, _, _ = ctx.selected_ci_diag(W, ctx.to_device(d), len(d), dX, max_restarts=200)
        return e

    plain, t_plain, tm_p = run("plain")
    plain2, t_np, _ = run("plain_nopruning")
    pset = set(map(tuple, plain.tolist()))
    same_np = len(pset & set(map(tuple, plain2.tolist()))) / len(plain)
    E_plain = energy(plain)
    res = {}
    for kind, keeps in (("nn", (0.05, 0.2, 0.5)), ("coeff", (0.05,)), ("perfect", (0.05,))):
        for keep in keeps:
            got, t, tm = run(kind, keep)
            r = {"search_ms": t, "speedup_vs_plain": t_plain / t,
                 "speedup_vs_plain_nopruning": t_np / t,
                 "recall": len(pset & set(map(tuple, got.tolist()))) / len(plain),
                 "raw_candidates": tm["asci_npairs"], "blocks_kept": tm["asci_nblocks_kept"],
                 "score_ms": tm["asci_nn_ms"]}
            if kind == "nn" and keep in (0.05, 0.5):
                r["dE_mHa"] = (energy(got) - E_plain) * 1e3
            if kind == "nn":
                # network forward over every block: 656,384 flop each
                r["network_TFLOPs"] = nblk * 656384 / (tm["asci_nn_ms"] * 1e-3) / 1e12
            res[f"{kind}_keep{keep}"] = r
    return {
        "workload": "64 orbitals, 20a20b; core 1,000 of the 2,000,000-det selected-CI ground "
                    "state; ndets_max 2,000,000",
        "raw_candidates": tm_p["asci_npairs"], "unique_candidates": tm_p["asci_nunique"],
        "blocks": int(nblk), "training_blocks": int(len(yt)), "train_s": train_s,
        "plain_search_ms": t_plain, "plain_nopruning_search_ms": t_np,
        "plain_nopruning_recall": same_np,
        "plain_phases_ms": {"generate": tm_p["asci_contrib_ms"], "sort_accumulate":
                            tm_p["asci_sort_ms"], "topk": tm_p["asci_topk_ms"]},
        "E_plain": E_plain, "screens": res,
        "bar": "dE <= 0.1 mHa: met by no measured keep fraction below 1 (every block kept "
               "reproduces the plain search)",
    }


def spmv_stream_floor(ctx, N, row_len, reps=10):
    """The roofline kernel's own streaming floor: the sliced SpMV on a matrix
    of the CAS(12e,12o) shape (N rows of row_len entries) whose column indices all point
    at x[0], so every x gather hits cache and only the 12 streamed bytes per
    entry remain. Returns (ms per launch, GB/s of its algorithmic bytes)."""
    nnz = N * row_len
    rowptr = np.arange(0, nnz + 1, row_len, dtype=np.int64)
    colind = np.zeros(nnz, dtype=np.int32)
    nzval = np.ones(nnz, dtype=np.float64)
    A = ctx.csr_from_host(N, N, rowptr, colind, nzval)
    del colind, nzval
    assert ctx.prepare_spmv(A, always=True)
    x = ctx.to_device(np.ones(N))
    y = ctx.empty(N, np.float64).zero()
    ctx.spmv(A, 1, 1.0, x, N, 0.0, y, N)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.spmv(A, 1, 1.0, x, N, 0.0, y, N)
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    del A, x, y
    return ms, spmv_bytes(nnz, N, True) / (ms * 1e-3) / 1e9


def write_floor(ctx, nbytes, reps=5):
    """The device's write-only rate on a buffer as large as the CSR a build
    writes: mhip_memset (the runtime's fill kernel) over nbytes, average of
    reps launches after one warm-up. Returns (ms, GB/s)."""
    buf = ctx.empty(nbytes // 8 + 1, np.float64)
    buf.zero()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        buf.zero()
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    del buf
    return ms, nbytes / (ms * 1e-3) / 1e9


def c1_section(C, ctx):
    """CAS(10e,10o) (63,504 determinants, seeded integrals n = 10, seed 10) timed
    in full on both sides: the GPU's selected_ci_diag (build + restarted
    Davidson to 1e-8) and, on the host, the oracle's serial double loop
    (double_loop.hpp:30-98, one pass per row as the reference pushes its
    entries) alone, then its selected_ci_diag (that build + the restarted
    Davidson with the reference's row-parallel gespmbv). No extrapolation."""
    from macis_amd import dets as D
    from macis_amd.synthetic import seeded_integrals
    from oracle import oracle as O

    n = 10
    T, V = seeded_integrals(n, 10)
    dets = D.generate_hilbert_space(1, n, 5, 5)
    N = len(dets)
    eps = np.finfo(float).eps
    ctx.set_integrals(n, T, V)
    dd = ctx.to_device(dets)
    dX = ctx.empty(N, np.float64)
    runs = []
    for _ in range(4):
        dX.zero()
        ctx.synchronize()
        t0 = time.perf_counter()
        E, nnz, _ = ctx.selected_ci_diag(1, dd, N, dX, h_el_tol=eps, max_m=20, res_tol=1e-8,
                                         max_restarts=200)
        ctx.synchronize()
        runs.append(((time.perf_counter() - t0) * 1e3, ctx.timings()["ham_total_ms"]))
    gpu_ms, gpu_build_ms = sorted(runs[1:])[1]
    orc = O.Ham(T, V)
    t0 = time.perf_counter()
    _, ci, _ = orc.csr(1, dets, thresh=eps)
    cpu_build_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    E_cpu, _, _ = orc.selected_ci_diag(1, dets, max_restarts=200)
    cpu_solve_s = time.perf_counter() - t0
    return {
        "workload": "CAS(10e,10o), 63,504 determinants, H build + Davidson (subspace 20, "
                    "restarted, tol 1e-8), seeded integrals n=10 seed 10",
        "nnz": nnz, "E0": E, "E0_cpu": E_cpu, "abs_dE": abs(E - E_cpu),
        "gpu_solve_ms": gpu_ms, "gpu_build_ms": gpu_build_ms,
        "gpu_build_nnz_per_s": nnz / (gpu_build_ms * 1e-3),
        "cpu_build_s": cpu_build_s, "cpu_build_nnz_per_s": len(ci) / cpu_build_s,
        "cpu_build_threads": 1, "cpu_solve_s": cpu_solve_s,
        "cpu_solve_threads": O.lib().orc_num_threads(),
        "solve_speedup": cpu_solve_s / (gpu_ms * 1e-3),
        "note": "both sides timed in full (no sampling); cpu_solve includes its own serial build",
    }


def c5_only(C, ctx):
    """The c5 section on its own: the selected-CI list and ground state first."""
    from macis_amd.synthetic import seeded_integrals, selected_ci_dets

    n, nd = 64, 2_000_000
    T, V = seeded_integrals(n, 64)
    dets = selected_ci_dets(n, 20, 20, nd, seed=64, energies=0.25 * np.arange(n))
    ctx.set_integrals(n, T, V)
    dX = ctx.empty(nd, np.float64).zero()
    E, _, _ = ctx.selected_ci_diag(2, ctx.to_device(dets), nd, dX, max_restarts=200)
    return c5_section(C, ctx, {"dets": dets, "X": dX.numpy(), "E0": E})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 256-instance batched CAS-CI section")
    ap.add_argument("--no-selected-ci", action="store_true",
                    help="skip the 64-orbital 2M-determinant selected-CI section")
    ap.add_argument("--max-restarts", type=int, default=200)
    ap.add_argument("--no-asci", action="store_true",
                    help="skip the learned pre-screen / ASCI step comparison")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="one 1 GiB device-to-device copy first (counter calibration)")
    ap.add_argument("--no-c5", action="store_true",
                    help="skip the 64-orbital learned block screen section")
    ap.add_argument("--only", choices=["selected_ci", "batch", "asci", "mlp", "c5", "c1"],
                    help="run only this section and print its JSON object (for profiling; "
                         "not the benchmark line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the hot path has no "
              "multi-GPU launch); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from macis_amd import capi
    from macis_amd import dets as D
    from macis_amd.synthetic import seeded_integrals

    C = capi()
    if args.only:
        ctx = C.Context(0)
        run = {"selected_ci": lambda: selected_ci_section(C, ctx, cpu=not args.no_cpu_baseline),
               "batch": lambda: batch_section(C, ctx),
               "asci": lambda: asci_step_comparison(C, ctx),
               "mlp": lambda: mlp_throughput(C, ctx),
               "c5": lambda: c5_only(C, ctx),
               "c1": lambda: c1_section(C, ctx)}[args.only]
        print(json.dumps({"section": args.only, args.only: run()}), flush=True)
        return
    T, V = seeded_integrals(N_ORB, SEED)
    dets = D.generate_hilbert_space(1, N_ORB, N_ALPHA, N_BETA)
    N = len(dets)
    eps = np.finfo(float).eps

    ctx = C.Context(0)
    ctx.set_integrals(N_ORB, T, V)
    d_dets = ctx.to_device(dets)
    d_X = ctx.empty(N, np.float64)
    ctx.synchronize()
    if args.calibrate_copy:
        a = ctx.empty(1 << 27, np.float64).zero()
        b = ctx.empty(1 << 27, np.float64)
        C.check(C.lib().mhip_memcpy_d2d(ctx.h, b.ptr, a.ptr, a.nbytes), "copy")
        ctx.synchronize()
        del a, b

    def step():
        d_X.zero()
        E, nnz, _ = ctx.selected_ci_diag(1, d_dets, N, d_X, h_el_tol=eps, max_m=20,
                                         res_tol=1e-8, max_restarts=args.max_restarts)
        return E, nnz, ctx.timings()

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    t0 = time.perf_counter()
    recs = [step() for _ in range(args.steps)]
    ctx.synchronize()
    wall = time.perf_counter() - t0

    E, nnz, _ = recs[-1]
    build_ms = sum(r[2]["ham_total_ms"] for r in recs)
    spmv_ms = sum(r[2]["dav_spmv_ms"] for r in recs)
    spmv_calls = sum(r[2]["dav_spmv_calls"] for r in recs)
    dav_ms = sum(r[2]["dav_total_ms"] for r in recs)
    avg_spmv_ms = spmv_ms / max(spmv_calls, 1)
    sliced = all(r[2]["dav_sliced"] == 1 for r in recs)
    bytes_spmv = spmv_bytes(nnz, N, sliced)
    spmv_gbs = bytes_spmv / (avg_spmv_ms * 1e-3) / 1e9
    value = nnz * args.steps / (build_ms * 1e-3)
    last = recs[-1][2]
    out = {
        "metric": "Ham-gen nnz/s; SpMV HBM GB/s; ASCI step time NN-screened vs plain at 0.1 mHa",
        "value": value,
        "unit": "nnz/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded integrals, full CAS(12e,12o) determinant list)",
        "config": {
            "workload": "CAS(12e,12o) FCI, 853,776 dets, H build + Davidson (subspace 20, "
                        "restarted, tol 1e-8)",
            "ndets": N,
            "nnz": nnz,
            "norb": N_ORB,
            "integral_seed": SEED,
        },
        "hamgen": {
            "ms_per_build": build_ms / args.steps,
            "nnz_per_s": value,
            "bounds_ms": last["ham_bounds_ms"],
            "count_ms": last["ham_count_ms"],
            "scan_ms": last["ham_scan_ms"],
            "fill_ms": last["ham_fill_ms"],
            "element_ms": last["ham_elem_ms"],
            "compact_ms": last["ham_compact_ms"],
            "written_GB": (nnz * 12 + (N + 1) * 8) / 1e9,
        },
        "spmv": {
            "avg_ms": avg_spmv_ms,
            "calls_per_step": spmv_calls / args.steps,
            "algorithmic_GB": bytes_spmv / 1e9,
            "hbm_GBps": spmv_gbs,
            "frac_of_peak": spmv_gbs / HBM_PEAK_GBS,
            # what the layout actually streams: 16-bit column offsets in the
            # narrow slices of the sliced copy (x read and y written added)
            "streamed_GB": (last["dav_spmv_matrix_bytes"] + 2 * N * 8) / 1e9,
            "streamed_GBps": (last["dav_spmv_matrix_bytes"] + 2 * N * 8) / (avg_spmv_ms * 1e6),
        },
        "davidson": {
            "ms_per_solve": dav_ms / args.steps,
            "iterations": last["dav_iters"],
            "restarts": last["dav_restarts"],
            "sliced_copy_ms": last["dav_layout_ms"],
            "residual": last["dav_residual"],
            "E0": E,
        },
        "roofline": {
            "kernel": ("spmv_sell_kernel (FP64 SpMV on the 64-row sliced copy of the CSR "
                       "Hamiltonian)" if sliced else "spmv_kernel (FP64 CSR SpMV)"),
            "bound": "hbm",
            "achieved": spmv_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": spmv_gbs / HBM_PEAK_GBS,
            "traffic": SPMV_TRAFFIC_GB * 1e9 if SPMV_TRAFFIC_GB and sliced else None,
        },
    }
    out["rdms"] = rdm_timing(ctx, T, V, d_dets, N, nnz, d_X, E)
    # what writing the build's bytes alone takes (the fill's ceiling)
    wf_ms, wf_gbs = write_floor(ctx, nnz * 12 + (N + 1) * 8)
    out["hamgen"]["write_floor_ms"] = wf_ms
    out["hamgen"]["write_floor_GBs"] = wf_gbs
    out["hamgen"]["fill_GBs"] = (nnz * 12 + (N + 1) * 8) / (last["ham_fill_ms"] * 1e-3) / 1e9
    if sliced and nnz % N == 0:
        # every row of the complete space has the same length
        fl_ms, fl_gbs = spmv_stream_floor(ctx, N, nnz // N)
        out["spmv"]["stream_floor_ms"] = fl_ms
        out["spmv"]["stream_floor_GBs"] = fl_gbs
        out["spmv"]["frac_of_stream_floor"] = fl_ms / avg_spmv_ms
    if not args.no_asci:
        out["mlp"] = mlp_throughput(C, ctx)
        out["asci"] = asci_step_comparison(C, ctx)
    if not args.no_batch:
        out["batch"] = batch_section(C, ctx)
    if not args.no_selected_ci:
        sel = {}
        out["selected_ci"] = selected_ci_section(C, ctx, cpu=not args.no_cpu_baseline, keep=sel)
        if not args.no_c5:
            out["c5"] = c5_section(C, ctx, sel)
            del sel
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(T, V, dets)
        out["cpu_fair"] = cpu_fair_baseline(T, V, dets)
        out["cpu_spmv"] = cpu_spmv_baseline(ctx, T, V, d_dets, N)
        assert out["cpu_spmv"]["sample"].startswith(f"full CAS(12e,12o) SpMV ({nnz} nonzeros")
        out["cpu_spmv"]["gpu_speedup"] = out["cpu_spmv"]["ms_per_spmv"] / avg_spmv_ms
        out["c1"] = c1_section(C, ctx)
        out["davidson"]["E0_oracle"] = out["cpu_spmv"]["oracle_E0"]
        out["davidson"]["abs_dE_vs_oracle"] = abs(E - out["cpu_spmv"]["oracle_E0"])
        out["cpu_rdms"] = cpu_rdm_baseline(dets, d_X.numpy())
        out["cpu_rdms"]["gpu_speedup"] = out["cpu_rdms"]["value"] / out["rdms"]["ms"]
    out["host"] = {"cpu": cpu_model(), "nproc": os.cpu_count(),
                   "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
