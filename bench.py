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
N_ORB, N_ALPHA, N_BETA, SEED = 12, 6, 6, 12


def spmv_bytes(nnz, n):
    """Algorithmic bytes of one CSR SpMV: values + int32 columns + int64 row
    pointers + x read + y write (y is read when beta != 0; Davidson passes 0)."""
    return nnz * (8 + 4) + (n + 1) * 8 + 2 * n * 8


def cpu_baseline(T, V, dets, budget_s=12.0):
    """Oracle restatement of the reference's serial double loop
    (double_loop.hpp:45-91) on a contiguous sample of bra rows x all kets."""
    from oracle import oracle as O

    orc = O.Ham(T, V)
    eps = np.finfo(float).eps
    rows, nnz, t_used = 0, 0, 0.0
    start = len(dets) // 3
    chunk = 64
    while t_used < budget_s and rows < 20000:
        t0 = time.perf_counter()
        cnt, _, _ = orc.csr_rows(1, dets, start + rows, start + rows + chunk, thresh=eps)
        t_used += time.perf_counter() - t0
        rows += chunk
        nnz += int(cnt.sum())
    return {
        "value": nnz / t_used,
        "unit": "nnz/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{rows} contiguous bra rows x {len(dets)} kets of CAS(12e,12o), "
                  f"serial double loop (the reference's loop has no threading), "
                  f"{t_used:.1f} s; nnz/s extrapolates to the full build",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--max-restarts", type=int, default=200)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="one 1 GiB device-to-device copy first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the hot path has no "
              "multi-GPU launch); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from macis_amd import capi
    from macis_amd import dets as D
    from macis_amd.synthetic import seeded_integrals

    C = capi()
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
    bytes_spmv = spmv_bytes(nnz, N)
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
        },
        "davidson": {
            "ms_per_solve": dav_ms / args.steps,
            "iterations": last["dav_iters"],
            "restarts": last["dav_restarts"],
            "residual": last["dav_residual"],
            "E0": E,
        },
        "roofline": {
            "kernel": "spmv_kernel (FP64 CSR SpMV, wavefront per row)",
            "bound": "hbm",
            "achieved": spmv_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": spmv_gbs / HBM_PEAK_GBS,
            "traffic": None,
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(T, V, dets)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
