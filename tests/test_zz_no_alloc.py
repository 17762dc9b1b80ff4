"""mhip_gemm_tn_noalloc (block partials in the device's static scratch
array, no allocation) against mhip_gemm_tn (partials in a fresh device
buffer): the same kernels, so the products must agree bit for bit, on two
calls in a row with the second smaller, and both must equal A^T B."""
import ctypes

import numpy as np
import pytest

from macis_amd import _capi

from conftest import announce

pytestmark = pytest.mark.gpu


def _call(name, ctx, N, k, dA, m, dB, dC):
    return getattr(_capi.lib(), name)(ctx.h, N, k, dA.ptr, N, m, dB.ptr, N, dC.ptr, k)


def test_gemm_tn_noalloc_matches_gemm_tn(ctx):
    rng = np.random.default_rng(7)
    # the first call uses every partial slot of the scratch (256 blocks),
    # the second far fewer: stale partials of the first must not leak in
    for N, k, m in [(1_000_000, 21, 21), (3000, 7, 5)]:
        announce(f"gemm_tn_noalloc N={N} k={k} m={m}")
        A = rng.standard_normal((k, N))  # column j of the N x k operand = row j here
        B = rng.standard_normal((m, N))
        dA, dB = ctx.to_device(A), ctx.to_device(B)
        dC1 = ctx.to_device(np.full(k * m, np.nan))
        dC2 = ctx.to_device(np.full(k * m, np.nan))
        _capi.check(_call("mhip_gemm_tn", ctx, N, k, dA, m, dB, dC1), "gemm_tn")
        _capi.check(_call("mhip_gemm_tn_noalloc", ctx, N, k, dA, m, dB, dC2), "gemm_tn_noalloc")
        C1 = dC1.numpy().reshape(m, k).T
        C2 = dC2.numpy().reshape(m, k).T
        assert np.array_equal(C1.view(np.uint64), C2.view(np.uint64))
        ref = A @ B.T
        assert np.max(np.abs(C2 - ref)) <= 1e-12 * np.sqrt(N) * 10


def test_gemm_tn_noalloc_rejects_bad_shapes(ctx):
    announce("gemm_tn_noalloc argument checks")
    dA = ctx.to_device(np.zeros(64 * 33))
    dC = ctx.to_device(np.zeros(33 * 33))
    rc = _capi.lib().mhip_gemm_tn_noalloc(ctx.h, 64, 33, dA.ptr, 64, 4, dA.ptr, 64, dC.ptr, 33)
    assert rc == -1
    assert b"[1, 32]" in _capi.lib().mhip_last_error()
