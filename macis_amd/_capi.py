"""ctypes binding of libmacis_hip.so (the C-ABI declared in include/macis_hip.h).

This is the Python-side caller used by the tests and the benchmark; the
C++ mirror of the reference API (macis_amd/include/macis/...) is the other
caller. Importing raises if the HIP library has not been built: there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmacis_hip.so")

MHIP_SUCCESS = 0
MHIP_ERR_NOT_CONVERGED = -5
MHIP_ERR_INDEX_OVERFLOW = -8

_p = C.c_void_p
_pp = C.POINTER(C.c_void_p)
_i64 = C.c_int64
_i32 = C.c_int
_d = C.c_double
_pi64 = C.POINTER(C.c_int64)
_pd = C.POINTER(C.c_double)


class Timings(C.Structure):
    _fields_ = [
        ("ham_bounds_ms", C.c_double),
        ("ham_count_ms", C.c_double),
        ("ham_scan_ms", C.c_double),
        ("ham_fill_ms", C.c_double),
        ("ham_elem_ms", C.c_double),
        ("ham_compact_ms", C.c_double),
        ("ham_total_ms", C.c_double),
        ("ham_nnz_struct", C.c_int64),
        ("ham_nnz", C.c_int64),
        ("dav_spmv_ms", C.c_double),
        ("dav_spmv_calls", C.c_int64),
        ("dav_total_ms", C.c_double),
        ("dav_iters", C.c_int64),
        ("dav_restarts", C.c_int64),
        ("dav_residual", C.c_double),
        ("asci_contrib_ms", C.c_double),
        ("asci_sort_ms", C.c_double),
        ("asci_topk_ms", C.c_double),
        ("asci_total_ms", C.c_double),
        ("asci_npairs", C.c_int64),
        ("asci_nunique", C.c_int64),
        ("asci_nn_ms", C.c_double),
        ("asci_nsurvivors", C.c_int64),
        ("ham_path", C.c_int64),
        ("dav_layout_ms", C.c_double),
        ("dav_sliced", C.c_int64),
        ("asci_nblocks", C.c_int64),
        ("asci_nblocks_kept", C.c_int64),
        ("dav_spmv_matrix_bytes", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ASCISettings(C.Structure):
    """mhip_asci_settings (ASCISettings, determinant_search.hpp:34-51)."""

    _fields_ = [
        ("ncdets_max", C.c_int64),
        ("h_el_tol", C.c_double),
        ("rv_prune_tol", C.c_double),
        ("pair_size_max", C.c_int64),
        ("just_singles", C.c_int32),
    ]

    @classmethod
    def defaults(cls, **kw):
        s = cls(ncdets_max=100, h_el_tol=1e-8, rv_prune_tol=1e-8,
                pair_size_max=500000000, just_singles=0)
        for k, v in kw.items():
            setattr(s, k, v)
        return s


_SIGS = {
    "mhip_get_version": (_i32, []),
    "mhip_error_string": (C.c_char_p, [_i32]),
    "mhip_last_error": (C.c_char_p, []),
    "mhip_context_create": (_i32, [_i32, _pp]),
    "mhip_context_destroy": (_i32, [_p]),
    "mhip_context_get_stream": (_i32, [_p, _pp]),
    "mhip_synchronize": (_i32, [_p]),
    "mhip_get_timings": (_i32, [_p, C.POINTER(Timings)]),
    "mhip_context_set_build_path": (_i32, [_p, _i32]),
    "mhip_context_release_cached": (_i32, [_p]),
    "mhip_malloc": (_i32, [_p, C.c_size_t, _pp]),
    "mhip_free": (_i32, [_p, _p]),
    "mhip_memcpy_h2d": (_i32, [_p, _p, _p, C.c_size_t]),
    "mhip_memcpy_d2h": (_i32, [_p, _p, _p, C.c_size_t]),
    "mhip_memcpy_d2d": (_i32, [_p, _p, _p, C.c_size_t]),
    "mhip_memset": (_i32, [_p, _p, _i32, C.c_size_t]),
    "mhip_set_integrals": (_i32, [_p, _i64, _p, _p]),
    "mhip_ham_build": (_i32, [_p, _i32, _p, _i64, _p, _i64, _d, _pp]),
    "mhip_ham_diagonal": (_i32, [_p, _i32, _p, _i64, _p]),
    "mhip_form_rdms": (_i32, [_p, _i32, _p, _i64, _p, _i64, _p, _p, _p]),
    "mhip_csr_info": (_i32, [_p, _pi64, _pi64, _pi64]),
    "mhip_csr_device_arrays": (_i32, [_p, _pp, _pp, _pp]),
    "mhip_csr_copy_to_host_i32": (_i32, [_p, _p, _p, _p, _p]),
    "mhip_csr_copy_to_host_i64": (_i32, [_p, _p, _p, _p, _p]),
    "mhip_csr_create_from_host": (_i32, [_p, _i64, _i64, _i64, _i64, _p, _p, _p, _pp]),
    "mhip_csr_destroy": (_i32, [_p]),
    "mhip_spmv": (_i32, [_p, _p, _i64, _d, _p, _i64, _d, _p, _i64]),
    "mhip_extract_diagonal": (_i32, [_p, _p, _p]),
    "mhip_csr_prepare_spmv": (_i32, [_p, _p, _i32, _p]),
    "mhip_gemm_tn": (_i32, [_p, _i64, _i32, _p, _i64, _i32, _p, _i64, _p, _i32]),
    "mhip_gemm_tn_noalloc": (_i32, [_p, _i64, _i32, _p, _i64, _i32, _p, _i64, _p, _i32]),
    "mhip_davidson": (_i32, [_p, _p, _i64, _i64, _d, _p, _p, _pd, _pi64]),
    "mhip_selected_ci_diag": (_i32, [_p, _i32, _p, _i64, _d, _i64, _i64, _d, _p, _pd, _pi64, _pp]),
    "mhip_asci_search": (_i32, [_p, _i32, C.POINTER(ASCISettings), _i64, _p, _i64, _d, _p, _i64,
                                _p, _pi64]),
    "mhip_mlp_create": (_i32, [_p, C.c_uint64, _pp]),
    "mhip_mlp_destroy": (_i32, [_p]),
    "mhip_mlp_num_params": (_i64, []),
    "mhip_mlp_get_params": (_i32, [_p, _p, _p]),
    "mhip_mlp_set_params": (_i32, [_p, _p, _p]),
    "mhip_mlp_get_grads": (_i32, [_p, _p, _p]),
    "mhip_mlp_predict": (_i32, [_p, _p, _i32, _p, _i64, _p]),
    "mhip_mlp_train_step": (_i32, [_p, _p, _i32, _p, _p, _i64, C.c_float, C.POINTER(C.c_float)]),
    "mhip_gemm_bf16_nt": (_i32, [_p, _i32, _i32, _i32, _p, _p, _p]),
    "mhip_asci_candidates": (_i32, [_p, _i32, C.POINTER(ASCISettings), _p, _i64, _d, _p, _i64,
                                    _p, _p, _i64, _pi64]),
    "mhip_asci_search_screened": (_i32, [_p, _i32, C.POINTER(ASCISettings), _p, C.c_float,
                                         C.c_float, _d, _i64, _p, _i64, _d, _p, _i64, _p, _pi64]),
    "mhip_asci_search_blocks": (_i32, [_p, _i32, C.POINTER(ASCISettings), _p, _d, _i64, _p, _i64,
                                       _d, _p, _i64, _p, _pi64]),
    "mhip_asci_blocks": (_i32, [_p, _i32, C.POINTER(ASCISettings), _p, _i64, _d, _i64, _p, _p,
                                _i64, _pi64]),
    "mhip_batch_create": (_i32, [_p, _i32, _p, _p, _p, _p, _p, _pp]),
    "mhip_batch_create_w": (_i32, [_p, _i32, _i32, _p, _p, _p, _p, _p, _pp]),
    "mhip_batch_destroy": (_i32, [_p]),
    "mhip_batch_info": (_i32, [_p, _pi64, _pi64]),
    "mhip_batch_build": (_i32, [_p, _p, _d]),
    "mhip_batch_tune": (_i32, [_p, _p, _p]),
    "mhip_batch_spmv_layout": (_i32, [_p, _p, _i32, C.POINTER(C.c_int)]),
    "mhip_batch_davidson": (_i32, [_p, _p, _i64, _i64, _d, _i32, _p, _p, _pd]),
    "mhip_batch_diagonal": (_i32, [_p, _p, _p]),
    "mhip_batch_set_guess": (_i32, [_p, _p, _p]),
    "mhip_batch_vectors": (_i32, [_p, _p, _p]),
    "mhip_batch_instance_csr": (_i32, [_p, _p, _i32, _p, _p, _p, _pi64]),
}

_lib = None


def lib():
    """Load the HIP library; raises if it is missing (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    """Names of the entry points declared in include/macis_hip.h."""
    import re

    hdr = os.path.join(os.path.dirname(_HERE), "include", "macis_hip.h")
    with open(hdr) as f:
        txt = f.read()
    return sorted(set(re.findall(r"\b(mhip_[a-z0-9_]+)\s*\(", txt)))


class MhipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def check(rc, what=""):
    if rc != MHIP_SUCCESS:
        L = lib()
        msg = L.mhip_last_error().decode() or L.mhip_error_string(rc).decode()
        raise MhipError(rc, f"{what}: {msg} (code {rc})")


class DeviceArray:
    """Owning device buffer with a numpy dtype/shape (freed on __del__)."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().mhip_malloc(ctx.h, self.nbytes, C.byref(p)), "mhip_malloc")
        self.ptr = p.value or 0

    def __del__(self):
        if getattr(self, "ptr", 0) and self.ctx.h:
            lib().mhip_free(self.ctx.h, self.ptr)
            self.ptr = 0

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            check(lib().mhip_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes), "d2h")
        return out

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        if self.nbytes:
            check(lib().mhip_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data, self.nbytes), "h2d")
        return self

    def zero(self):
        if self.nbytes:
            check(lib().mhip_memset(self.ctx.h, self.ptr, 0, self.nbytes), "memset")
        return self


class CSR:
    """Device CSR handle (mhip_csr)."""

    def __init__(self, ctx, h):
        self.ctx = ctx
        self.h = h
        m, n, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().mhip_csr_info(h, C.byref(m), C.byref(n), C.byref(nnz)), "csr_info")
        self.m, self.n, self.nnz = m.value, n.value, nnz.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().mhip_csr_destroy(self.h)
            self.h = None

    def to_host(self):
        """(rowptr int64, colind int32, nzval float64) copied to the host."""
        rp = np.empty(self.m + 1, dtype=np.int64)
        ci64 = np.empty(max(self.nnz, 1), dtype=np.int64)
        nz = np.empty(max(self.nnz, 1), dtype=np.float64)
        check(lib().mhip_csr_copy_to_host_i64(self.ctx.h, self.h, rp.ctypes.data,
                                              ci64.ctypes.data, nz.ctypes.data), "copy_to_host")
        return rp, ci64[: self.nnz].astype(np.int32), nz[: self.nnz]

    def to_host_i32(self):
        rp = np.empty(self.m + 1, dtype=np.int32)
        ci = np.empty(max(self.nnz, 1), dtype=np.int32)
        nz = np.empty(max(self.nnz, 1), dtype=np.float64)
        check(lib().mhip_csr_copy_to_host_i32(self.ctx.h, self.h, rp.ctypes.data,
                                              ci.ctypes.data, nz.ctypes.data), "copy_to_host_i32")
        return rp, ci[: self.nnz], nz[: self.nnz]

    def to_host_native(self):
        """(rowptr int64, colind int32, nzval float64): the device arrays
        copied as they are stored (no index widening)."""
        d_rp, d_ci, d_nz = self.device_arrays()
        rp = np.empty(self.m + 1, dtype=np.int64)
        ci = np.empty(max(self.nnz, 1), dtype=np.int32)
        nz = np.empty(max(self.nnz, 1), dtype=np.float64)
        L = lib()
        check(L.mhip_memcpy_d2h(self.ctx.h, rp.ctypes.data, d_rp, rp.nbytes), "copy rowptr")
        if self.nnz:
            check(L.mhip_memcpy_d2h(self.ctx.h, ci.ctypes.data, d_ci, 4 * self.nnz), "copy colind")
            check(L.mhip_memcpy_d2h(self.ctx.h, nz.ctypes.data, d_nz, 8 * self.nnz), "copy nzval")
        return rp, ci[: self.nnz], nz[: self.nnz]

    def device_arrays(self):
        rp, ci, nz = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(lib().mhip_csr_device_arrays(self.h, C.byref(rp), C.byref(ci), C.byref(nz)), "arrays")
        return rp.value, ci.value, nz.value


class Context:
    """mhip_context: one device, one stream."""

    def __init__(self, device=0):
        p = C.c_void_p()
        check(lib().mhip_context_create(device, C.byref(p)), "context_create")
        self.h = p.value

    def close(self):
        if getattr(self, "h", None):
            lib().mhip_context_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- memory
    def empty(self, shape, dtype):
        return DeviceArray(self, shape, dtype)

    def to_device(self, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or np.asarray(a).dtype)
        return DeviceArray(self, a.shape, a.dtype).upload(a)

    def synchronize(self):
        check(lib().mhip_synchronize(self.h), "synchronize")

    def stream(self):
        s = C.c_void_p()
        check(lib().mhip_context_get_stream(self.h, C.byref(s)), "get_stream")
        return s.value

    def timings(self):
        t = Timings()
        check(lib().mhip_get_timings(self.h, C.byref(t)), "get_timings")
        return t.as_dict()

    BUILD_AUTO, BUILD_TILED, BUILD_PRODUCT, BUILD_SPIN_GROUPS = 0, 1, 2, 3

    def set_build_path(self, path):
        check(lib().mhip_context_set_build_path(self.h, path), "set_build_path")

    def release_cached(self):
        check(lib().mhip_context_release_cached(self.h), "release_cached")

    # ---- hot path
    def set_integrals(self, norb, T, V):
        T = np.ascontiguousarray(T, dtype=np.float64)
        V = np.ascontiguousarray(V, dtype=np.float64)
        assert T.size == norb**2 and V.size == norb**4
        check(lib().mhip_set_integrals(self.h, norb, T.ctypes.data, V.ctypes.data), "set_integrals")

    def ham_build(self, W, d_bra, nbra, d_ket=None, nket=None, thresh=1e-16):
        if d_ket is None:
            d_ket, nket = d_bra, nbra
        out = C.c_void_p()
        check(lib().mhip_ham_build(self.h, W, _ptr(d_bra), nbra, _ptr(d_ket), nket, thresh,
                                   C.byref(out)), "ham_build")
        return CSR(self, out.value)

    def ham_diagonal(self, W, d_dets, n, d_out):
        check(lib().mhip_ham_diagonal(self.h, W, _ptr(d_dets), n, _ptr(d_out)), "ham_diagonal")

    def form_rdms(self, W, d_bra, nbra, d_C, d_ordm, d_trdm=None, d_ket=None, nket=None):
        """Adds the 1-RDM (and 2-RDM when d_trdm is given) into device arrays."""
        d_ket = d_bra if d_ket is None else d_ket
        nket = nbra if nket is None else nket
        check(lib().mhip_form_rdms(self.h, W, _ptr(d_bra), nbra, _ptr(d_ket), nket, _ptr(d_C),
                                   _ptr(d_ordm), _ptr(d_trdm) if d_trdm is not None else None),
              "form_rdms")

    def csr_from_host(self, m, n, rowptr, colind, nzval, indexing=0):
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        colind = np.ascontiguousarray(colind, dtype=np.int32)
        nzval = np.ascontiguousarray(nzval, dtype=np.float64)
        out = C.c_void_p()
        nnz = len(nzval)
        check(lib().mhip_csr_create_from_host(self.h, m, n, nnz, indexing, rowptr.ctypes.data,
                                              colind.ctypes.data, nzval.ctypes.data,
                                              C.byref(out)), "csr_create_from_host")
        return CSR(self, out.value)

    def spmv(self, A, K, alpha, d_V, LDV, beta, d_AV, LDAV):
        check(lib().mhip_spmv(self.h, A.h, K, alpha, _ptr(d_V), LDV, beta, _ptr(d_AV), LDAV),
              "spmv")

    def extract_diagonal(self, A, d_D):
        check(lib().mhip_extract_diagonal(self.h, A.h, _ptr(d_D)), "extract_diagonal")

    def prepare_spmv(self, A, always=False, panels=False, xtiles=False):
        """Attach the sliced SpMV copy to A (always: whatever the padding), or
        with panels=True the column-panel split, with xtiles=True the
        panel-major copy with LDS-staged x tiles; True if A has it afterwards."""
        out = C.c_int(0)
        policy = 3 if xtiles else 2 if panels else (1 if always else 0)
        check(lib().mhip_csr_prepare_spmv(self.h, A.h, policy, C.byref(out)),
              "csr_prepare_spmv")
        return bool(out.value)

    def gemm_tn(self, N, k, d_A, lda, m, d_B, ldb, d_C, ldc):
        check(lib().mhip_gemm_tn(self.h, N, k, _ptr(d_A), lda, m, _ptr(d_B), ldb, _ptr(d_C), ldc),
              "gemm_tn")

    def davidson(self, A, max_m, tol, d_X, d_D=None, max_restarts=0):
        E = C.c_double()
        it = C.c_int64()
        check(lib().mhip_davidson(self.h, A.h, max_m, max_restarts, tol,
                                  _ptr(d_D) if d_D is not None else None, _ptr(d_X),
                                  C.byref(E), C.byref(it)), "davidson")
        return it.value, E.value

    def selected_ci_diag(self, W, d_dets, ndets, d_C, h_el_tol=np.finfo(float).eps, max_m=20,
                         res_tol=1e-8, max_restarts=0, keep_H=False):
        E = C.c_double()
        nnz = C.c_int64()
        H = C.c_void_p()
        check(lib().mhip_selected_ci_diag(self.h, W, _ptr(d_dets), ndets, h_el_tol, max_m,
                                          max_restarts, res_tol, _ptr(d_C), C.byref(E),
                                          C.byref(nnz), C.byref(H) if keep_H else None),
              "selected_ci_diag")
        return E.value, nnz.value, (CSR(self, H.value) if keep_H else None)

    def asci_candidates(self, W, settings, d_cdets, ncdets, E_ASCI, d_C, norb, capacity):
        """(dets (n, W) DeviceArray, |rv| DeviceArray, n) of the scored candidates."""
        d_out = self.empty((max(capacity, 1), W), np.uint64)
        a_out = self.empty(max(capacity, 1), np.float64)
        n = C.c_int64()
        check(lib().mhip_asci_candidates(self.h, W, C.byref(settings), _ptr(d_cdets), ncdets,
                                         E_ASCI, _ptr(d_C), norb, d_out.ptr, a_out.ptr, capacity,
                                         C.byref(n)), "asci_candidates")
        return d_out, a_out, n.value

    def asci_blocks(self, W, settings, d_cdets, ncdets, E_ASCI, norb, targets=True):
        """(block dets (n, W) DeviceArray, targets DeviceArray or None, n): the
        excitation blocks of the core (mhip_asci_blocks)."""
        n = C.c_int64()
        check(lib().mhip_asci_blocks(self.h, W, C.byref(settings), _ptr(d_cdets), ncdets, E_ASCI,
                                     norb, None, None, 0, C.byref(n)), "asci_blocks")
        nb = n.value
        d_out = self.empty((max(nb, 1), W), np.uint64)
        t_out = self.empty(max(nb, 1), np.float32) if targets else None
        check(lib().mhip_asci_blocks(self.h, W, C.byref(settings), _ptr(d_cdets), ncdets, E_ASCI,
                                     norb, d_out.ptr, t_out.ptr if targets else None, nb,
                                     C.byref(n)), "asci_blocks")
        return d_out, t_out, nb

    def asci_search_blocks(self, W, settings, d_scores, keep_fraction, ndets_max, d_cdets,
                           ncdets, E_ASCI, d_C, norb, d_out):
        n = C.c_int64()
        check(lib().mhip_asci_search_blocks(self.h, W, C.byref(settings), _ptr(d_scores),
                                            keep_fraction, ndets_max, _ptr(d_cdets), ncdets,
                                            E_ASCI, _ptr(d_C), norb, _ptr(d_out), C.byref(n)),
              "asci_search_blocks")
        return n.value

    def asci_search_screened(self, W, settings, mlp, keep_fraction, ndets_max, d_cdets, ncdets,
                             E_ASCI, d_C, norb, d_out, scale=1.0, shift=0.0):
        n = C.c_int64()
        check(lib().mhip_asci_search_screened(self.h, W, C.byref(settings), mlp.h, scale, shift,
                                              keep_fraction, ndets_max, _ptr(d_cdets), ncdets,
                                              E_ASCI, _ptr(d_C), norb, _ptr(d_out), C.byref(n)),
              "asci_search_screened")
        return n.value

    def asci_search(self, W, settings, ndets_max, d_cdets, ncdets, E_ASCI, d_C, norb, d_out):
        n = C.c_int64()
        check(lib().mhip_asci_search(self.h, W, C.byref(settings), ndets_max, _ptr(d_cdets),
                                     ncdets, E_ASCI, _ptr(d_C), norb, _ptr(d_out), C.byref(n)),
              "asci_search")
        return n.value


class Batch:
    """Batched CAS-CI instances (mhip_batch)."""

    def __init__(self, ctx, norbs, det_lists, Ts, Vs, W=1):
        """det_lists: per instance (ndets,) uint64 words (W=1) or (ndets, 2)
        [alpha, beta] words (W=2)."""
        self.ctx = ctx
        n = len(norbs)
        norb = np.ascontiguousarray(norbs, dtype=np.int32)
        offs = np.zeros(n + 1, dtype=np.int64)
        offs[1:] = np.cumsum([np.asarray(d).size // W for d in det_lists])
        dets = np.ascontiguousarray(np.concatenate([np.asarray(d).reshape(-1) for d in det_lists]),
                                    dtype=np.uint64)
        self._T = [np.ascontiguousarray(t, dtype=np.float64) for t in Ts]
        self._V = [np.ascontiguousarray(v, dtype=np.float64) for v in Vs]
        Tp = (C.c_void_p * n)(*[t.ctypes.data for t in self._T])
        Vp = (C.c_void_p * n)(*[v.ctypes.data for v in self._V])
        p = C.c_void_p()
        check(lib().mhip_batch_create_w(ctx.h, W, n, norb.ctypes.data, offs.ctypes.data,
                                        dets.ctypes.data, Tp, Vp, C.byref(p)), "batch_create")
        self.h = p.value
        self.ninst = n
        self.offsets = offs
        self.ndets = int(offs[-1])

    def __del__(self):
        if getattr(self, "h", None):
            lib().mhip_batch_destroy(self.h)
            self.h = None

    def info(self):
        a, b = C.c_int64(), C.c_int64()
        check(lib().mhip_batch_info(self.h, C.byref(a), C.byref(b)), "batch_info")
        return a.value, b.value

    def build(self, thresh):
        check(lib().mhip_batch_build(self.ctx.h, self.h, thresh), "batch_build")

    def tune(self):
        lanes = np.zeros(self.ninst, np.int32)
        check(lib().mhip_batch_tune(self.ctx.h, self.h, lanes.ctypes.data), "batch_tune")
        return lanes

    def spmv_layout(self, layout=-2):
        """Forces (1 sliced, 0 CSR schedule, -1 default) or, with -2, only
        queries the SpMV layout; returns the current one."""
        cur = C.c_int()
        check(lib().mhip_batch_spmv_layout(self.ctx.h, self.h, layout, C.byref(cur)),
              "batch_spmv_layout")
        return cur.value

    def diagonal(self):
        """H_ii of all instances, caller order (concatenated)."""
        D = np.zeros(self.ndets)
        check(lib().mhip_batch_diagonal(self.ctx.h, self.h, D.ctypes.data), "batch_diagonal")
        return D

    def set_guess(self, X):
        """Davidson guess (concatenated, caller order); None: diagonal guess."""
        if X is None:
            check(lib().mhip_batch_set_guess(self.ctx.h, self.h, None), "batch_set_guess")
            return
        X = np.ascontiguousarray(X, dtype=np.float64)
        assert X.size == self.ndets
        check(lib().mhip_batch_set_guess(self.ctx.h, self.h, X.ctypes.data), "batch_set_guess")

    def vectors(self):
        X = np.zeros(self.ndets)
        check(lib().mhip_batch_vectors(self.ctx.h, self.h, X.ctypes.data), "batch_vectors")
        return X

    def instance_csr(self, inst):
        """(rowptr int64, colind int32, nzval) of one instance's block."""
        nnz = C.c_int64()
        check(lib().mhip_batch_instance_csr(self.ctx.h, self.h, inst, None, None, None,
                                            C.byref(nnz)), "batch_instance_csr")
        m = int(self.offsets[inst + 1] - self.offsets[inst])
        rp = np.zeros(m + 1, np.int64)
        ci = np.zeros(max(nnz.value, 1), np.int32)
        nz = np.zeros(max(nnz.value, 1), np.float64)
        check(lib().mhip_batch_instance_csr(self.ctx.h, self.h, inst, rp.ctypes.data,
                                            ci.ctypes.data, nz.ctypes.data, C.byref(nnz)),
              "batch_instance_csr")
        return rp, ci[: nnz.value], nz[: nnz.value]

    def davidson(self, max_m=20, max_restarts=0, tol=1e-8, use_graphs=True):
        E = np.zeros(self.ninst)
        res = np.zeros(self.ninst)
        ms = C.c_double()
        check(lib().mhip_batch_davidson(self.ctx.h, self.h, max_m, max_restarts, tol,
                                        int(use_graphs), E.ctypes.data, res.ctypes.data,
                                        C.byref(ms)), "batch_davidson")
        return E, res, ms.value


class MLP:
    """Learned pre-screen network (mhip_mlp)."""

    def __init__(self, ctx, seed=0):
        self.ctx = ctx
        p = C.c_void_p()
        check(lib().mhip_mlp_create(ctx.h, seed, C.byref(p)), "mlp_create")
        self.h = p.value
        self.nparam = lib().mhip_mlp_num_params()

    def __del__(self):
        if getattr(self, "h", None):
            lib().mhip_mlp_destroy(self.h)
            self.h = None

    def params(self):
        out = np.empty(self.nparam, np.float32)
        check(lib().mhip_mlp_get_params(self.ctx.h, self.h, out.ctypes.data), "get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().mhip_mlp_set_params(self.ctx.h, self.h, p.ctypes.data), "set_params")

    def grads(self):
        out = np.empty(self.nparam, np.float32)
        check(lib().mhip_mlp_get_grads(self.ctx.h, self.h, out.ctypes.data), "get_grads")
        return out

    def predict(self, W, d_dets, n, d_out):
        check(lib().mhip_mlp_predict(self.ctx.h, self.h, W, _ptr(d_dets), n, _ptr(d_out)), "predict")

    def train_step(self, W, d_dets, d_target, n, lr, want_loss=True):
        """One Adam step; returns the mini-batch loss (None when not asked
        for: the step then does not wait for the device)."""
        loss = C.c_float()
        check(lib().mhip_mlp_train_step(self.ctx.h, self.h, W, _ptr(d_dets), _ptr(d_target), n,
                                        lr, C.byref(loss) if want_loss else None), "train_step")
        return loss.value if want_loss else None


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, DeviceArray):
        return x.ptr
    return int(x)
