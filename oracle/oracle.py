"""ctypes wrapper of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module. It restates the reference algorithms in C (oracle.c, each
function citing the dfablan/MACIS file:line it follows) and is the checker
for the HIP path, never part of it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

ORC_ERR_NOT_CONVERGED = -4
ORC_ERR_DID_NOT_GROW = -5
ORC_ERR_REFINE = -6

_p = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int
_d = C.c_double
_u32 = C.c_uint32


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "orc_ham_create": (_p, [_i64, _p, _p]),
            "orc_ham_destroy": (None, [_p]),
            "orc_ham_array": (_p, [_p, _i32]),
            "orc_matrix_element": (_d, [_p, _i32, _p, _p]),
            "orc_single_orbital_ens": (None, [_p, _i64, _p, _i64, _p, _i64, _p]),
            "orc_fast_diag_single": (_d, [_p, _d, _d, _u32, _u32, _d]),
            "orc_fast_diag_ss_double": (_d, [_p, _d, _d, _d, _d, _u32, _u32, _u32, _u32, _d]),
            "orc_fast_diag_os_double": (_d, [_p, _d, _d, _d, _d, _u32, _u32, _u32, _u32, _d]),
            "orc_single_excitation_sign": (_d, [C.c_uint64, _u32, _u32]),
            "orc_doubles_sign_indices": (_d, [C.c_uint64, C.c_uint64, C.c_uint64, _p, _p, _p, _p]),
            "orc_generate_hilbert_space": (_i64, [_i32, _i64, _i64, _i64, _p]),
            "orc_generate_cisd_hilbert_space": (_i64, [_i32, _i64, _p, _p]),
            "orc_csr_build": (_i32, [_p, _i32, _p, _i64, _p, _i64, _d, _p, _p, _p, _p]),
            "orc_csr_rows": (_i32, [_p, _i32, _p, _i64, _i64, _p, _i64, _d, _p, _p, _p]),
            "orc_csr_rows_parallel": (_i32, [_p, _i32, _p, _i64, _i64, _p, _i64, _d, _p, _p, _p]),
            "orc_num_threads": (_i32, []),
            "orc_csr_verify_rows": (_i64, [_p, _i32, _p, _i64, _i64, _p, _i64, _d, _p, _p, _p,
                                           _p]),
            "orc_form_rdms": (None, [_i64, _i32, _p, _i64, _p, _i64, _p, _p, _p]),
            "orc_free": (None, [_p]),
            "orc_gespmbv": (None, [_i64, _d, _i64, _p, _p, _p, _i64, _p, _i64, _d, _p, _i64]),
            "orc_extract_diagonal": (None, [_i64, _p, _p, _p, _i64, _p]),
            "orc_syev_lower": (None, [_i64, _p, _i64, _p]),
            "orc_diagonal_guess": (None, [_i64, _p, _p]),
            "orc_davidson": (_i32, [_i64, _i64, _p, _p, _p, _p, _d, _p, _p, _p]),
            "orc_selected_ci_diag": (_i32, [_p, _i32, _p, _i64, _d, _i64, _d, _p, _p, _p]),
            "orc_davidson_restarted": (_i32, [_i64, _i64, _i64, _p, _p, _p, _p, _d, _p, _p, _p, _p]),
            "orc_selected_ci_diag_restarted": (_i32, [_p, _i32, _p, _i64, _d, _i64, _i64, _d, _p, _p, _p]),
            "orc_asci_default_settings": (None, [_p]),
            "orc_asci_contributions": (_i64, [_p, _i32, _p, _p, _i64, _d, _p, _i64, _p, _p]),
            "orc_asci_search": (_i32, [_p, _i32, _p, _i64, _p, _i64, _d, _p, _i64, _p, _p, _p]),
            "orc_asci_blocks": (_i64, [_p, _i32, _p, _p, _i64, _d, _i64, _p, _p]),
            "orc_asci_search_blocks": (_i32, [_p, _i32, _p, _p, _d, _i64, _p, _i64, _d, _p, _i64,
                                              _p, _p, _p]),
            "orc_asci_iter": (_i32, [_p, _i32, _p, _i64, _d, _d, _i64, _d, _p, _p, _p, _i64, _p]),
            "orc_asci_grow": (_i32, [_p, _i32, _p, _i64, _d, _d, _p, _p, _p, _p, _i64]),
            "orc_asci_refine": (_i32, [_p, _i32, _p, _i64, _d, _d, _p, _p, _p, _p, _i64]),
            "orc_asci_grow_rot": (_i32, [_p, _i32, _p, _i64, _d, _d, _p, _p, _p, _p, _i64, _i64]),
            "orc_natural_orbital_rotation": (_i32, [_p, _i32, _p, _i64, _p, _p]),
            "orc_two_index_transform": (None, [_i64, _p, _p, _p]),
            "orc_four_index_transform": (None, [_i64, _p, _p, _p]),
            "orc_mlp_num_params": (_i64, []),
            "orc_mlp_loss_grad": (_d, [_p, _i32, _p, _p, _i64, _i32, _p, _p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class ASCISettings(C.Structure):
    """ASCISettings (determinant_search.hpp:34-51) -- non-MPI fields."""

    _fields_ = [
        ("ntdets_max", C.c_int64),
        ("ntdets_min", C.c_int64),
        ("ncdets_max", C.c_int64),
        ("h_el_tol", C.c_double),
        ("rv_prune_tol", C.c_double),
        ("pair_size_max", C.c_int64),
        ("just_singles", C.c_int32),
        ("grow_factor", C.c_int64),
        ("max_refine_iter", C.c_int64),
        ("refine_energy_tol", C.c_double),
    ]

    @classmethod
    def defaults(cls, **kw):
        s = cls()
        lib().orc_asci_default_settings(C.byref(s))
        for k, v in kw.items():
            setattr(s, k, v)
        return s


class Ham:
    """Oracle HamiltonianGenerator: owns copies of T, V and intermediates."""

    def __init__(self, T, V):
        T = np.ascontiguousarray(T, dtype=np.float64)
        V = np.ascontiguousarray(V, dtype=np.float64)
        self.norb = int(round(T.size ** 0.5))
        assert T.size == self.norb**2 and V.size == self.norb**4
        # T/V are column-major in the reference; callers pass flat arrays in
        # that order (T.ravel(order="F")).
        self._h = lib().orc_ham_create(self.norb, _ptr(T.ravel()), _ptr(V.ravel()))
        if not self._h:
            raise MemoryError("orc_ham_create failed")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_ham_destroy(h)
            self._h = None

    def array(self, which):
        n = self.norb
        size = {0: n**2, 1: n**4, 2: n**4, 3: n**3, 4: n**3, 5: n**2, 6: n**2}[which]
        p = lib().orc_ham_array(self._h, which)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), shape=(size,)).copy()

    def matrix_element(self, W, bra, ket):
        b = np.ascontiguousarray(bra, dtype=np.uint64).reshape(-1)
        k = np.ascontiguousarray(ket, dtype=np.uint64).reshape(-1)
        return lib().orc_matrix_element(self._h, W, _ptr(b), _ptr(k))

    def single_orbital_ens(self, ss_occ, os_occ):
        ss = np.ascontiguousarray(ss_occ, dtype=np.uint32)
        os_ = np.ascontiguousarray(os_occ, dtype=np.uint32)
        out = np.zeros(self.norb)
        lib().orc_single_orbital_ens(self._h, self.norb, _ptr(ss), len(ss), _ptr(os_), len(os_), _ptr(out))
        return out

    def fast_diag_single(self, *a):
        return lib().orc_fast_diag_single(self._h, *a)

    def fast_diag_ss_double(self, *a):
        return lib().orc_fast_diag_ss_double(self._h, *a)

    def fast_diag_os_double(self, *a):
        return lib().orc_fast_diag_os_double(self._h, *a)

    def csr(self, W, bra, ket=None, thresh=1e-16):
        """make_csr_hamiltonian_block -> (rowptr int64, colind int32, nzval)."""
        bra = np.ascontiguousarray(bra, dtype=np.uint64)
        ket = bra if ket is None else np.ascontiguousarray(ket, dtype=np.uint64)
        nb, nk = bra.size // W, ket.size // W
        rp, ci, nz, nnz = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_int64()
        rc = lib().orc_csr_build(self._h, W, _ptr(bra), nb, _ptr(ket), nk, thresh,
                                 C.byref(rp), C.byref(ci), C.byref(nz), C.byref(nnz))
        if rc:
            raise RuntimeError(f"orc_csr_build rc={rc}")
        n = nnz.value
        rowptr = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_int64)), shape=(nb + 1,)).copy()
        colind = np.ctypeslib.as_array(C.cast(ci, C.POINTER(C.c_int32)), shape=(max(n, 1),))[:n].copy()
        nzval = np.ctypeslib.as_array(C.cast(nz, C.POINTER(C.c_double)), shape=(max(n, 1),))[:n].copy()
        for p in (rp, ci, nz):
            lib().orc_free(p)
        return rowptr, colind, nzval

    def csr_rows_into(self, W, bra, row_begin, row_end, col, val, ket=None, thresh=1e-16):
        """One pass of the serial double loop over rows [row_begin,row_end)
        writing into caller buffers (large enough for every kept entry, as
        the reference's push_back never reallocates more than amortised);
        returns the row counts."""
        bra = np.ascontiguousarray(bra, dtype=np.uint64)
        ket = bra if ket is None else np.ascontiguousarray(ket, dtype=np.uint64)
        counts = np.zeros(row_end - row_begin, dtype=np.int64)
        lib().orc_csr_rows(self._h, W, _ptr(bra), row_begin, row_end, _ptr(ket), ket.size // W,
                           thresh, _ptr(counts), _ptr(col), _ptr(val))
        return counts

    def csr_rows_parallel(self, W, bra, row_begin, row_end, ket=None, thresh=1e-16, col=None,
                          val=None):
        """Rows [row_begin,row_end) by the OpenMP row-parallel double loop
        (count pass, prefix sum, fill pass): (counts, colind, nzval); with
        caller buffers col/val (large enough) one call fills them and the
        counts are returned."""
        bra = np.ascontiguousarray(bra, dtype=np.uint64)
        ket = bra if ket is None else np.ascontiguousarray(ket, dtype=np.uint64)
        counts = np.zeros(row_end - row_begin, dtype=np.int64)
        if col is not None:
            lib().orc_csr_rows_parallel(self._h, W, _ptr(bra), row_begin, row_end, _ptr(ket),
                                        ket.size // W, thresh, _ptr(counts), _ptr(col), _ptr(val))
            return counts
        lib().orc_csr_rows_parallel(self._h, W, _ptr(bra), row_begin, row_end, _ptr(ket),
                                    ket.size // W, thresh, _ptr(counts), None, None)
        tot = int(counts.sum())
        col = np.zeros(max(tot, 1), dtype=np.int32)
        val = np.zeros(max(tot, 1), dtype=np.float64)
        lib().orc_csr_rows_parallel(self._h, W, _ptr(bra), row_begin, row_end, _ptr(ket),
                                    ket.size // W, thresh, _ptr(counts), _ptr(col), _ptr(val))
        return counts, col[:tot], val[:tot]

    def csr_rows(self, W, bra, row_begin, row_end, ket=None, thresh=1e-16):
        """Rows [row_begin,row_end) of the block: (counts, colind, nzval)."""
        bra = np.ascontiguousarray(bra, dtype=np.uint64)
        ket = bra if ket is None else np.ascontiguousarray(ket, dtype=np.uint64)
        nk = ket.size // W
        counts = np.zeros(row_end - row_begin, dtype=np.int64)
        lib().orc_csr_rows(self._h, W, _ptr(bra), row_begin, row_end, _ptr(ket), nk,
                           thresh, _ptr(counts), None, None)
        tot = int(counts.sum())
        col = np.zeros(max(tot, 1), dtype=np.int32)
        val = np.zeros(max(tot, 1), dtype=np.float64)
        lib().orc_csr_rows(self._h, W, _ptr(bra), row_begin, row_end, _ptr(ket), nk,
                           thresh, _ptr(counts), _ptr(col), _ptr(val))
        return counts, col[:tot], val[:tot]

    def csr_verify_rows(self, W, bra, row_begin, row_end, rowptr, colind, nzval, ket=None,
                        thresh=1e-16):
        """Rows [row_begin,row_end) of a given CSR (rowptr int64 indexed by
        absolute row, colind int32, nzval) against the double loop, rows over
        OpenMP threads: (number of differing rows, first such row or -1)."""
        bra = np.ascontiguousarray(bra, dtype=np.uint64)
        ket = bra if ket is None else np.ascontiguousarray(ket, dtype=np.uint64)
        assert rowptr.dtype == np.int64 and colind.dtype == np.int32 and nzval.dtype == np.float64
        assert len(rowptr) >= row_end + 1
        first = C.c_int64()
        bad = lib().orc_csr_verify_rows(self._h, W, _ptr(bra), row_begin, row_end, _ptr(ket),
                                        ket.size // W, thresh, _ptr(rowptr), _ptr(colind),
                                        _ptr(nzval), C.byref(first))
        if bad < 0:
            raise RuntimeError("orc_csr_verify_rows: bad arguments")
        return int(bad), int(first.value)

    def selected_ci_diag(self, W, dets, h_el_tol=np.finfo(float).eps, max_m=20,
                         res_tol=1e-8, C0=None, max_restarts=0):
        dets = np.ascontiguousarray(dets, dtype=np.uint64)
        n = dets.size // W
        X = np.zeros(n) if C0 is None else np.array(C0, dtype=np.float64)
        E = C.c_double()
        nnz = C.c_int64()
        rc = lib().orc_selected_ci_diag_restarted(self._h, W, _ptr(dets), n, h_el_tol, max_m,
                                                  max_restarts, res_tol, _ptr(X), C.byref(E),
                                                  C.byref(nnz))
        if rc == ORC_ERR_NOT_CONVERGED:
            raise RuntimeError("Davidson Did Not Converge!")
        if rc:
            raise RuntimeError(f"orc_selected_ci_diag rc={rc}")
        return E.value, X, nnz.value

    def asci_contributions(self, W, settings, cdets, E0, C0):
        cdets = np.ascontiguousarray(cdets, dtype=np.uint64)
        C0 = np.ascontiguousarray(C0, dtype=np.float64)
        nc = cdets.size // W
        n = lib().orc_asci_contributions(self._h, W, C.byref(settings), _ptr(cdets), nc, E0,
                                         _ptr(C0), self.norb, None, None)
        states = np.zeros(max(n, 1) * W, dtype=np.uint64)
        rv = np.zeros(max(n, 1))
        lib().orc_asci_contributions(self._h, W, C.byref(settings), _ptr(cdets), nc, E0,
                                     _ptr(C0), self.norb, _ptr(states), _ptr(rv))
        return states[: n * W].reshape(n, W), rv[:n]

    def asci_search(self, W, settings, ndets_max, cdets, E0, C0):
        cdets = np.ascontiguousarray(cdets, dtype=np.uint64)
        C0 = np.ascontiguousarray(C0, dtype=np.float64)
        nc = cdets.size // W
        out = np.zeros(ndets_max * W, dtype=np.uint64)
        rv = np.zeros(ndets_max)
        nn = C.c_int64()
        rc = lib().orc_asci_search(self._h, W, C.byref(settings), ndets_max, _ptr(cdets), nc,
                                   E0, _ptr(C0), self.norb, _ptr(out), C.byref(nn), _ptr(rv))
        if rc:
            raise RuntimeError(f"orc_asci_search rc={rc}")
        n = nn.value
        return out[: n * W].reshape(n, W), rv[: n - nc]

    def asci_blocks(self, W, settings, cdets, E0):
        """(block determinants [nblk, W], targets [nblk]) of the core."""
        cdets = np.ascontiguousarray(cdets, dtype=np.uint64)
        nc = cdets.size // W
        n = lib().orc_asci_blocks(self._h, W, C.byref(settings), _ptr(cdets), nc, E0, self.norb,
                                  None, None)
        bd = np.zeros(max(n, 1) * W, dtype=np.uint64)
        t = np.zeros(max(n, 1))
        lib().orc_asci_blocks(self._h, W, C.byref(settings), _ptr(cdets), nc, E0, self.norb,
                              _ptr(bd), _ptr(t))
        return bd[: n * W].reshape(n, W), t[:n]

    def asci_search_blocks(self, W, settings, scores, keep_fraction, ndets_max, cdets, E0, C0):
        cdets = np.ascontiguousarray(cdets, dtype=np.uint64)
        C0 = np.ascontiguousarray(C0, dtype=np.float64)
        sc = np.ascontiguousarray(scores, dtype=np.float32)
        nc = cdets.size // W
        out = np.zeros(ndets_max * W, dtype=np.uint64)
        rv = np.zeros(ndets_max)
        nn = C.c_int64()
        rc = lib().orc_asci_search_blocks(self._h, W, C.byref(settings), _ptr(sc), keep_fraction,
                                          ndets_max, _ptr(cdets), nc, E0, _ptr(C0), self.norb,
                                          _ptr(out), C.byref(nn), _ptr(rv))
        if rc:
            raise RuntimeError(f"orc_asci_search_blocks rc={rc}")
        n = nn.value
        return out[: n * W].reshape(n, W), rv[: n - nc]

    def asci_iter(self, W, settings, ndets_max, E0, wfn, X, max_sub=20, res_tol=1e-8,
                  matel_tol=np.finfo(float).eps):
        cap = max(ndets_max, len(X))
        w = np.zeros(cap * W, dtype=np.uint64)
        w[: np.size(wfn)] = np.asarray(wfn, dtype=np.uint64).ravel()
        x = np.zeros(cap)
        x[: len(X)] = X
        nw = C.c_int64(len(X))
        E = C.c_double()
        rc = lib().orc_asci_iter(self._h, W, C.byref(settings), max_sub, res_tol, matel_tol,
                                 ndets_max, E0, _ptr(w), _ptr(x), C.byref(nw), self.norb,
                                 C.byref(E))
        if rc:
            raise RuntimeError(f"orc_asci_iter rc={rc}")
        n = nw.value
        return E.value, w[: n * W].reshape(n, W), x[:n]

    def _grow_or_refine(self, fn, W, settings, E0, wfn, X, max_sub, res_tol, matel_tol):
        cap = max(int(settings.ntdets_max), len(X))
        w = np.zeros(cap * W, dtype=np.uint64)
        w[: np.size(wfn)] = np.asarray(wfn, dtype=np.uint64).ravel()
        x = np.zeros(cap)
        x[: len(X)] = X
        nw = C.c_int64(len(X))
        E = C.c_double(E0)
        rc = fn(self._h, W, C.byref(settings), max_sub, res_tol, matel_tol, C.byref(E),
                _ptr(w), _ptr(x), C.byref(nw), self.norb)
        if rc:
            raise RuntimeError(f"asci rc={rc}")
        n = nw.value
        return E.value, w[: n * W].reshape(n, W), x[:n]

    def asci_grow(self, W, settings, E0, wfn, X, max_sub=20, res_tol=1e-8,
                  matel_tol=np.finfo(float).eps):
        return self._grow_or_refine(lib().orc_asci_grow, W, settings, E0, wfn, X,
                                    max_sub, res_tol, matel_tol)

    def asci_refine(self, W, settings, E0, wfn, X, max_sub=20, res_tol=1e-8,
                    matel_tol=np.finfo(float).eps):
        return self._grow_or_refine(lib().orc_asci_refine, W, settings, E0, wfn, X,
                                    max_sub, res_tol, matel_tol)

    def asci_grow_rot(self, W, settings, E0, wfn, X, rot_size_start=1000, max_sub=20,
                      res_tol=1e-8, matel_tol=np.finfo(float).eps):
        """asci_grow with grow_with_rot (grow.hpp:16-173); rotates this Ham."""
        def fn(*a):
            return lib().orc_asci_grow_rot(*a, rot_size_start)
        return self._grow_or_refine(fn, W, settings, E0, wfn, X, max_sub, res_tol, matel_tol)

    def natural_orbital_rotation(self, W, wfn, X):
        """Rotate this Ham to the natural orbitals of (wfn, X); occupations."""
        w = np.ascontiguousarray(wfn, dtype=np.uint64)
        x = np.ascontiguousarray(X, dtype=np.float64)
        occ = np.zeros(self.norb)
        rc = lib().orc_natural_orbital_rotation(self._h, W, _ptr(w), w.size // W, _ptr(x), _ptr(occ))
        if rc:
            raise RuntimeError(f"rotation rc={rc}")
        return occ


def two_index_transform(X, U):
    """Y(p,q) = U(i,p) X(i,j) U(j,q), flat column-major (transform.cxx:22-49)."""
    n = U.shape[0]
    x = np.ascontiguousarray(X, dtype=np.float64).reshape(-1)
    u = np.asfortranarray(U, dtype=np.float64).reshape(-1, order="F")
    y = np.empty(n * n)
    lib().orc_two_index_transform(n, _ptr(x), _ptr(u), _ptr(y))
    return y


def four_index_transform(X, U):
    """Y(p,q,r,s) = X(i,j,k,l) U(i,p) U(j,q) U(k,r) U(l,s) (transform.cxx:51-115)."""
    n = U.shape[0]
    x = np.ascontiguousarray(X, dtype=np.float64).reshape(-1)
    u = np.asfortranarray(U, dtype=np.float64).reshape(-1, order="F")
    y = np.empty(n ** 4)
    lib().orc_four_index_transform(n, _ptr(x), _ptr(u), _ptr(y))
    return y


def form_rdms(norb, W, bra, C, ket=None, want_trdm=True, ordm=None, trdm=None):
    """form_rdms (double_loop.hpp:116-165): spin-summed 1-RDM (norb, norb)
    and 2-RDM (norb,)*4, both column-major (Fortran-ordered views), added
    into ordm / trdm when given."""
    bra = np.ascontiguousarray(bra, dtype=np.uint64)
    ket = bra if ket is None else np.ascontiguousarray(ket, dtype=np.uint64)
    C = np.ascontiguousarray(C, dtype=np.float64)
    n = int(norb)
    o = np.zeros(n * n) if ordm is None else np.ascontiguousarray(ordm).reshape(-1, order="F").copy()
    t = None
    if want_trdm:
        t = np.zeros(n ** 4) if trdm is None else np.ascontiguousarray(trdm).reshape(-1, order="F").copy()
    lib().orc_form_rdms(n, W, _ptr(bra), bra.size // W, _ptr(ket), ket.size // W, _ptr(C), _ptr(o),
                        _ptr(t) if t is not None else None)
    return o.reshape((n, n), order="F"), (t.reshape((n,) * 4, order="F") if t is not None else None)


def single_excitation_sign(state, p, q):
    return lib().orc_single_excitation_sign(int(state), p, q)


def doubles_sign_indices(bra, ket, ex):
    o = [C.c_uint32() for _ in range(4)]
    s = lib().orc_doubles_sign_indices(int(bra), int(ket), int(ex), *[C.byref(x) for x in o])
    return tuple(x.value for x in o) + (s,)


def hilbert_space(W, norb, na, nb):
    n = lib().orc_generate_hilbert_space(W, norb, na, nb, None)
    out = np.zeros(n * W, dtype=np.uint64)
    lib().orc_generate_hilbert_space(W, norb, na, nb, _ptr(out))
    return out.reshape(n, W)


def cisd_space(W, norb, state):
    st = np.ascontiguousarray(state, dtype=np.uint64).reshape(-1)
    n = lib().orc_generate_cisd_hilbert_space(W, norb, _ptr(st), None)
    out = np.zeros(n * W, dtype=np.uint64)
    lib().orc_generate_cisd_hilbert_space(W, norb, _ptr(st), _ptr(out))
    return out.reshape(n, W)


def num_threads():
    """Threads of the oracle's OpenMP regions (gespmbv, csr_rows_parallel)."""
    return int(lib().orc_num_threads())


def gespmbv(K, alpha, m, rowptr, colind, nzval, indexing, V, LDV, beta, AV, LDAV):
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    colind = np.ascontiguousarray(colind, dtype=np.int32)
    nzval = np.ascontiguousarray(nzval, dtype=np.float64)
    V = np.ascontiguousarray(V, dtype=np.float64)
    assert AV.dtype == np.float64 and AV.flags.c_contiguous
    lib().orc_gespmbv(K, alpha, m, _ptr(rowptr), _ptr(colind), _ptr(nzval), indexing,
                      _ptr(V), LDV, beta, _ptr(AV), LDAV)
    return AV


def extract_diagonal(m, rowptr, colind, nzval, indexing=0):
    D = np.zeros(m)
    lib().orc_extract_diagonal(m, _ptr(np.ascontiguousarray(rowptr, dtype=np.int64)),
                               _ptr(np.ascontiguousarray(colind, dtype=np.int32)),
                               _ptr(np.ascontiguousarray(nzval, dtype=np.float64)),
                               indexing, _ptr(D))
    return D


def syev_lower(A):
    A = np.array(A, dtype=np.float64, order="F")
    n = A.shape[0]
    W = np.zeros(n)
    buf = A.ravel(order="F").copy()
    lib().orc_syev_lower(n, _ptr(buf), n, _ptr(W))
    return W, buf.reshape(n, n, order="F")


def davidson(N, max_m, rowptr, colind, nzval, D, tol, X, max_restarts=0):
    """davidson (davidson.hpp:128-225); max_restarts > 0 selects the
    labelled restarted variant. Returns (iterations, E, X)."""
    X = np.array(X, dtype=np.float64)
    E = C.c_double()
    it = C.c_int64()
    rs = C.c_int64()
    rc = lib().orc_davidson_restarted(N, max_m, max_restarts,
                                      _ptr(np.ascontiguousarray(rowptr, dtype=np.int64)),
                                      _ptr(np.ascontiguousarray(colind, dtype=np.int32)),
                                      _ptr(np.ascontiguousarray(nzval, dtype=np.float64)),
                                      _ptr(np.ascontiguousarray(D, dtype=np.float64)), tol,
                                      _ptr(X), C.byref(E), C.byref(it), C.byref(rs))
    if rc == ORC_ERR_NOT_CONVERGED:
        raise RuntimeError("Davidson Did Not Converge!")
    if rc:
        raise RuntimeError(f"orc_davidson rc={rc}")
    return it.value, E.value, X


def mlp_loss_grad(P, W, dets, target=None, bf16_points=True, want_grad=True):
    """MLP oracle: (loss, grad float64 or None, predictions)."""
    P = np.ascontiguousarray(P, dtype=np.float32)
    dets = np.ascontiguousarray(dets, dtype=np.uint64)
    n = dets.size // W
    t = None if target is None else np.ascontiguousarray(target, dtype=np.float32)
    g = np.zeros(lib().orc_mlp_num_params()) if want_grad else None
    pred = np.zeros(n)
    loss = lib().orc_mlp_loss_grad(_ptr(P), W, _ptr(dets), _ptr(t), n, int(bf16_points),
                                   _ptr(g), _ptr(pred))
    return loss, g, pred
