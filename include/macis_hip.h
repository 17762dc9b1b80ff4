/*
 * macis_hip.h -- C-ABI of the MI355X (gfx950) determinant-CI hot path.
 *
 * Plain pointers and sizes only; no C++ or framework types cross this
 * boundary. Every entry point returns MHIP_SUCCESS (0) or a negative
 * MHIP_ERR_* code; mhip_last_error() gives the thread's last message. The
 * C++ mirror of the reference API (macis_amd/include/macis/...) maps the
 * codes to the reference's std::runtime_error messages.
 *
 * Reference seams replaced (paths inside dfablan/MACIS):
 *   - HamiltonianGenerator<N> ctor + generate_integral_intermediates
 *       include/macis/hamiltonian_generator/impl.hpp:14-70
 *   - HamiltonianGenerator<N>::make_csr_hamiltonian_block_{32,64}bit_
 *       include/macis/hamiltonian_generator.hpp:66-72, implemented by
 *       include/macis/hamiltonian_generator/double_loop.hpp:30-114
 *   - sparsexx::spblas::gespmbv
 *       src/sparsexx/include/sparsexx/spblas/spmbv.hpp:37-73
 *   - sparsexx::extract_diagonal_elements
 *       src/sparsexx/include/sparsexx/util/submatrix.hpp:264-295
 *   - macis::davidson / diagonal_guess / SparseMatrixOperator
 *       include/macis/solvers/davidson.hpp:29-225
 *   - macis::selected_ci_diag / serial_selected_ci_diag
 *       include/macis/solvers/selected_ci_diag.hpp:75-207
 *   - macis::asci_search (world_size == 1)
 *       include/macis/asci/determinant_search.hpp:394-663
 *
 * Determinant layout (std::bitset<N> words, bitset_operations.hpp:353-375):
 *   words_per_det = 1 -> wfn_t<64>:  alpha in bits 0-31, beta in bits 32-63
 *   words_per_det = 2 -> wfn_t<128>: word 0 alpha, word 1 beta
 * Integrals are column-major: T(p,q) = T[p + n*q],
 *   V(p,q,r,s) = V[p + n*q + n^2*r + n^3*s]   (types.hpp:22-33).
 *
 * Device CSR layout (one GPU, HBM resident): rowptr int64[m+1] (0-based),
 * colind int32[nnz], nzval double[nnz]; columns ascending within a row.
 *
 * Threading: a context owns one HIP stream and is used by one host thread
 * at a time (the reference generator is not re-entrant either,
 * hamiltonian_generator.hpp:40-64).
 */
#ifndef MACIS_HIP_H
#define MACIS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MHIP_VERSION 10000

/* ---- status codes ------------------------------------------------------ */
#define MHIP_SUCCESS 0
#define MHIP_ERR_INVALID_ARG -1
#define MHIP_ERR_HIP -2
#define MHIP_ERR_OUT_OF_MEMORY -3
#define MHIP_ERR_NO_INTEGRALS -4
#define MHIP_ERR_NOT_CONVERGED -5   /* davidson.hpp:221 */
#define MHIP_ERR_NO_GUESS -6        /* davidson.hpp:134 */
#define MHIP_ERR_UNSUPPORTED -7     /* hamiltonian_generator.hpp:153 */
#define MHIP_ERR_INDEX_OVERFLOW -8  /* nnz does not fit the index type */
#define MHIP_ERR_NO_DEVICE -9

typedef struct mhip_context mhip_context;
typedef struct mhip_csr mhip_csr;

/* Phase timings of the last call on a context, milliseconds measured with
 * HIP events on the context's stream. */
typedef struct mhip_timings {
  double ham_bounds_ms;   /* ket tile AND/OR bounds */
  double ham_count_ms;    /* structural count pass (XOR/popcount scan) */
  double ham_scan_ms;     /* device exclusive scan of row counts */
  double ham_fill_ms;     /* structural fill pass (ascending columns) */
  double ham_elem_ms;     /* Slater-Condon elements + threshold marks */
  double ham_compact_ms;  /* threshold compaction (0 when nothing dropped) */
  double ham_total_ms;    /* host wall time of the whole build call */
  int64_t ham_nnz_struct; /* pairs with popcount(bra^ket) <= 4 */
  int64_t ham_nnz;        /* kept after |h| > thresh */
  double dav_spmv_ms;     /* summed SpMV time inside the last Davidson */
  int64_t dav_spmv_calls;
  double dav_total_ms;
  int64_t dav_iters;
  int64_t dav_restarts;
  double dav_residual;
  double asci_contrib_ms; /* candidate generation */
  double asci_sort_ms;    /* sort + accumulate + dedupe */
  double asci_topk_ms;    /* exclusion of core dets + top-k selection */
  double asci_total_ms;
  int64_t asci_npairs;    /* raw candidates */
  int64_t asci_nunique;   /* unique candidates after accumulation */
  double asci_nn_ms;      /* block screen: scoring (network) + block selection */
  int64_t asci_nsurvivors; /* unique non-core candidates entering the top-k */
  int64_t ham_path;       /* structure pass of the last build: MHIP_BUILD_* */
  double dav_layout_ms;   /* sliced copy / panel split made for the last Davidson */
  int64_t dav_sliced;     /* 1: its SpMVs ran on the sliced copy, 2: on column panels,
                             3: on the x-tile panels */
  int64_t asci_nblocks;      /* excitation blocks of the core (block-screened search) */
  int64_t asci_nblocks_kept; /* blocks whose candidates were generated */
  /* matrix bytes one SpMV of the last Davidson streams from its layout: the
   * sliced copy's values, column words (16-bit offsets in narrow slices,
   * 32-bit columns elsewhere) and slot bases, or the CSR's 12 bytes per
   * nonzero and its row pointers (x and y not counted) */
  double dav_spmv_matrix_bytes;
} mhip_timings;

/* Structure passes of the Hamiltonian build. AUTO takes the product-space
 * walk when the determinant list is a complete product of spin strings
 * (every run of equal outer string carries the same inner-string sequence,
 * as the lists of generate_hilbert_space and their bitset_less-sorted form
 * are); otherwise, for square lists of at least 32,768 determinants with one
 * alpha and one beta electron count, the spin-string group scan (each row
 * scans the determinants sharing its alpha or its beta string and those of
 * the single-excitation neighbour strings, then its columns are sorted);
 * otherwise the tiled XOR/popcount double loop. All give the same CSR.
 * SPIN_GROUPS forces the group scan (MHIP_ERR_INVALID_ARG when the list is
 * not square or its electron counts vary). */
#define MHIP_BUILD_AUTO 0
#define MHIP_BUILD_TILED 1
#define MHIP_BUILD_PRODUCT 2
#define MHIP_BUILD_SPIN_GROUPS 3

/* Settings of asci_search (ASCISettings, determinant_search.hpp:34-51).
 * pair_size_max / rv_prune_tol prune as the reference does (:120-137). */
typedef struct mhip_asci_settings {
  int64_t ncdets_max;
  double h_el_tol;
  double rv_prune_tol;
  int64_t pair_size_max;
  int32_t just_singles;
} mhip_asci_settings;

/* ---- library / context --------------------------------------------------- */
int mhip_get_version(void);
const char* mhip_error_string(int code);
/* Message of this host thread's last call ("" when it succeeded). */
const char* mhip_last_error(void);

/* Creates a context on `device` with its own HIP stream. */
int mhip_context_create(int device, mhip_context** ctx);
int mhip_context_destroy(mhip_context* ctx);
/* hipStream_t of the context (for events / interop). */
int mhip_context_get_stream(mhip_context* ctx, void** stream);
int mhip_synchronize(mhip_context* ctx);
int mhip_get_timings(mhip_context* ctx, mhip_timings* out);
/* Selects the structure passes of later builds on this context (no
 * counterpart in the reference, which has one serial loop): MHIP_BUILD_AUTO
 * (default), MHIP_BUILD_TILED, MHIP_BUILD_PRODUCT (fails with
 * MHIP_ERR_INVALID_ARG when a square build's list is not a product space) or
 * MHIP_BUILD_SPIN_GROUPS (see above). */
int mhip_context_set_build_path(mhip_context* ctx, int path);
/* Frees device storage the context keeps for reuse (the last internal
 * Hamiltonian of mhip_selected_ci_diag and the Davidson subspace vectors);
 * no counterpart in the reference. */
int mhip_context_release_cached(mhip_context* ctx);

/* Device memory helpers, so callers need no HIP runtime of their own. */
int mhip_malloc(mhip_context* ctx, size_t bytes, void** dptr);
int mhip_free(mhip_context* ctx, void* dptr);
int mhip_memcpy_h2d(mhip_context* ctx, void* dst, const void* src,
                    size_t bytes);
int mhip_memcpy_d2h(mhip_context* ctx, void* dst, const void* src,
                    size_t bytes);
int mhip_memcpy_d2d(mhip_context* ctx, void* dst, const void* src,
                    size_t bytes);
int mhip_memset(mhip_context* ctx, void* dst, int value, size_t bytes);

/* ---- integrals ----------------------------------------------------------- */
/* HamiltonianGenerator(T, V) + generate_integral_intermediates
 * (impl.hpp:14-70). T (norb^2) and V (norb^4) are HOST pointers, copied;
 * G, G_red, V_red, G2_red and V2_red are formed and kept on the device.
 * Calling again re-uploads (asci_grow rotates integrals in place,
 * grow.hpp:104-122). norb <= 64. */
int mhip_set_integrals(mhip_context* ctx, int64_t norb, const double* T,
                       const double* V);

/* ---- Hamiltonian generation (K1) ------------------------------------------ */
/* make_csr_hamiltonian_block over bra x ket with |h| > thresh, built on the
 * device (double_loop.hpp:30-98): structural count, device scan, ascending
 * fill, Slater-Condon elements, threshold compaction. d_bra / d_ket are
 * DEVICE arrays of words_per_det uint64 per determinant. Zero determinants
 * give empty rows and are never columns. nbra == 0 or nket == 0 gives an
 * empty matrix (csr_hamiltonian.hpp:30-35). The CSR is owned by *out. */
/* The tiled passes stage columns in a device pool; the environment variable
 * MHIP_STAGE_POOL_CHUNKS caps it (128-entry chunks; 0 disables staging), the
 * build falling back to a count and a fill walk when it overflows. */
int mhip_ham_build(mhip_context* ctx, int words_per_det,
                   const uint64_t* d_bra, int64_t nbra, const uint64_t* d_ket,
                   int64_t nket, double thresh, mhip_csr** out);

/* Diagonal Slater-Condon elements H_ii of each determinant
 * (matrix_element_diag, matrix_elements.hpp:108-135) into DEVICE d_diag. */
int mhip_ham_diagonal(mhip_context* ctx, int words_per_det,
                      const uint64_t* d_dets, int64_t ndets, double* d_diag);

/* form_rdms (double_loop.hpp:116-165, rdm_contributions util/rdms.hpp:15-132):
 * ADDS the spin-summed 1-RDM ordm(p,q) = d_ordm[p + norb q] and 2-RDM
 * trdm(p,q,r,s) = d_trdm[p + norb q + norb^2 r + norb^3 s] (column-major, as
 * the reference's matrix_span / rank4_span) of the pairs bra i x ket j with
 * popcount(bra ^ ket) <= 4 and |C[i] C[j]| > 1e-16, norb from
 * mhip_set_integrals. One coefficient array serves both sides, as in the
 * reference, so d_C holds max(nbra, nket) values. d_trdm may be NULL (1-RDM
 * only, with the reference's branch order: same-spin doubles then reach the
 * diagonal branch). All arrays are DEVICE arrays; the caller zeroes the
 * outputs, which must be ordinary (coarse-grained) device memory such as
 * mhip_malloc returns: the sums use hardware FP64 atomics, equal to the
 * reference's up to summation order. */
int mhip_form_rdms(mhip_context* ctx, int words_per_det, const uint64_t* d_bra,
                   int64_t nbra, const uint64_t* d_ket, int64_t nket,
                   const double* d_C, double* d_ordm, double* d_trdm);

/* ---- CSR handles ----------------------------------------------------------- */
int mhip_csr_info(const mhip_csr* csr, int64_t* m, int64_t* n, int64_t* nnz);
/* Device arrays of the handle (borrowed; valid until mhip_csr_destroy). */
int mhip_csr_device_arrays(const mhip_csr* csr, const int64_t** d_rowptr,
                           const int32_t** d_colind, const double** d_nzval);
/* Copy to HOST arrays in the reference csr_matrix<double,int32_t> /
 * <double,int64_t> layout (hamiltonian_generator.hpp:27-28). The int32
 * variant fails with MHIP_ERR_INDEX_OVERFLOW when nnz > INT32_MAX. */
int mhip_csr_copy_to_host_i32(mhip_context* ctx, const mhip_csr* csr,
                              int32_t* rowptr, int32_t* colind,
                              double* nzval);
int mhip_csr_copy_to_host_i64(mhip_context* ctx, const mhip_csr* csr,
                              int64_t* rowptr, int64_t* colind,
                              double* nzval);
/* Upload a host CSR (rowptr int64[m+1], colind int32[nnz]) with base
 * `indexing` (0 or 1, csr_matrix.hpp:65-82) into a device handle. */
int mhip_csr_create_from_host(mhip_context* ctx, int64_t m, int64_t n,
                              int64_t nnz, int64_t indexing,
                              const int64_t* rowptr, const int32_t* colind,
                              const double* nzval, mhip_csr** out);
int mhip_csr_destroy(mhip_csr* csr);

/* ---- sparse BLAS (K2) ------------------------------------------------------ */
/* gespmbv: AV[:,k] = alpha*A*V[:,k] + beta*AV[:,k] for k < K, on DEVICE
 * column-major blocks (spmbv.hpp:37-73). */
int mhip_spmv(mhip_context* ctx, const mhip_csr* A, int64_t K, double alpha,
              const double* d_V, int64_t LDV, double beta, double* d_AV,
              int64_t LDAV);
/* Attaches a sliced copy (SELL-64: 64-row slices stored slot-major, rows
 * padded to their slice's longest) of a square matrix to the handle, used
 * from then on by every single-vector product with it
 * (mhip_spmv columns, Davidson). No counterpart in the reference (sparsexx
 * multiplies its CSR directly); for callers that apply one matrix many
 * times. The copy needs 12 bytes per stored entry and is made only when the
 * device keeps 1 GiB free beside it; with MHIP_SPMV_SLICED_AUTO (what
 * Davidson uses) also only when padding adds at most a quarter of the
 * nonzeros. *sliced (may be NULL) says whether the handle has one. Products
 * equal the CSR ones up to the order of the row sums. */
#define MHIP_SPMV_SLICED_AUTO 0
#define MHIP_SPMV_SLICED_ALWAYS 1
/* MHIP_SPMV_COLUMN_PANELS attaches instead a column-panel split of a square
 * matrix without a sliced copy: the columns are cut into 8 panels of about
 * equal nonzero counts and each of the GPU's 8 XCDs multiplies one panel's
 * entries of every row, so its L2 holds one panel of x (for x vectors larger
 * than an L2; per row 8 int32 panel offsets and 8 partial sums, 96 bytes,
 * no copy of the entries). Davidson attaches it by itself to matrices that
 * get no sliced copy and have at least 262,144 columns. *sliced receives 2.
 * Products equal the CSR ones up to the order of the row sums. */
#define MHIP_SPMV_COLUMN_PANELS 2
/* MHIP_SPMV_X_TILES attaches instead a panel-major copy of a square matrix
 * without a sliced copy: columns cut into panels of 12,288 (96 KB of x, held
 * in LDS while a workgroup streams the panel's entries of its 2,048 rows),
 * entries stored as 16-bit panel-local columns and values (10 bytes per
 * nonzero, plus 4 bytes per row and panel). Only on request: on the
 * 2,000,000-determinant selected-CI matrix it is slower than the column
 * panels (one 96 KB tile fits a CU's LDS, so a CU's waves synchronise on
 * every piece of entries); a Davidson on a matrix that has it uses it.
 * *sliced receives 3. Products equal the CSR ones up to the order of the
 * row sums (per row: panels in order, each panel's entries in a fixed
 * order) and are the same bits call to call. */
#define MHIP_SPMV_X_TILES 3
int mhip_csr_prepare_spmv(mhip_context* ctx, const mhip_csr* A, int policy, int* sliced);
/* extract_diagonal_elements (submatrix.hpp:264-295) into DEVICE d_D. */
int mhip_extract_diagonal(mhip_context* ctx, const mhip_csr* A, double* d_D);

/* ---- Davidson (K3-K5) ------------------------------------------------------ */
/* davidson(N, max_m, SparseMatrixOperator(A), D, tol, X) with everything
 * on the device (davidson.hpp:128-225). d_X (DEVICE, length N) is the guess
 * in and the Ritz vector out; d_D may be NULL (extracted from A). With
 * max_restarts == 0 the reference algorithm is followed exactly and
 * MHIP_ERR_NOT_CONVERGED is returned when max_m is exhausted; with
 * max_restarts > 0 (a labelled variant, not in the reference) an exhausted
 * subspace restarts from [X, previous Ritz vector, next preconditioned
 * residual] with A applied by combination of the stored products. E receives
 * LAM(0), iters the iteration count (davidson.hpp:224). max_m is clamped to
 * N (davidson.hpp:140) and may be up to 1024; subspaces wider than 32
 * columns are projected in 32-column blocks. */
int mhip_davidson(mhip_context* ctx, const mhip_csr* A, int64_t max_m,
                  int64_t max_restarts, double tol, const double* d_D,
                  double* d_X, double* E, int64_t* iters);

/* Tall-skinny FP64 GEMM on MFMA: C(k x m, ldc) = A^T B with A N x k (lda)
 * and B N x m (ldb), all DEVICE, k, m <= 32. This is the projection of
 * lobpcgxx::rayleigh_ritz (rayleigh_ritz.hpp:73-74) and of the CGS2 inner
 * products (davidson.hpp:114-121) that Davidson runs on the device. */
int mhip_gemm_tn(mhip_context* ctx, int64_t N, int k, const double* d_A,
                 int64_t lda, int m, const double* d_B, int64_t ldb,
                 double* d_C, int ldc);

/* mhip_gemm_tn without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The block partials
 * (at most 256 blocks x 1024 doubles = 2 MiB, the maximum of every N) live in
 * a static 2 MiB device array per GPU, set up when a context is created on
 * that GPU and held until the process ends. Same arguments, limits and
 * result as mhip_gemm_tn. Calls that overlap on one GPU (two contexts or
 * threads at once) share the array and race. */
int mhip_gemm_tn_noalloc(mhip_context* ctx, int64_t N, int k, const double* d_A,
                         int64_t lda, int m, const double* d_B, int64_t ldb,
                         double* d_C, int ldc);

/* selected_ci_diag (selected_ci_diag.hpp:124-207, serial path): build H over
 * DEVICE d_dets with threshold h_el_tol, extract the diagonal, guess rule
 * (reuse d_C if max|C| > 1/ndets, else unit vector at argmin D), Davidson.
 * d_C (DEVICE, length ndets) is guess in / eigenvector out. If H_out is not
 * NULL the Hamiltonian handle is returned to the caller; otherwise its
 * device storage is kept by the context and reused by the next build of the
 * same or smaller size (mhip_context_release_cached frees it). */
int mhip_selected_ci_diag(mhip_context* ctx, int words_per_det,
                          const uint64_t* d_dets, int64_t ndets,
                          double h_el_tol, int64_t max_m, int64_t max_restarts,
                          double res_tol, double* d_C, double* E,
                          int64_t* nnz, mhip_csr** H_out);

/* ---- ASCI search (K6-K8) ---------------------------------------------------- */
/* asci_search (determinant_search.hpp:394-663, world_size == 1): scores
 * all singles/doubles of the DEVICE core dets d_cdets (coefficients DEVICE
 * d_C, length ncdets) against E_ASCI, accumulates duplicates, removes core
 * dets and keeps the ndets_max - ncdets largest |rv|. Writes the new list
 * (top-k, then the core dets) to DEVICE d_new_dets (capacity ndets_max) and
 * its length to *n_new. Top-k ties are broken by determinant order. */
int mhip_asci_search(mhip_context* ctx, int words_per_det,
                     const mhip_asci_settings* settings, int64_t ndets_max,
                     const uint64_t* d_cdets, int64_t ncdets, double E_ASCI,
                     const double* d_C, int64_t norb, uint64_t* d_new_dets,
                     int64_t* n_new);

/* ---- batched CAS-CI (no counterpart in the reference, which solves one
 * instance per compute_casci_rdms call, util/cas.hpp:32-64) ------------------
 * ninst independent problems, each with its own norb (<= 32 for wfn_t<64>,
 * <= 64 for wfn_t<128> through mhip_batch_create_w),
 * HOST determinants dets[det_offsets[b] .. det_offsets[b+1]) and HOST
 * integrals T[b] (norb^2), V[b] (norb^4). All instances share one
 * block-diagonal device CSR built by one launch per stage. */
typedef struct mhip_batch mhip_batch;
int mhip_batch_create(mhip_context* ctx, int ninst, const int32_t* norb,
                      const int64_t* det_offsets, const uint64_t* dets,
                      const double* const* T, const double* const* V,
                      mhip_batch** out);
/* mhip_batch_create for any determinant width: words_per_det 1 (wfn_t<64>,
 * norb <= 32) or 2 (wfn_t<128>, norb <= 64); dets holds words_per_det
 * uint64 per determinant (word 0 alpha, word 1 beta for wfn_t<128>), as in
 * the per-instance path (compute_casci_rdms works for any N, cas.hpp:32-64). */
int mhip_batch_create_w(mhip_context* ctx, int words_per_det, int ninst,
                        const int32_t* norb, const int64_t* det_offsets,
                        const uint64_t* dets, const double* const* T,
                        const double* const* V, mhip_batch** out);
int mhip_batch_destroy(mhip_batch* batch);
int mhip_batch_info(const mhip_batch* batch, int64_t* ndets, int64_t* nnz);
/* Hamiltonians of all instances with |h| > thresh (double_loop.hpp:30-98
 * per instance). */
int mhip_batch_build(mhip_context* ctx, mhip_batch* batch, double thresh);
/* Autotuner: picks the SpMV lanes per row (8/16/32/64) and the rows per
 * Hamiltonian work item for every instance-size bucket (largest first);
 * lanes_out (optional) receives one entry per bucket. It then times that CSR
 * schedule against the sliced copy of the block-diagonal matrix (when the
 * copy's padding rule allows one) and keeps the faster layout. */
int mhip_batch_tune(mhip_context* ctx, mhip_batch* batch, int32_t* lanes_out);
/* SpMV layout of the batched Davidson: layout 1 forces the sliced copy
 * (MHIP_ERR_UNSUPPORTED when the padding rule refuses it), 0 the CSR
 * schedule, -1 back to the default (the sliced copy when allowed), -2 only
 * queries. *current (optional) receives 1 (sliced), 0 (CSR schedule) or -1
 * (undecided: the default applies at the next Davidson). */
int mhip_batch_spmv_layout(mhip_context* ctx, mhip_batch* batch, int layout, int* current);
/* Davidson of every instance in lockstep (davidson.hpp:128-225 per
 * instance; max_restarts > 0 selects the restarted variant), iterations
 * captured as hipGraphs when use_graphs != 0. E and residual (optional) are
 * indexed by instance; ms_out (optional) gets the wall time in ms. Returns
 * MHIP_ERR_NOT_CONVERGED if any instance did not converge. */
int mhip_batch_davidson(mhip_context* ctx, mhip_batch* batch, int64_t max_m,
                        int64_t max_restarts, double tol, int use_graphs,
                        double* E, double* residual, double* ms_out);

/* Diagonal H_ii of every instance of the built batch (HOST, caller order:
 * det_offsets). */
int mhip_batch_diagonal(mhip_context* ctx, mhip_batch* batch, double* D);
/* Davidson guess for later mhip_batch_davidson calls (HOST, caller order,
 * each instance's part normalised by the caller, as davidson.hpp:151 copies
 * the guess into V(:,0) unchanged); NULL returns to the diagonal guess of
 * davidson.hpp:61-68 (unit vector at each instance's first minimum of D). */
int mhip_batch_set_guess(mhip_context* ctx, mhip_batch* batch, const double* X);
/* Ritz vectors of the last mhip_batch_davidson (HOST, caller order). */
int mhip_batch_vectors(mhip_context* ctx, mhip_batch* batch, double* X);
/* Instance `inst` (caller index) of the built block-diagonal Hamiltonian
 * as a HOST CSR with instance-local indices (rowptr int64[ndets_b + 1] from
 * 0, colind int32, nzval): *nnz receives its nonzero count; with rowptr NULL
 * only the count is returned. */
int mhip_batch_instance_csr(mhip_context* ctx, mhip_batch* batch, int inst, int64_t* rowptr,
                            int32_t* colind, double* nzval, int64_t* nnz);

/* ---- learned ASCI pre-screen (no counterpart in the reference) ---------------
 * An MLP 128 -> 512 -> 512 -> 1 with GELU on the occupation bits of a
 * determinant (bits 0-63 alpha, 64-127 beta for wfn_t<128>; wfn_t<64> uses
 * the 64 bits of its word), predicting the block targets of
 * mhip_asci_blocks (log of the block's largest |h / (E - H_kk)|). BF16
 * operands on MFMA with FP32 accumulation; FP32 master weights, Adam.
 * Parameters are one FP32 block: W1 (512x128, row-major) | b1 (512) |
 * W2 (512x512) | b2 (512) | w3 (512) | b3 (1). */
typedef struct mhip_mlp mhip_mlp;
int mhip_mlp_create(mhip_context* ctx, uint64_t seed, mhip_mlp** out);
int mhip_mlp_destroy(mhip_mlp* mlp);
int64_t mhip_mlp_num_params(void);
int mhip_mlp_get_params(mhip_context* ctx, const mhip_mlp* mlp, float* host);
int mhip_mlp_set_params(mhip_context* ctx, mhip_mlp* mlp, const float* host);
/* Gradients of the last mhip_mlp_train_step (same layout as the params). */
int mhip_mlp_get_grads(mhip_context* ctx, const mhip_mlp* mlp, float* host);
/* d_out[i] = prediction for DEVICE determinant i (FP32, DEVICE). */
int mhip_mlp_predict(mhip_context* ctx, mhip_mlp* mlp, int words_per_det,
                     const uint64_t* d_dets, int64_t n, float* d_out);
/* One Adam step on the mean-squared error over a DEVICE mini-batch
 * (n <= 4194304); lr == 0 forms the gradients only. */
int mhip_mlp_train_step(mhip_context* ctx, mhip_mlp* mlp, int words_per_det,
                        const uint64_t* d_dets, const float* d_target,
                        int64_t n, float lr, float* loss);
/* Scored candidates of asci_search before the top-k: the unique non-core
 * determinants with their accumulated |rv|. DEVICE outputs of `capacity`
 * entries; *n_out = count. */
int mhip_asci_candidates(mhip_context* ctx, int words_per_det,
                         const mhip_asci_settings* settings,
                         const uint64_t* d_cdets, int64_t ncdets, double E_ASCI,
                         const double* d_C, int64_t norb, uint64_t* d_dets_out,
                         double* d_absrv_out, int64_t capacity, int64_t* n_out);
/* Excitation blocks of a core (the units of the block screen). The
 * candidates of asci_search's core determinant c (determinant_contributions
 * .hpp:25-160) fall into one block per hole set, in the order
 *   [alpha single holes (na) | beta single holes (nb) |
 *    alpha-alpha hole pairs (na(na-1)/2) | beta-beta hole pairs |
 *    alpha-beta hole pairs (na*nb, alpha hole major)]
 * (the singles only with just_singles); hole indices follow the occupied
 * list in ascending orbital order, pairs (x < y) lexicographically. Blocks
 * are numbered core det by core det. A block's determinant is the core
 * determinant with its holes cleared (the network's input); its target is
 * log max |h / (E_ASCI - H_kk)| over the block's candidates that pass the
 * h_el_tol screen (-INF if none): the coefficient-free part of the
 * candidates' scores. DEVICE outputs of `capacity` entries (either may be
 * NULL; with both NULL only *n_blocks is set, which is also set when the
 * capacity is too small). */
int mhip_asci_blocks(mhip_context* ctx, int words_per_det, const mhip_asci_settings* settings,
                     const uint64_t* d_cdets, int64_t ncdets, double E_ASCI, int64_t norb,
                     uint64_t* d_block_dets, float* d_target, int64_t capacity,
                     int64_t* n_blocks);
/* asci_search over the best-scored blocks only: the ceil(keep_fraction *
 * nblocks) blocks with the largest DEVICE d_block_scores (ties: lower block
 * index first) are generated, their exact contributions accumulated, and
 * the top-k picked as in mhip_asci_search (core dets excluded, ties in
 * determinant order, core appended). Contributions of dropped blocks are
 * not generated: that is the work the screen removes. No pair_size_max
 * pruning replay. keep_fraction 1 gives the plain search's candidates. */
int mhip_asci_search_blocks(mhip_context* ctx, int words_per_det,
                            const mhip_asci_settings* settings, const float* d_block_scores,
                            double keep_fraction, int64_t ndets_max, const uint64_t* d_cdets,
                            int64_t ncdets, double E_ASCI, const double* d_C, int64_t norb,
                            uint64_t* d_new_dets, int64_t* n_new);
/* Learned block screen: mhip_asci_search_blocks with block scores
 * pred * pred_scale + pred_shift + log|C_c| from the network's prediction
 * `pred` on the block's determinant (a network trained on standardised
 * mhip_asci_blocks targets passes their standard deviation and mean). The
 * plain mhip_asci_search is the fallback. */
int mhip_asci_search_screened(mhip_context* ctx, int words_per_det,
                              const mhip_asci_settings* settings, mhip_mlp* mlp,
                              float pred_scale, float pred_shift, double keep_fraction,
                              int64_t ndets_max, const uint64_t* d_cdets, int64_t ncdets,
                              double E_ASCI, const double* d_C, int64_t norb,
                              uint64_t* d_new_dets, int64_t* n_new);
/* BF16 GEMM of the MLP kernels, for parity tests: C (FP32, M x N) = A B^T,
 * A (M x K) and B (N x K) row-major BF16 bit patterns, all DEVICE; M and N
 * multiples of 128, K of 32. */
int mhip_gemm_bf16_nt(mhip_context* ctx, int M, int N, int K,
                      const uint16_t* d_A, const uint16_t* d_B, float* d_C);

#ifdef __cplusplus
}
#endif

#endif /* MACIS_HIP_H */
