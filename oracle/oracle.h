/*
 * CPU oracle for the MACIS determinant-CI hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference
 * algorithms (dfablan/MACIS, include/macis/...) used as the parity checker for
 * the HIP path and as the timed CPU baseline. Product code never links,
 * loads or calls it; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg do.
 *
 * Determinant layout (the reference's std::bitset<N> word layout,
 * bitset_operations.hpp:353-375):
 *   W = 1 (wfn_t<64>):  one uint64, alpha string in bits 0-31, beta in 32-63.
 *   W = 2 (wfn_t<128>): two uint64, word 0 = alpha string, word 1 = beta.
 *
 * Integrals are column-major: T(p,q) = T[p + n*q],
 * V(p,q,r,s) = V[p + n*q + n^2*r + n^3*s]  (types.hpp:22-33).
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---------------------------------------------------- */
#define ORC_OK 0
#define ORC_ERR_ARG -1
#define ORC_ERR_NOMEM -2
#define ORC_ERR_NO_GUESS -3        /* "Davidson: No Guess Provided" */
#define ORC_ERR_NOT_CONVERGED -4   /* "Davidson Did Not Converge!" */
#define ORC_ERR_DID_NOT_GROW -5    /* "Wavefunction didn't grow enough..." */
#define ORC_ERR_REFINE -6          /* "ACCI Refine did not converge" */
#define ORC_ERR_REFINE_SIZE -7     /* "Wavefunction size can't change in refinement" */

/* ---- Hamiltonian generator (hamiltonian_generator.hpp:17-188) -------- */
typedef struct orc_ham orc_ham;

/* Copies T (n*n) and V (n^4) and forms G, G_red, V_red, G2_red, V2_red
 * exactly as generate_integral_intermediates (impl.hpp:25-70). */
orc_ham* orc_ham_create(int64_t norb, const double* T, const double* V);
void orc_ham_destroy(orc_ham* h);
int64_t orc_ham_norb(const orc_ham* h);
/* Read-only views of the intermediates (for tests). which: 0=T 1=V 2=G
 * 3=G_red 4=V_red 5=G2_red 6=V2_red. */
const double* orc_ham_array(const orc_ham* h, int which);

/* H_ij for two determinants (matrix_elements.hpp:14-63). */
double orc_matrix_element(const orc_ham* h, int W, const uint64_t* bra,
                          const uint64_t* ket);

/* single_orbital_ens (fast_diagonals.hpp:31-51): ens[norb]. occ lists are
 * ascending orbital indices. */
void orc_single_orbital_ens(const orc_ham* h, int64_t norb,
                            const uint32_t* ss_occ, int64_t nss,
                            const uint32_t* os_occ, int64_t nos, double* ens);
double orc_fast_diag_single(const orc_ham* h, double hol_en, double par_en,
                            uint32_t orb_hol, uint32_t orb_par, double Hii);
double orc_fast_diag_ss_double(const orc_ham* h, double h1, double h2,
                               double p1, double p2, uint32_t oh1,
                               uint32_t oh2, uint32_t op1, uint32_t op2,
                               double Hii);
double orc_fast_diag_os_double(const orc_ham* h, double hu, double hd,
                               double pu, double pd, uint32_t ohu,
                               uint32_t ohd, uint32_t opu, uint32_t opd,
                               double Hii);

/* ---- bit / sign primitives (sd_operations.hpp:497-551) --------------- */
double orc_single_excitation_sign(uint64_t state, uint32_t p, uint32_t q);
/* Returns sign; writes o1,v1,o2,v2 (doubles_sign_indices, :528-543). */
double orc_doubles_sign_indices(uint64_t bra, uint64_t ket, uint64_t ex,
                                uint32_t* o1, uint32_t* v1, uint32_t* o2,
                                uint32_t* v2);

/* ---- determinant spaces (sd_operations.hpp:421-476) ------------------- */
/* Writes C(norb,na)*C(norb,nb) dets (W words each) in generate_hilbert_space
 * order; returns the count, or the required count if out==NULL. */
int64_t orc_generate_hilbert_space(int W, int64_t norb, int64_t nalpha,
                                   int64_t nbeta, uint64_t* out);
/* CISD space around `state` (generate_cisd_hilbert_space); out==NULL gives
 * the count. */
int64_t orc_generate_cisd_hilbert_space(int W, int64_t norb,
                                        const uint64_t* state, uint64_t* out);

/* ---- CSR Hamiltonian (double_loop.hpp:30-98) -------------------------- */
/* Builds the CSR block over bra x ket with |h| > thresh. On success the
 * malloc'ed arrays are returned through the out pointers (free with
 * orc_free). rowptr has nbra+1 entries. */
int orc_csr_build(const orc_ham* h, int W, const uint64_t* bra, int64_t nbra,
                  const uint64_t* ket, int64_t nket, double thresh,
                  int64_t** rowptr, int32_t** colind, double** nzval,
                  int64_t* nnz);
/* Rows [row_begin,row_end) of the same block, written into caller arrays
 * sized by a previous call with colind==NULL (which only fills counts). */
int orc_csr_rows(const orc_ham* h, int W, const uint64_t* bra,
                 int64_t row_begin, int64_t row_end, const uint64_t* ket,
                 int64_t nket, double thresh, int64_t* row_counts,
                 int32_t* colind, double* nzval);
/* orc_csr_rows with rows spread over OpenMP threads (count pass, prefix
 * sum, fill pass); row_counts is required. Same CSR. */
int orc_csr_rows_parallel(const orc_ham* h, int W, const uint64_t* bra, int64_t row_begin,
                          int64_t row_end, const uint64_t* ket, int64_t nket, double thresh,
                          int64_t* row_counts, int32_t* colind, double* nzval);
/* Rows [row_begin,row_end) of a given CSR (rowptr indexed by absolute row)
 * checked against the double loop entry by entry, rows over OpenMP threads;
 * returns the number of differing rows, *first_bad the lowest (-1 if none). */
int64_t orc_csr_verify_rows(const orc_ham* h, int W, const uint64_t* bra,
                            int64_t row_begin, int64_t row_end,
                            const uint64_t* ket, int64_t nket, double thresh,
                            const int64_t* rowptr, const int32_t* colind,
                            const double* nzval, int64_t* first_bad);
/* Threads an OpenMP parallel region of the oracle uses. */
int orc_num_threads(void);
void orc_free(void* p);

/* ---- sparse BLAS (spmbv.hpp:37-73; submatrix.hpp:264-295) -------------- */
void orc_gespmbv(int64_t K, double alpha, int64_t m, const int64_t* rowptr,
                 const int32_t* colind, const double* nzval, int64_t indexing,
                 const double* V, int64_t LDV, double beta, double* AV,
                 int64_t LDAV);
void orc_extract_diagonal(int64_t m, const int64_t* rowptr,
                          const int32_t* colind, const double* nzval,
                          int64_t indexing, double* D);

/* ---- dense symmetric eigensolver (replaces lapack::syev, Lower) ------- */
/* Cyclic Jacobi on the lower triangle of the column-major n x n matrix A
 * (lda). On return A holds eigenvectors (columns) and W ascending
 * eigenvalues. */
void orc_syev_lower(int64_t n, double* A, int64_t lda, double* W);

/* ---- Davidson (davidson.hpp:61-225) ----------------------------------- */
void orc_diagonal_guess(int64_t N, const double* D, double* X);
int orc_davidson(int64_t N, int64_t max_m, const int64_t* rowptr,
                 const int32_t* colind, const double* nzval, const double* D,
                 double tol, double* X, double* E, int64_t* iters);
/* Restarted variant (not in the reference): max_restarts == 0 is
 * orc_davidson exactly; otherwise an exhausted subspace restarts from the
 * current Ritz vector and the pending preconditioned residual. */
int orc_davidson_restarted(int64_t N, int64_t max_m, int64_t max_restarts,
                           const int64_t* rowptr, const int32_t* colind,
                           const double* nzval, const double* D, double tol,
                           double* X, double* E, int64_t* iters,
                           int64_t* restarts);
/* serial_selected_ci_diag + selected_ci_diag (selected_ci_diag.hpp:75-207):
 * builds H over dets with threshold h_el_tol, then Davidson. C (length
 * ndets) is the guess in / eigenvector out; pass all-zero for the diagonal
 * guess. nnz_out may be NULL. */
int orc_selected_ci_diag(const orc_ham* h, int W, const uint64_t* dets,
                         int64_t ndets, double h_el_tol, int64_t max_m,
                         double res_tol, double* C, double* E,
                         int64_t* nnz_out);

int orc_selected_ci_diag_restarted(const orc_ham* h, int W,
                                   const uint64_t* dets, int64_t ndets,
                                   double h_el_tol, int64_t max_m,
                                   int64_t max_restarts, double res_tol,
                                   double* C, double* E, int64_t* nnz_out);

/* ---- ASCI (determinant_search.hpp:34-51, 53-141, 394-663) ------------- */
typedef struct orc_asci_settings {
  int64_t ntdets_max;
  int64_t ntdets_min;
  int64_t ncdets_max;
  double h_el_tol;
  double rv_prune_tol;
  int64_t pair_size_max;
  int32_t just_singles;
  int64_t grow_factor;
  int64_t max_refine_iter;
  double refine_energy_tol;
} orc_asci_settings;

void orc_asci_default_settings(orc_asci_settings* s);

/* Raw candidate scores before accumulation (asci_contributions_standard
 * without pruning): returns the number of pairs; with out arrays NULL only
 * counts. states: W words per pair. */
int64_t orc_asci_contributions(const orc_ham* h, int W,
                               const orc_asci_settings* s,
                               const uint64_t* cdets, int64_t ncdets,
                               double E_ASCI, const double* C, int64_t norb,
                               uint64_t* states, double* rv);

/* Full asci_search. new_dets must hold ndets_max dets; *n_new receives
 * the number written (top-k followed by the core dets). rv_out (optional,
 * length ndets_max) receives the final -|rv| score of each top-k det. */
int orc_asci_search(const orc_ham* h, int W, const orc_asci_settings* s,
                    int64_t ndets_max, const uint64_t* cdets, int64_t ncdets,
                    double E_ASCI, const double* C, int64_t norb,
                    uint64_t* new_dets, int64_t* n_new, double* rv_out);

/* Excitation blocks of a core (the block screen; no reference counterpart):
 * per core det, one block per hole set in the order [alpha single holes |
 * beta single holes | alpha-alpha hole pairs | beta-beta hole pairs |
 * alpha-beta hole pairs]. block_dets (W words each) = the core det with the
 * holes cleared; target = log max |h / (E_ASCI - H_kk)| over the block's
 * screened-in candidates (-INFINITY if none). Either may be NULL; returns
 * the number of blocks. */
int64_t orc_asci_blocks(const orc_ham* h, int W, const orc_asci_settings* s,
                        const uint64_t* cdets, int64_t ncdets, double E_ASCI, int64_t norb,
                        uint64_t* block_dets, double* target);
/* asci_search over the ceil(keep_fraction * nblocks) blocks with the largest
 * scores (ties: lower block index first); no pair_size_max pruning. */
int orc_asci_search_blocks(const orc_ham* h, int W, const orc_asci_settings* s,
                           const float* scores, double keep_fraction, int64_t ndets_max,
                           const uint64_t* cdets, int64_t ncdets, double E_ASCI, const double* C,
                           int64_t norb, uint64_t* new_dets, int64_t* n_new, double* rv_out);

/* asci_iter (iteration.hpp:16-67). wfn/X are in/out, capacity ndets_max;
 * *nwfn is the current size in and the new size out. */
int orc_asci_iter(const orc_ham* h, int W, const orc_asci_settings* s,
                  int64_t ci_max_subspace, double ci_res_tol,
                  double ci_matel_tol, int64_t ndets_max, double E0,
                  uint64_t* wfn, double* X, int64_t* nwfn, int64_t norb,
                  double* E_out);
/* asci_grow (grow.hpp:16-173, no rotation) and asci_refine
 * (refine.hpp:14-68). Buffers must hold s->ntdets_max dets. */
int orc_asci_grow(const orc_ham* h, int W, const orc_asci_settings* s,
                  int64_t ci_max_subspace, double ci_res_tol,
                  double ci_matel_tol, double* E0, uint64_t* wfn, double* X,
                  int64_t* nwfn, int64_t norb);
int orc_asci_refine(const orc_ham* h, int W, const orc_asci_settings* s,
                    int64_t ci_max_subspace, double ci_res_tol,
                    double ci_matel_tol, double* E0, uint64_t* wfn,
                    double* X, int64_t* nwfn, int64_t norb);

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
extern "C" {
#endif
/* ---- MLP pre-screen oracle (no reference counterpart) ----------------- */
/* Forward/backward of the 128 -> 512 -> 512 -> 1 GELU(tanh) network in
 * float64 over parameters P (layout of macis_hip.h). With bf16_points the
 * weights, activations and back-propagated pre-activations are rounded to
 * bf16 where the device kernels round them. Returns the mean squared error
 * (0 without targets); grad (nparam) and pred (n) are optional. */
/* two_index_transform / four_index_transform (transform.cxx:22-115),
 * square (norb_new == norb_old == n), leading dimension n; Y may alias X. */
void orc_two_index_transform(int64_t n, const double* X, const double* U, double* Y);
void orc_four_index_transform(int64_t n, const double* X, const double* U, double* Y);
/* Natural-orbital rotation of asci_grow (grow.hpp:73-114) applied to h in
 * place; occupations (descending) written when non-NULL. */
int orc_natural_orbital_rotation(orc_ham* h, int W, const uint64_t* wfn, int64_t nwfn,
                                 const double* X, double* occupations);
/* asci_grow with grow_with_rot (grow.hpp:16-173); h is rotated in place. */
int orc_asci_grow_rot(orc_ham* h, int W, const orc_asci_settings* s, int64_t ci_max_subspace,
                      double ci_res_tol, double ci_matel_tol, double* E0, uint64_t* wfn,
                      double* X, int64_t* nwfn, int64_t norb, int64_t rot_size_start);
/* form_rdms (double_loop.hpp:116-165); accumulates into ordm / trdm
 * (column-major, norb^2 / norb^4); trdm may be NULL. */
void orc_form_rdms(int64_t norb, int W, const uint64_t* bra, int64_t nbra,
                   const uint64_t* ket, int64_t nket, const double* C, double* ordm,
                   double* trdm);
int64_t orc_mlp_num_params(void);
double orc_mlp_loss_grad(const float* P, int W, const uint64_t* dets,
                         const float* target, int64_t n, int bf16_points,
                         double* grad, double* pred);
#ifdef __cplusplus
}
#endif
