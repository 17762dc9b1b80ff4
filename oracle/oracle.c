/*
 * CPU oracle for the MACIS determinant-CI hot path -- TEST INFRASTRUCTURE.
 *
 * A plain-C restatement of the reference algorithms, function by function,
 * each citing the reference file:line it follows (paths relative to the
 * dfablan/MACIS source tree). It is the checker for the HIP path and the
 * timed CPU baseline; product code never links or calls it.
 *
 * Floating-point policy: built with -ffp-contract=off so every sum is
 * evaluated in the reference's source order without fused multiply-adds
 * (matrix elements involve only additions and multiplications by +-1).
 *
 * Pinning (see tests/test_oracle_*.py):
 *   - matrix elements / intermediates: H2O cc-pVDZ known answers (E_HF,
 *     excited diagonals, MP2, Brillouin), brute-force second quantisation
 *     on small spaces, dense eigensolves;
 *   - CSR build: h2o.ccpvdz.cisd.rowptr.bin fixture (exact), FCI row counts,
 *     pattern symmetry;
 *   - gespmbv: the 6x6 known answer;
 *   - Davidson: H2O CISD E0 fixture and dense eigensolves;
 *   - ASCI grow/refine: asci.cxx energies; search: brute-force rescoring;
 *   - form_rdms: h2o.ccpvdz.cas88.rdms.bin fixture (entry by entry), second
 *     quantisation on small spaces, the HF identities and traces of
 *     tests/double_loop.cxx:180-223, tr(D1 T) + tr(D2 V) = C^T H C.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ===================================================================== */
/* determinant helpers                                                    */
/* ===================================================================== */
typedef struct {
  uint64_t a; /* alpha string */
  uint64_t b; /* beta string  */
} det_t;

static inline det_t load_det(int W, const uint64_t* p) {
  det_t d;
  if(W == 1) {
    d.a = p[0] & 0xffffffffULL;
    d.b = p[0] >> 32;
  } else {
    d.a = p[0];
    d.b = p[1];
  }
  return d;
}

static inline void store_det(int W, det_t d, uint64_t* p) {
  if(W == 1)
    p[0] = d.a | (d.b << 32);
  else {
    p[0] = d.a;
    p[1] = d.b;
  }
}

static inline int spin_bits(int W) { return W == 1 ? 32 : 64; }
static inline int popc(uint64_t x) { return __builtin_popcountll(x); }
static inline uint32_t ffs0(uint64_t x) { return (uint32_t)__builtin_ctzll(x); }
/* full_mask<N>(i): bits [0, i)  (bitset_operations.hpp:225-230) */
static inline uint64_t lowmask(uint32_t i) {
  return i >= 64 ? ~0ULL : ((1ULL << i) - 1ULL);
}

/* bits_to_indices (bitset_operations.hpp:294-306) */
static int bits_to_indices(uint64_t x, uint32_t* idx) {
  int c = 0;
  while(x) {
    idx[c++] = ffs0(x);
    x &= x - 1;
  }
  return c;
}

/* bitset_to_occ_vir (sd_operations.hpp:82-97): occupied = set bits,
 * virtual = the first (norb - nocc) clear bits, ascending. */
static void occ_vir(int64_t norb, uint64_t s, uint32_t* occ, int* nocc,
                    uint32_t* vir, int* nvir) {
  *nocc = bits_to_indices(s, occ);
  int nv = (int)norb - *nocc;
  uint64_t t = ~s;
  for(int i = 0; i < nv; ++i) {
    uint32_t a = ffs0(t);
    vir[i] = a;
    t &= t - 1;
  }
  *nvir = nv;
}

/* ===================================================================== */
/* sign primitives (sd_operations.hpp:497-551)                            */
/* ===================================================================== */
double orc_single_excitation_sign(uint64_t state, uint32_t p, uint32_t q) {
  /* :503-513 -- parity of occupied orbitals strictly between p and q */
  uint64_t m;
  if(p > q)
    m = state & (lowmask(p) ^ lowmask(q + 1));
  else
    m = state & (lowmask(q) ^ lowmask(p + 1));
  return (popc(m) & 1) ? -1. : 1.;
}

double orc_doubles_sign_indices(uint64_t bra, uint64_t ket, uint64_t ex,
                                uint32_t* o1, uint32_t* v1, uint32_t* o2,
                                uint32_t* v2) {
  /* :528-543 -- two sequential singles on the updated ket */
  *o1 = ffs0(ket & ex);
  *v1 = ffs0(bra & ex);
  double sign = orc_single_excitation_sign(ket, *v1, *o1);
  ket ^= (1ULL << *o1) | (1ULL << *v1);
  ex ^= (1ULL << *o1) | (1ULL << *v1);
  *o2 = ffs0(ket & ex);
  *v2 = ffs0(bra & ex);
  sign *= orc_single_excitation_sign(ket, *v2, *o2);
  return sign;
}

/* ===================================================================== */
/* Hamiltonian generator                                                  */
/* ===================================================================== */
struct orc_ham {
  int64_t n;
  double* T;
  double* V;
  double* G;
  double* G_red;
  double* V_red;
  double* G2_red;
  double* V2_red;
};

static void ham_intermediates(orc_ham* h);

orc_ham* orc_ham_create(int64_t n, const double* T, const double* V) {
  /* HamiltonianGenerator ctor + generate_integral_intermediates
   * (impl.hpp:14-70) */
  if(n <= 0 || !T || !V) return NULL;
  orc_ham* h = (orc_ham*)calloc(1, sizeof(orc_ham));
  if(!h) return NULL;
  const int64_t n2 = n * n, n3 = n2 * n, n4 = n3 * n;
  h->n = n;
  h->T = (double*)malloc(sizeof(double) * n2);
  h->V = (double*)malloc(sizeof(double) * n4);
  h->G = (double*)malloc(sizeof(double) * n4);
  h->G_red = (double*)malloc(sizeof(double) * n3);
  h->V_red = (double*)malloc(sizeof(double) * n3);
  h->G2_red = (double*)malloc(sizeof(double) * n2);
  h->V2_red = (double*)malloc(sizeof(double) * n2);
  if(!h->T || !h->V || !h->G || !h->G_red || !h->V_red || !h->G2_red ||
     !h->V2_red) {
    orc_ham_destroy(h);
    return NULL;
  }
  memcpy(h->T, T, sizeof(double) * n2);
  memcpy(h->V, V, sizeof(double) * n4);
  ham_intermediates(h);
  return h;
}

/* generate_integral_intermediates (impl.hpp:25-70) from h->V */
static void ham_intermediates(orc_ham* h) {
  const int64_t n = h->n, n2 = n * n, n3 = n2 * n;
  const double* V = h->V;
#define IDX4(i, j, k, l) ((i) + n * (j) + n2 * (k) + n3 * (l))
  /* G(i,j,k,l) = V(i,j,k,l) - V(i,l,k,j)   (:36-44) */
  for(int64_t i = 0; i < n; ++i)
    for(int64_t j = 0; j < n; ++j)
      for(int64_t k = 0; k < n; ++k)
        for(int64_t l = 0; l < n; ++l)
          h->G[IDX4(i, j, k, l)] = V[IDX4(i, j, k, l)] - V[IDX4(i, l, k, j)];
  /* G_red(k,i,j) = G(k,k,i,j); V_red(k,i,j) = V(k,k,i,j)   (:46-57) */
  for(int64_t j = 0; j < n; ++j)
    for(int64_t i = 0; i < n; ++i)
      for(int64_t k = 0; k < n; ++k) {
        h->G_red[k + n * i + n2 * j] = h->G[IDX4(k, k, i, j)];
        h->V_red[k + n * i + n2 * j] = V[IDX4(k, k, i, j)];
      }
  /* G2_red(i,j) = 0.5*G(i,i,j,j); V2_red(i,j) = V(i,i,j,j)   (:59-69) */
  for(int64_t j = 0; j < n; ++j)
    for(int64_t i = 0; i < n; ++i) {
      h->G2_red[i + n * j] = 0.5 * h->G[IDX4(i, i, j, j)];
      h->V2_red[i + n * j] = V[IDX4(i, i, j, j)];
    }
#undef IDX4
}

void orc_ham_destroy(orc_ham* h) {
  if(!h) return;
  free(h->T);
  free(h->V);
  free(h->G);
  free(h->G_red);
  free(h->V_red);
  free(h->G2_red);
  free(h->V2_red);
  free(h);
}

int64_t orc_ham_norb(const orc_ham* h) { return h->n; }

const double* orc_ham_array(const orc_ham* h, int which) {
  switch(which) {
    case 0: return h->T;
    case 1: return h->V;
    case 2: return h->G;
    case 3: return h->G_red;
    case 4: return h->V_red;
    case 5: return h->G2_red;
    case 6: return h->V2_red;
    default: return NULL;
  }
}

/* matrix_element_4 (matrix_elements.hpp:65-71) */
static double me_4(const orc_ham* h, uint64_t bra, uint64_t ket, uint64_t ex) {
  uint32_t o1, v1, o2, v2;
  const int64_t n = h->n, n2 = n * n, n3 = n2 * n;
  double sign = orc_doubles_sign_indices(bra, ket, ex, &o1, &v1, &o2, &v2);
  return sign * h->G[v1 + n * o1 + n2 * v2 + n3 * o2];
}

/* matrix_element_22 (matrix_elements.hpp:73-84) */
static double me_22(const orc_ham* h, det_t bra, det_t ket, uint64_t exa,
                    uint64_t exb) {
  const int64_t n = h->n, n2 = n * n, n3 = n2 * n;
  uint32_t o1 = ffs0(ket.a & exa), v1 = ffs0(bra.a & exa);
  double sa = orc_single_excitation_sign(ket.a, v1, o1);
  uint32_t o2 = ffs0(ket.b & exb), v2 = ffs0(bra.b & exb);
  double sb = orc_single_excitation_sign(ket.b, v2, o2);
  double sign = sa * sb;
  return sign * h->V[v1 + n * o1 + n2 * v2 + n3 * o2];
}

/* matrix_element_2 (matrix_elements.hpp:86-106) */
static double me_2(const orc_ham* h, uint64_t bra, uint64_t ket, uint64_t ex,
                   const uint32_t* occ_same, int nsame,
                   const uint32_t* occ_othr, int nothr) {
  const int64_t n = h->n, n2 = n * n;
  uint32_t o1 = ffs0(ket & ex), v1 = ffs0(bra & ex);
  double sign = orc_single_excitation_sign(ket, v1, o1);
  double h_el = h->T[v1 + n * o1];
  const double* G_ov = h->G_red + n * v1 + n2 * o1;
  for(int p = 0; p < nsame; ++p) h_el += G_ov[occ_same[p]];
  const double* V_ov = h->V_red + n * v1 + n2 * o1;
  for(int p = 0; p < nothr; ++p) h_el += V_ov[occ_othr[p]];
  return sign * h_el;
}

/* matrix_element_diag (matrix_elements.hpp:108-135) */
static double me_diag(const orc_ham* h, const uint32_t* oa, int na,
                      const uint32_t* ob, int nb) {
  const int64_t n = h->n;
  double h_el = 0;
  for(int i = 0; i < na; ++i) h_el += h->T[oa[i] * (n + 1)];
  for(int i = 0; i < nb; ++i) h_el += h->T[ob[i] * (n + 1)];
  for(int q = 0; q < na; ++q)
    for(int p = 0; p < na; ++p) h_el += h->G2_red[oa[p] + n * oa[q]];
  for(int q = 0; q < nb; ++q)
    for(int p = 0; p < nb; ++p) h_el += h->G2_red[ob[p] + n * ob[q]];
  for(int q = 0; q < nb; ++q)
    for(int p = 0; p < na; ++p) h_el += h->V2_red[oa[p] + n * ob[q]];
  return h_el;
}

/* matrix_element dispatch (matrix_elements.hpp:32-63) */
static double me_dispatch(const orc_ham* h, det_t bra, det_t ket,
                          const uint32_t* oa, int na, const uint32_t* ob,
                          int nb) {
  const uint64_t exa = bra.a ^ ket.a, exb = bra.b ^ ket.b;
  const int ca = popc(exa), cb = popc(exb);
  if(ca + cb > 4) return 0.;
  if(ca == 4) return me_4(h, bra.a, ket.a, exa);
  if(cb == 4) return me_4(h, bra.b, ket.b, exb);
  if(ca == 2 && cb == 2) return me_22(h, bra, ket, exa, exb);
  if(ca == 2) return me_2(h, bra.a, ket.a, exa, oa, na, ob, nb);
  if(cb == 2) return me_2(h, bra.b, ket.b, exb, ob, nb, oa, na);
  return me_diag(h, oa, na, ob, nb);
}

double orc_matrix_element(const orc_ham* h, int W, const uint64_t* bra_p,
                          const uint64_t* ket_p) {
  /* matrix_elements.hpp:14-30 */
  uint32_t oa[64], ob[64];
  det_t bra = load_det(W, bra_p), ket = load_det(W, ket_p);
  int na = bits_to_indices(bra.a, oa), nb = bits_to_indices(bra.b, ob);
  return me_dispatch(h, bra, ket, oa, na, ob, nb);
}

/* ---- fast diagonals (fast_diagonals.hpp:14-127) ---------------------- */
void orc_single_orbital_ens(const orc_ham* h, int64_t norb,
                            const uint32_t* ss_occ, int64_t nss,
                            const uint32_t* os_occ, int64_t nos, double* ens) {
  const int64_t n = h->n;
  for(int64_t i = 0; i < norb; ++i) {
    double e = h->T[i * (n + 1)];
    for(int64_t q = 0; q < nss; ++q)
      e += h->G2_red[i + n * ss_occ[q]] + h->G2_red[ss_occ[q] + n * i];
    e -= h->G2_red[i * (n + 1)];
    for(int64_t q = 0; q < nos; ++q) e += h->V2_red[i + n * os_occ[q]];
    ens[i] = e;
  }
}

#define G2(p, q) (h->G2_red[(p) + h->n * (q)])
#define V2(p, q) (h->V2_red[(p) + h->n * (q)])
double orc_fast_diag_single(const orc_ham* h, double hol_en, double par_en,
                            uint32_t oh, uint32_t op, double Hii) {
  /* :53-60 */
  return Hii + par_en - hol_en - G2(op, oh) - G2(oh, op);
}

double orc_fast_diag_ss_double(const orc_ham* h, double h1, double h2,
                               double p1, double p2, uint32_t oh1,
                               uint32_t oh2, uint32_t op1, uint32_t op2,
                               double Hii) {
  /* :72-85 */
  return Hii + p1 + p2 - h1 - h2 + G2(oh1, oh2) + G2(oh2, oh1) +
         G2(op1, op2) + G2(op2, op1) - G2(op1, oh1) - G2(oh1, op1) -
         G2(op2, oh1) - G2(oh1, op2) - G2(op1, oh2) - G2(oh2, op1) -
         G2(op2, oh2) - G2(oh2, op2);
}

double orc_fast_diag_os_double(const orc_ham* h, double hu, double hd,
                               double pu, double pd, uint32_t ohu,
                               uint32_t ohd, uint32_t opu, uint32_t opd,
                               double Hii) {
  /* :101-112 */
  return Hii + pu + pd - hu - hd + V2(ohu, ohd) + V2(opu, opd) -
         G2(opu, ohu) - G2(ohu, opu) - G2(opd, ohd) - G2(ohd, opd) -
         V2(opu, ohd) - V2(ohu, opd);
}
#undef G2
#undef V2

/* ===================================================================== */
/* determinant spaces (sd_operations.hpp:129-476)                         */
/* ===================================================================== */
/* std::prev_permutation on a 0/1 array */
static int prev_perm(unsigned char* a, int64_t n) {
  if(n < 2) return 0;
  int64_t i = n - 1;
  while(i > 0 && a[i - 1] <= a[i]) --i;
  if(i == 0) {
    for(int64_t l = 0, r = n - 1; l < r; ++l, --r) {
      unsigned char t = a[l];
      a[l] = a[r];
      a[r] = t;
    }
    return 0;
  }
  int64_t j = n - 1;
  while(a[j] >= a[i - 1]) --j;
  unsigned char t = a[i - 1];
  a[i - 1] = a[j];
  a[j] = t;
  for(int64_t l = i, r = n - 1; l < r; ++l, --r) {
    t = a[l];
    a[l] = a[r];
    a[r] = t;
  }
  return 1;
}

/* generate_combs (sd_operations.hpp:440-458) */
static int64_t gen_combs(int64_t nbits, int64_t nset, uint64_t* out) {
  unsigned char v[128];
  for(int64_t i = 0; i < nbits; ++i) v[i] = (unsigned char)(i < nset);
  int64_t c = 0;
  do {
    uint64_t t = 0;
    for(int64_t i = 0; i < nbits; ++i)
      if(v[i]) t |= (1ULL << i);
    if(out) out[c] = t;
    ++c;
  } while(prev_perm(v, nbits));
  return c;
}

int64_t orc_generate_hilbert_space(int W, int64_t norb, int64_t na,
                                   int64_t nb, uint64_t* out) {
  /* :460-476 -- alpha outer, beta inner */
  if(norb > spin_bits(W) || na > norb || nb > norb) return -1;
  int64_t nA = gen_combs(norb, na, NULL), nB = gen_combs(norb, nb, NULL);
  if(!out) return nA * nB;
  uint64_t* A = (uint64_t*)malloc(sizeof(uint64_t) * nA);
  uint64_t* B = (uint64_t*)malloc(sizeof(uint64_t) * nB);
  gen_combs(norb, na, A);
  gen_combs(norb, nb, B);
  int64_t c = 0;
  for(int64_t i = 0; i < nA; ++i)
    for(int64_t j = 0; j < nB; ++j) {
      det_t d = {A[i], B[j]};
      store_det(W, d, out + W * c);
      ++c;
    }
  free(A);
  free(B);
  return c;
}

/* generate_singles_doubles on one spin string (append_singles :129-143,
 * append_doubles :145-165). Returns counts. */
static void spin_singles_doubles(int64_t norb, uint64_t s, uint64_t* sing,
                                 int64_t* ns, uint64_t* dbl, int64_t* nd) {
  uint32_t occ[64], vir[64];
  int no, nv;
  occ_vir(norb, s, occ, &no, vir, &nv);
  int64_t c = 0;
  for(int a = 0; a < nv; ++a)
    for(int i = 0; i < no; ++i) {
      if(sing) sing[c] = s ^ ((1ULL << occ[i]) | (1ULL << vir[a]));
      ++c;
    }
  *ns = c;
  c = 0;
  for(int a = 0; a < nv; ++a)
    for(int i = 0; i < no; ++i)
      for(int b = a + 1; b < nv; ++b)
        for(int j = i + 1; j < no; ++j) {
          if(dbl)
            dbl[c] = s ^ ((1ULL << occ[i]) | (1ULL << vir[a]) |
                          (1ULL << occ[j]) | (1ULL << vir[b]));
          ++c;
        }
  *nd = c;
}

int64_t orc_generate_cisd_hilbert_space(int W, int64_t norb,
                                        const uint64_t* state_p,
                                        uint64_t* out) {
  /* generate_cisd_hilbert_space (:421-430) over
   * generate_singles_doubles_spin (:345-419) */
  det_t st = load_det(W, state_p);
  int64_t nsa, nda, nsb, ndb;
  spin_singles_doubles(norb, st.a, NULL, &nsa, NULL, &nda);
  spin_singles_doubles(norb, st.b, NULL, &nsb, NULL, &ndb);
  int64_t total = 1 + nsa + nsb + nda + ndb + nsa * nsb;
  if(!out) return total;
  uint64_t* sa = (uint64_t*)malloc(sizeof(uint64_t) * (nsa + 1));
  uint64_t* da = (uint64_t*)malloc(sizeof(uint64_t) * (nda + 1));
  uint64_t* sb = (uint64_t*)malloc(sizeof(uint64_t) * (nsb + 1));
  uint64_t* db = (uint64_t*)malloc(sizeof(uint64_t) * (ndb + 1));
  spin_singles_doubles(norb, st.a, sa, &nsa, da, &nda);
  spin_singles_doubles(norb, st.b, sb, &nsb, db, &ndb);
  int64_t c = 0;
  store_det(W, st, out + W * c++);
  for(int64_t i = 0; i < nsa; ++i) {
    det_t d = {sa[i], st.b};
    store_det(W, d, out + W * c++);
  }
  for(int64_t i = 0; i < nsb; ++i) {
    det_t d = {st.a, sb[i]};
    store_det(W, d, out + W * c++);
  }
  for(int64_t i = 0; i < nda; ++i) {
    det_t d = {da[i], st.b};
    store_det(W, d, out + W * c++);
  }
  for(int64_t i = 0; i < ndb; ++i) {
    det_t d = {st.a, db[i]};
    store_det(W, d, out + W * c++);
  }
  for(int64_t i = 0; i < nsa; ++i)
    for(int64_t j = 0; j < nsb; ++j) {
      det_t d = {sa[i], sb[j]};
      store_det(W, d, out + W * c++);
    }
  free(sa);
  free(da);
  free(sb);
  free(db);
  return c;
}

/* ===================================================================== */
/* CSR Hamiltonian (double_loop.hpp:30-98)                                */
/* ===================================================================== */
typedef struct {
  int32_t* col;
  double* val;
  int64_t n, cap;
} entry_buf;

static int buf_push(entry_buf* b, int32_t c, double v) {
  if(b->n == b->cap) {
    int64_t nc = b->cap ? 2 * b->cap : 1024;
    int32_t* c2 = (int32_t*)realloc(b->col, sizeof(int32_t) * nc);
    if(!c2) return ORC_ERR_NOMEM;
    b->col = c2;
    double* v2 = (double*)realloc(b->val, sizeof(double) * nc);
    if(!v2) return ORC_ERR_NOMEM;
    b->val = v2;
    b->cap = nc;
  }
  b->col[b->n] = c;
  b->val[b->n] = v;
  b->n++;
  return ORC_OK;
}

/* One bra row of the double loop (:45-89). Calls sink for each kept
 * entry in ascending ket order; returns the row count. */
static int64_t csr_row(const orc_ham* h, int W, det_t bra, const uint64_t* ket,
                       int64_t nket, double thresh, int32_t* col, double* val,
                       entry_buf* buf, int* err) {
  int64_t nrow = 0;
  if(!(popc(bra.a) + popc(bra.b))) return 0; /* zero bra: empty row (:50) */
  uint32_t oa[64], ob[64];
  int na = bits_to_indices(bra.a, oa), nb = bits_to_indices(bra.b, ob);
  for(int64_t j = 0; j < nket; ++j) {
    det_t k = load_det(W, ket + W * j);
    if(!(popc(k.a) + popc(k.b))) continue; /* zero ket (:62) */
    if(popc(bra.a ^ k.a) + popc(bra.b ^ k.b) > 4) continue; /* :66-67 */
    double h_el = me_dispatch(h, bra, k, oa, na, ob, nb);
    if(fabs(h_el) > thresh) { /* :76-80 */
      if(col) {
        col[nrow] = (int32_t)j;
        val[nrow] = h_el;
      }
      if(buf && buf_push(buf, (int32_t)j, h_el) != ORC_OK) *err = 1;
      ++nrow;
    }
  }
  return nrow;
}

int orc_csr_build(const orc_ham* h, int W, const uint64_t* bra, int64_t nbra,
                  const uint64_t* ket, int64_t nket, double thresh,
                  int64_t** rowptr_out, int32_t** colind_out,
                  double** nzval_out, int64_t* nnz_out) {
  if(!h || (W != 1 && W != 2) || nbra < 0 || nket < 0) return ORC_ERR_ARG;
  int64_t* rowptr = (int64_t*)malloc(sizeof(int64_t) * (nbra + 1));
  if(!rowptr) return ORC_ERR_NOMEM;
  entry_buf buf = {NULL, NULL, 0, 0};
  int err = 0;
  rowptr[0] = 0;
  for(int64_t i = 0; i < nbra; ++i) {
    det_t b = load_det(W, bra + W * i);
    int64_t nrow =
        nket ? csr_row(h, W, b, ket, nket, thresh, NULL, NULL, &buf, &err) : 0;
    rowptr[i + 1] = rowptr[i] + nrow;
    if(err) {
      free(rowptr);
      free(buf.col);
      free(buf.val);
      return ORC_ERR_NOMEM;
    }
  }
  *rowptr_out = rowptr;
  *colind_out = buf.col ? buf.col : (int32_t*)malloc(sizeof(int32_t));
  *nzval_out = buf.val ? buf.val : (double*)malloc(sizeof(double));
  *nnz_out = buf.n;
  return ORC_OK;
}

int orc_csr_rows(const orc_ham* h, int W, const uint64_t* bra,
                 int64_t row_begin, int64_t row_end, const uint64_t* ket,
                 int64_t nket, double thresh, int64_t* row_counts,
                 int32_t* colind, double* nzval) {
  if(!h || (W != 1 && W != 2) || row_end < row_begin) return ORC_ERR_ARG;
  int64_t off = 0;
  int err = 0;
  for(int64_t i = row_begin; i < row_end; ++i) {
    det_t b = load_det(W, bra + W * i);
    int64_t c = csr_row(h, W, b, ket, nket, thresh, colind ? colind + off : NULL,
                        colind ? nzval + off : NULL, NULL, &err);
    if(row_counts) row_counts[i - row_begin] = c;
    off += c;
  }
  return ORC_OK;
}

int orc_csr_rows_parallel(const orc_ham* h, int W, const uint64_t* bra, int64_t row_begin,
                          int64_t row_end, const uint64_t* ket, int64_t nket, double thresh,
                          int64_t* row_counts, int32_t* colind, double* nzval) {
  /* The double loop of double_loop.hpp:45-91 with rows spread over OpenMP
   * threads (not in the reference, whose loop is serial): a counting pass,
   * a prefix sum, then every row filled at its offset. Same CSR as
   * orc_csr_rows. */
  if(!h || (W != 1 && W != 2) || row_end < row_begin || !row_counts) return ORC_ERR_ARG;
  const int64_t n = row_end - row_begin;
#pragma omp parallel for schedule(dynamic, 16)
  for(int64_t i = 0; i < n; ++i) {
    int err = 0;
    det_t b = load_det(W, bra + W * (row_begin + i));
    row_counts[i] = csr_row(h, W, b, ket, nket, thresh, NULL, NULL, NULL, &err);
  }
  if(!colind) return ORC_OK;
  int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
  if(!off) return ORC_ERR_NOMEM;
  off[0] = 0;
  for(int64_t i = 0; i < n; ++i) off[i + 1] = off[i] + row_counts[i];
#pragma omp parallel for schedule(dynamic, 16)
  for(int64_t i = 0; i < n; ++i) {
    int err = 0;
    det_t b = load_det(W, bra + W * (row_begin + i));
    csr_row(h, W, b, ket, nket, thresh, colind + off[i], nzval + off[i], NULL, &err);
  }
  free(off);
  return ORC_OK;
}

/* Connection flags of a block of kets against one bra: flags[j] = 1 when
 * popcount(bra ^ ket_j) <= 4 (the test of double_loop.hpp:66-67, summed over
 * the two spin strings, which for one-word determinants is the popcount of
 * the whole word). An AVX-512 build of the same loop is used when the CPU
 * has vector popcounts; both give the same flags. */
#define VBLK 2048
__attribute__((target("avx512f,avx512bw,avx512vl,avx512vpopcntdq"))) static void
conn_flags_v(int W, uint64_t a, uint64_t b, const uint64_t* k, int64_t n, unsigned char* f) {
  if(W == 1) {
    const uint64_t w = a | (b << 32);
    for(int64_t j = 0; j < n; ++j) f[j] = __builtin_popcountll(w ^ k[j]) <= 4;
  } else {
    for(int64_t j = 0; j < n; ++j)
      f[j] = (__builtin_popcountll(a ^ k[2 * j]) + __builtin_popcountll(b ^ k[2 * j + 1])) <= 4;
  }
}
static void conn_flags_s(int W, uint64_t a, uint64_t b, const uint64_t* k, int64_t n,
                         unsigned char* f) {
  if(W == 1) {
    const uint64_t w = a | (b << 32);
    for(int64_t j = 0; j < n; ++j) f[j] = popc(w ^ k[j]) <= 4;
  } else {
    for(int64_t j = 0; j < n; ++j) f[j] = (popc(a ^ k[2 * j]) + popc(b ^ k[2 * j + 1])) <= 4;
  }
}

/* State of one bra row of the double loop (double_loop.hpp:45-89) while it
 * is compared, entry by entry in ascending ket order, against a given CSR
 * row col[0..len) / val[0..len). */
typedef struct {
  det_t bra;
  uint32_t oa[64], ob[64];
  int na, nb, bad, zero;
  int64_t pos, len;
  const int32_t* col;
  const double* val;
} vrow_t;

/* Kets j0 .. j0+n of the loop for one row: flagged kets go through the
 * loop's zero-ket test, matrix element and threshold as in csr_row, and each
 * kept entry must be the row's next stored one (same column, bitwise equal
 * value). */
static void verify_block(const orc_ham* h, int W, vrow_t* r, const uint64_t* ket, int64_t j0,
                         int64_t n, double thresh, int vec, unsigned char* flags) {
  if(r->bad || r->zero) return;
  if(vec)
    conn_flags_v(W, r->bra.a, r->bra.b, ket + W * j0, n, flags);
  else
    conn_flags_s(W, r->bra.a, r->bra.b, ket + W * j0, n, flags);
  for(int64_t q = 0; q < n; q += 8) {
    uint64_t m8;
    if(q + 8 <= n) {
      memcpy(&m8, flags + q, 8);
      if(!m8) continue;
    }
    const int64_t qe = q + 8 < n ? q + 8 : n;
    for(int64_t jj = q; jj < qe; ++jj) {
      if(!flags[jj]) continue;
      const int64_t j = j0 + jj;
      det_t k = load_det(W, ket + W * j);
      if(!(popc(k.a) + popc(k.b))) continue; /* zero ket (:62) */
      double h_el = me_dispatch(h, r->bra, k, r->oa, r->na, r->ob, r->nb);
      if(fabs(h_el) > thresh) { /* :76-80 */
        if(r->pos >= r->len || r->col[r->pos] != (int32_t)j ||
           memcmp(&r->val[r->pos], &h_el, sizeof(double)) != 0) {
          r->bad = 1;
          return;
        }
        ++r->pos;
      }
    }
  }
}

int64_t orc_csr_verify_rows(const orc_ham* h, int W, const uint64_t* bra,
                            int64_t row_begin, int64_t row_end,
                            const uint64_t* ket, int64_t nket, double thresh,
                            const int64_t* rowptr, const int32_t* colind,
                            const double* nzval, int64_t* first_bad) {
  /* Checks rows [row_begin,row_end) of a CSR (rowptr indexed by absolute
   * row, 0-based, entries colind/nzval at rowptr[i] .. rowptr[i+1]) against
   * the reference double loop without storing the loop's output. Threads
   * take 16 consecutive rows at a time and walk the kets in blocks of VBLK,
   * each block serving all 16 rows from cache; every row still sees the kets
   * in ascending order, as double_loop.hpp:45-89 does. Returns the number
   * of rows that differ; *first_bad receives the lowest such row (-1 if
   * none). */
  if(!h || (W != 1 && W != 2) || row_end < row_begin) return -1;
  enum { RB = 16 };
  const int64_t nrow = row_end - row_begin, ngrp = (nrow + RB - 1) / RB;
  int64_t bad = 0, first = INT64_MAX;
  const int vec = __builtin_cpu_supports("avx512vpopcntdq") && __builtin_cpu_supports("avx512bw") &&
                  __builtin_cpu_supports("avx512vl");
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : bad) reduction(min : first)
  for(int64_t g = 0; g < ngrp; ++g) {
    vrow_t rows[RB];
    unsigned char flags[VBLK];
    const int64_t i0 = row_begin + g * RB;
    const int nr = (int)(row_end - i0 < RB ? row_end - i0 : RB);
    for(int q = 0; q < nr; ++q) {
      vrow_t* r = &rows[q];
      const int64_t i = i0 + q;
      r->bra = load_det(W, bra + W * i);
      r->na = bits_to_indices(r->bra.a, r->oa);
      r->nb = bits_to_indices(r->bra.b, r->ob);
      r->zero = !(r->na + r->nb); /* zero bra: empty row (:50) */
      r->len = rowptr[i + 1] - rowptr[i];
      r->bad = r->len < 0;
      r->pos = 0;
      r->col = colind + rowptr[i];
      r->val = nzval + rowptr[i];
    }
    for(int64_t j0 = 0; j0 < nket; j0 += VBLK) {
      const int64_t n = nket - j0 < VBLK ? nket - j0 : VBLK;
      for(int q = 0; q < nr; ++q) verify_block(h, W, &rows[q], ket, j0, n, thresh, vec, flags);
    }
    for(int q = 0; q < nr; ++q) {
      const vrow_t* r = &rows[q];
      if(r->bad || r->pos != (r->zero ? 0 : r->len) || (r->zero && r->len)) {
        ++bad;
        if(i0 + q < first) first = i0 + q;
      }
    }
  }
  if(first_bad) *first_bad = bad ? first : -1;
  return bad;
}

int orc_num_threads(void) {
  int n = 1;
#pragma omp parallel
  {
#pragma omp single
    n = omp_get_num_threads();
  }
  return n;
}

void orc_free(void* p) { free(p); }

/* ===================================================================== */
/* sparse BLAS                                                            */
/* ===================================================================== */
void orc_gespmbv(int64_t K, double alpha, int64_t M, const int64_t* Arp,
                 const int32_t* Aci, const double* Anz, int64_t indexing,
                 const double* V, int64_t LDV, double beta, double* AV,
                 int64_t LDAV) {
  /* spmbv.hpp:37-73: rows in parallel (the reference's omp parallel for,
   * spmbv.hpp:54), each row a serial dot in CSR order, so the result does
   * not depend on the thread count */
  for(int64_t k = 0; k < K; ++k)
#pragma omp parallel for schedule(static)
    for(int64_t i = 0; i < M; ++i) {
      const int64_t j_st = Arp[i] - indexing, j_en = Arp[i + 1] - indexing;
      const double* V_k = V + k * LDV - indexing;
      double av = 0.;
      for(int64_t j = j_st; j < j_en; ++j) av += Anz[j] * V_k[Aci[j]];
      AV[i + k * LDAV] = alpha * av + beta * AV[i + k * LDAV];
    }
}

void orc_extract_diagonal(int64_t M, const int64_t* Arp, const int32_t* Aci,
                          const double* Anz, int64_t indexing, double* D) {
  /* submatrix.hpp:264-295 -- first column equal to i, else 0 */
  for(int64_t i = 0; i < M; ++i) {
    const int64_t j_st = Arp[i] - indexing, j_en = Arp[i + 1] - indexing;
    D[i] = 0.;
    for(int64_t j = j_st; j < j_en; ++j)
      if(Aci[j] == i + indexing) {
        D[i] = Anz[j];
        break;
      }
  }
}

/* ===================================================================== */
/* symmetric eigensolver: cyclic Jacobi (Golub & Van Loan, Alg. 8.5.3)     */
/* ===================================================================== */
void orc_syev_lower(int64_t n, double* A, int64_t lda, double* W) {
  double* a = (double*)malloc(sizeof(double) * n * n);
  double* v = (double*)calloc(n * n, sizeof(double));
  for(int64_t j = 0; j < n; ++j)
    for(int64_t i = j; i < n; ++i) {
      a[i + n * j] = A[i + lda * j];
      a[j + n * i] = A[i + lda * j];
    }
  for(int64_t i = 0; i < n; ++i) v[i + n * i] = 1.;
  for(int sweep = 0; sweep < 100; ++sweep) {
    double off = 0., tot = 0.;
    for(int64_t j = 0; j < n; ++j)
      for(int64_t i = 0; i < n; ++i) {
        tot += a[i + n * j] * a[i + n * j];
        if(i != j) off += a[i + n * j] * a[i + n * j];
      }
    if(off <= 1e-32 * tot || off == 0.) break;
    for(int64_t p = 0; p < n - 1; ++p)
      for(int64_t q = p + 1; q < n; ++q) {
        const double apq = a[p + n * q];
        if(apq == 0.) continue;
        { /* after four sweeps an element below the diagonals' rounding is
           * zeroed instead of rotated (Numerical Recipes, jacobi) */
          const double g = 100. * fabs(apq), app = fabs(a[p + n * p]), aqq = fabs(a[q + n * q]);
          if(sweep > 3 && app + g == app && aqq + g == aqq) {
            a[p + n * q] = a[q + n * p] = 0.;
            continue;
          }
        }
        const double tau = (a[q + n * q] - a[p + n * p]) / (2. * apq);
        const double t = (tau >= 0. ? 1. : -1.) / (fabs(tau) + sqrt(1. + tau * tau));
        const double c = 1. / sqrt(1. + t * t), s = t * c;
        for(int64_t k = 0; k < n; ++k) { /* A <- A J (columns p,q) */
          const double akp = a[k + n * p], akq = a[k + n * q];
          a[k + n * p] = c * akp - s * akq;
          a[k + n * q] = s * akp + c * akq;
        }
        for(int64_t k = 0; k < n; ++k) { /* A <- J^T A (rows p,q) */
          const double apk = a[p + n * k], aqk = a[q + n * k];
          a[p + n * k] = c * apk - s * aqk;
          a[q + n * k] = s * apk + c * aqk;
        }
        for(int64_t k = 0; k < n; ++k) { /* V <- V J */
          const double vkp = v[k + n * p], vkq = v[k + n * q];
          v[k + n * p] = c * vkp - s * vkq;
          v[k + n * q] = s * vkp + c * vkq;
        }
      }
  }
  /* sort ascending (selection sort; n <= a few dozen) */
  int64_t* ord = (int64_t*)malloc(sizeof(int64_t) * n);
  for(int64_t i = 0; i < n; ++i) ord[i] = i;
  for(int64_t i = 0; i < n; ++i)
    for(int64_t j = i + 1; j < n; ++j)
      if(a[ord[j] * (n + 1)] < a[ord[i] * (n + 1)]) {
        int64_t t = ord[i];
        ord[i] = ord[j];
        ord[j] = t;
      }
  for(int64_t j = 0; j < n; ++j) {
    W[j] = a[ord[j] * (n + 1)];
    for(int64_t i = 0; i < n; ++i) A[i + lda * j] = v[i + n * ord[j]];
  }
  free(ord);
  free(a);
  free(v);
}

/* ===================================================================== */
/* Davidson (davidson.hpp:61-225)                                         */
/* ===================================================================== */
void orc_diagonal_guess(int64_t N, const double* D, double* X) {
  /* :61-68 -- unit vector at the first minimum of D */
  int64_t mi = 0;
  for(int64_t i = 1; i < N; ++i)
    if(D[i] < D[mi]) mi = i;
  if(N > 0) X[mi] = 1.;
}

/* gram_schmidt (:111-126): CGS2 then normalise */
static void gram_schmidt(int64_t N, int64_t K, const double* V, int64_t LDV,
                         double* w) {
  double inner[1025]; /* K <= max_m + 1 <= 1025 */
  for(int pass = 0; pass < 2; ++pass) {
    for(int64_t j = 0; j < K; ++j) {
      double s = 0.;
      for(int64_t n = 0; n < N; ++n) s += V[n + j * LDV] * w[n];
      inner[j] = s;
    }
    for(int64_t n = 0; n < N; ++n) {
      double s = 0.;
      for(int64_t j = 0; j < K; ++j) s += V[n + j * LDV] * inner[j];
      w[n] = -1. * s + w[n];
    }
  }
  double nrm = 0.;
  for(int64_t n = 0; n < N; ++n) nrm += w[n] * w[n];
  nrm = sqrt(nrm);
  const double sc = 1. / nrm;
  for(int64_t n = 0; n < N; ++n) w[n] *= sc;
}

/* Coefficients of the restart basis of the restarted Davidson variant
 * (k x nb, column-major): q0 = c/|c| and, for nb == 2, q1 = the previous
 * Ritz coefficients (c_prev, 0) orthogonalised against q0 by two classical
 * Gram-Schmidt passes and normalised, or c2 (the second Ritz vector's
 * coefficients) treated the same way when c_prev is absent or its
 * component orthogonal to q0 has norm <= 1e-10. */
static void restart_coefficients(int64_t k, int64_t nb, const double* c, const double* cprev,
                                 const double* c2, double* Q) {
  double cn = 0.;
  for(int64_t j = 0; j < k; ++j) cn += c[j] * c[j];
  cn = 1. / sqrt(cn);
  for(int64_t j = 0; j < k; ++j) Q[j] = c[j] * cn;
  if(nb < 2) return;
  double* q1 = Q + k;
  for(int attempt = cprev ? 0 : 1; attempt < 2; ++attempt) {
    for(int64_t j = 0; j < k; ++j) q1[j] = attempt ? c2[j] : (j < k - 1 ? cprev[j] : 0.);
    for(int pass = 0; pass < 2; ++pass) {
      double s = 0.;
      for(int64_t j = 0; j < k; ++j) s += Q[j] * q1[j];
      for(int64_t j = 0; j < k; ++j) q1[j] -= s * Q[j];
    }
    double pn = 0.;
    for(int64_t j = 0; j < k; ++j) pn += q1[j] * q1[j];
    pn = sqrt(pn);
    if(pn > 1e-10) {
      for(int64_t j = 0; j < k; ++j) q1[j] /= pn;
      return;
    }
  }
}

int orc_davidson_restarted(int64_t N, int64_t max_m, int64_t max_restarts,
                           const int64_t* rp, const int32_t* ci,
                           const double* nz, const double* D, double tol,
                           double* X, double* E, int64_t* iters,
                           int64_t* restarts_out) {
  /* davidson (davidson.hpp:128-225). With max_restarts == 0 this is the
   * reference algorithm exactly: no restart, error when the subspace of
   * max_m vectors is exhausted. With max_restarts > 0 (a labelled variant,
   * not in the reference) an exhausted subspace restarts from
   * [X, Xp, t]: X the current Ritz vector, Xp the previous iteration's Ritz
   * vector orthogonalised against X (see restart_coefficients; dropped when
   * max_m < 3) and t the preconditioned, orthogonalised residual that would
   * have been the next basis vector. A*X and A*Xp are the same combinations
   * of the stored A*V columns (no product with A). */
  if(!X) return ORC_ERR_NO_GUESS;
  if(max_m > N) max_m = N;
  if(max_m > 1024) return ORC_ERR_ARG;
  const int64_t ld = max_m + 1;
  double* V = (double*)calloc((size_t)N * ld, sizeof(double));
  double* AV = (double*)calloc((size_t)N * ld, sizeof(double));
  double* tmp = (double*)calloc((size_t)N * 4, sizeof(double));  /* restart staging */
  /* k x k Ritz problem, eigenvalues, current / previous / second Ritz
   * coefficients and the two restart columns, for k <= max_m + 1 */
  double* small = (double*)calloc((size_t)ld * ld + 6 * (size_t)ld, sizeof(double));
  if(!V || !AV || !tmp || !small) {
    free(V);
    free(AV);
    free(tmp);
    free(small);
    return ORC_ERR_NOMEM;
  }
  double *C = small, *LAM = C + ld * ld, *ccur = LAM + ld, *cprev = ccur + ld,
         *c2cur = cprev + ld, *Q = c2cur + ld;
  memcpy(V, X, sizeof(double) * N);
  orc_gespmbv(1, 1., N, rp, ci, nz, 0, V, N, 0., AV, N); /* :153 */
  memcpy(V + N, AV, sizeof(double) * N);                   /* :156 */
  gram_schmidt(N, 1, V, N, V + N);                          /* :157 */
  int converged = 0;
  int64_t iter = 1, restarts = 0, i0 = 1, kcur = 0;
  int prev_valid = 0;
  for(;;) {
    for(int64_t i = i0; i < max_m; ++i, ++iter) {
      const int64_t k = i + 1;
      orc_gespmbv(1, 1., N, rp, ci, nz, 0, V + i * N, N, 0., AV + i * N, N);
      /* rayleigh_ritz (rayleigh_ritz.hpp:70-77): C = V^T AV, syev(Lower) */
      for(int64_t b = 0; b < k; ++b)
        for(int64_t a = 0; a < k; ++a) {
          double sum = 0.;
          for(int64_t n = 0; n < N; ++n) sum += V[n + a * N] * AV[n + b * N];
          C[a + k * b] = sum;
        }
      orc_syev_lower(k, C, k, LAM);
      memcpy(cprev, ccur, sizeof(double) * ld);
      prev_valid = i > i0; /* c_prev holds coefficients in this basis */
      for(int64_t j = 0; j < k; ++j) {
        ccur[j] = C[j];
        c2cur[j] = k > 1 ? C[j + k] : 0.;
      }
      kcur = k;
      double* R = V + (i + 1) * N;
      /* X = V*C(:,0); R = AV*C(:,0) - LAM0*X; ||R||   (:181-193) */
      double rn = 0.;
      for(int64_t n = 0; n < N; ++n) {
        double x = 0., r = 0.;
        for(int64_t j = 0; j < k; ++j) x += V[n + j * N] * C[j];
        for(int64_t j = 0; j < k; ++j) r += AV[n + j * N] * C[j];
        X[n] = x;
        R[n] = r + (-LAM[0]) * x;
        rn += R[n] * R[n];
      }
      rn = sqrt(rn);
      if(rn < tol) {
        converged = 1;
        break;
      }
      for(int64_t j = 0; j < N; ++j) R[j] = -R[j] / (D[j] - LAM[0]); /* :212 */
      gram_schmidt(N, k, V, N, R);                                    /* :217 */
    }
    if(converged || restarts >= max_restarts || max_m < 2) break;
    ++restarts;
    /* restart basis V*Q, Q = [q0, q1] orthonormal in the coefficient space:
     * q0 = c, q1 = (c_prev, 0) made orthogonal to q0 (two Gram-Schmidt
     * passes); the second Ritz vector's coefficients c2 stand in for c_prev
     * when the last two iterations did not share this basis or c_prev is
     * numerically parallel to c. One column only when max_m < 3. */
    const int64_t nb = max_m >= 3 ? 2 : 1;
    restart_coefficients(kcur, nb, ccur, prev_valid ? cprev : NULL, c2cur, Q);
    for(int64_t n = 0; n < N; ++n) {
      double v[2] = {0., 0.}, av[2] = {0., 0.};
      for(int64_t c = 0; c < nb; ++c)
        for(int64_t j = 0; j < kcur; ++j) {
          v[c] += V[n + j * N] * Q[j + c * kcur];
          av[c] += AV[n + j * N] * Q[j + c * kcur];
        }
      tmp[n] = v[0];
      tmp[n + N] = v[1];
      tmp[n + 2 * N] = av[0];
      tmp[n + 3 * N] = av[1];
    }
    for(int64_t c = 0; c < nb; ++c) {
      memcpy(V + c * N, tmp + c * N, sizeof(double) * N);
      memcpy(AV + c * N, tmp + (2 + c) * N, sizeof(double) * N);
    }
    memmove(V + nb * N, V + max_m * N, sizeof(double) * N);
    i0 = nb;
  }
  free(V);
  free(AV);
  free(tmp);
  *E = LAM[0];
  free(small);
  if(iters) *iters = iter;
  if(restarts_out) *restarts_out = restarts;
  return converged ? ORC_OK : ORC_ERR_NOT_CONVERGED;
}

int orc_davidson(int64_t N, int64_t max_m, const int64_t* rp,
                 const int32_t* ci, const double* nz, const double* D,
                 double tol, double* X, double* E, int64_t* iters) {
  return orc_davidson_restarted(N, max_m, 0, rp, ci, nz, D, tol, X, E, iters,
                                NULL);
}

int orc_selected_ci_diag_restarted(const orc_ham* h, int W,
                                   const uint64_t* dets, int64_t ndets,
                                   double h_el_tol, int64_t max_m,
                                   int64_t max_restarts, double res_tol,
                                   double* C, double* E, int64_t* nnz_out) {
  /* selected_ci_diag (:124-207) -> serial_selected_ci_diag (:75-122) */
  int64_t *rp = NULL, nnz = 0;
  int32_t* ci = NULL;
  double* nz = NULL;
  int rc = orc_csr_build(h, W, dets, ndets, dets, ndets, h_el_tol, &rp, &ci,
                         &nz, &nnz);
  if(rc) return rc;
  if(nnz_out) *nnz_out = nnz;
  double* D = (double*)malloc(sizeof(double) * (ndets + 1));
  orc_extract_diagonal(ndets, rp, ci, nz, rp[0], D);
  double max_c = 0.;
  for(int64_t i = 0; i < ndets; ++i)
    if(fabs(C[i]) > max_c) max_c = fabs(C[i]);
  if(!(max_c > 1. / (double)ndets)) orc_diagonal_guess(ndets, D, C); /* :99 */
  int64_t it = 0;
  rc = orc_davidson_restarted(ndets, max_m, max_restarts, rp, ci, nz, D,
                              res_tol, C, E, &it, NULL);
  free(D);
  free(rp);
  free(ci);
  free(nz);
  return rc;
}

int orc_selected_ci_diag(const orc_ham* h, int W, const uint64_t* dets,
                         int64_t ndets, double h_el_tol, int64_t max_m,
                         double res_tol, double* C, double* E,
                         int64_t* nnz_out) {
  return orc_selected_ci_diag_restarted(h, W, dets, ndets, h_el_tol, max_m, 0,
                                        res_tol, C, E, nnz_out);
}

/* ===================================================================== */
/* ASCI                                                                   */
/* ===================================================================== */
void orc_asci_default_settings(orc_asci_settings* s) {
  /* ASCISettings (determinant_search.hpp:34-51) */
  s->ntdets_max = 100000;
  s->ntdets_min = 100;
  s->ncdets_max = 100;
  s->h_el_tol = 1e-8;
  s->rv_prune_tol = 1e-8;
  s->pair_size_max = 500000000;
  s->just_singles = 0;
  s->grow_factor = 8;
  s->max_refine_iter = 6;
  s->refine_energy_tol = 1e-6;
}

typedef struct {
  det_t s;
  double rv;
  int64_t idx; /* generation order: makes every sort deterministic */
} pair_t;

typedef struct {
  pair_t* p;
  int64_t n, cap;
} pair_buf;

static int pb_push(pair_buf* b, det_t s, double rv) {
  if(b->n == b->cap) {
    int64_t nc = b->cap ? 2 * b->cap : 4096;
    pair_t* q = (pair_t*)realloc(b->p, sizeof(pair_t) * nc);
    if(!q) return ORC_ERR_NOMEM;
    b->p = q;
    b->cap = nc;
  }
  b->p[b->n].s = s;
  b->p[b->n].rv = rv;
  b->p[b->n].idx = b->n;
  b->n++;
  return ORC_OK;
}

/* bitset_less (bitset_operations.hpp:377-395): numeric order of the whole
 * word, i.e. beta string most significant. */
static inline int det_cmp(det_t x, det_t y) {
  if(x.b != y.b) return x.b < y.b ? -1 : 1;
  if(x.a != y.a) return x.a < y.a ? -1 : 1;
  return 0;
}
static int cmp_state_idx(const void* x_, const void* y_) {
  const pair_t *x = (const pair_t*)x_, *y = (const pair_t*)y_;
  int c = det_cmp(x->s, y->s);
  if(c) return c;
  return (x->idx > y->idx) - (x->idx < y->idx);
}
static int cmp_absrv_desc(const void* x_, const void* y_) {
  const pair_t *x = (const pair_t*)x_, *y = (const pair_t*)y_;
  double ax = fabs(x->rv), ay = fabs(y->rv);
  if(ax != ay) return ax > ay ? -1 : 1;
  return det_cmp(x->s, y->s);
}

/* sort_and_accumulate_asci_pairs (determinant_sort.hpp:39-83): sort by
 * determinant (generation order among equals), sum rv into the first copy,
 * drop duplicates. */
static void sort_and_accumulate(pair_buf* b) {
  if(!b->n) return;
  for(int64_t i = 0; i < b->n; ++i) b->p[i].idx = i;
  qsort(b->p, b->n, sizeof(pair_t), cmp_state_idx);
  int64_t w = 0;
  for(int64_t i = 1; i < b->n; ++i) {
    if(det_cmp(b->p[i].s, b->p[w].s) != 0)
      b->p[++w] = b->p[i];
    else
      b->p[w].rv += b->p[i].rv;
  }
  b->n = w + 1;
}

/* keep_only_largest_copy_asci_pairs (determinant_sort.hpp:85-119) */
static void keep_largest(pair_buf* b) {
  if(!b->n) return;
  for(int64_t i = 0; i < b->n; ++i) b->p[i].idx = i;
  qsort(b->p, b->n, sizeof(pair_t), cmp_state_idx);
  int64_t w = 0;
  for(int64_t i = 1; i < b->n; ++i) {
    if(det_cmp(b->p[i].s, b->p[w].s) != 0)
      b->p[++w] = b->p[i];
    else if(b->p[i].rv > b->p[w].rv)
      b->p[w].rv = b->p[i].rv;
  }
  b->n = w + 1;
}

/* Per-core-det data of the candidate formulas. */
typedef struct {
  det_t s;
  uint32_t oa[64], va[64], ob[64], vb[64];
  int na, nva, nb, nvb;
  double *eps_a, *eps_b;
  double h_diag;
} core_t;

static void core_load(const orc_ham* h, int W, const uint64_t* det, int64_t norb, core_t* k) {
  k->s = load_det(W, det);
  occ_vir(norb, k->s.a, k->oa, &k->na, k->va, &k->nva);
  occ_vir(norb, k->s.b, k->ob, &k->nb, k->vb, &k->nvb);
  orc_single_orbital_ens(h, norb, k->oa, k->na, k->ob, k->nb, k->eps_a);
  orc_single_orbital_ens(h, norb, k->ob, k->nb, k->oa, k->na, k->eps_b);
  k->h_diag = me_diag(h, k->oa, k->na, k->ob, k->nb);
}

/* One single i -> a of spin `spin` (append_singles_asci_contributions,
 * determinant_contributions.hpp:25-66): returns 0 when screened out by
 * |h| < tol (before the sign, :47), else the excited det and
 * val = sign * h / (E0 - H_kk^fast). */
static int cand_single(const orc_ham* h, const core_t* k, int spin, uint32_t i, uint32_t a,
                       double E0, double tol, det_t* ex, double* val) {
  const int64_t n = h->n, n2 = n * n;
  const uint64_t ss = spin ? k->s.b : k->s.a;
  const uint32_t* occ = spin ? k->ob : k->oa;
  const uint32_t* oth = spin ? k->oa : k->ob;
  const int no = spin ? k->nb : k->na, nt = spin ? k->na : k->nb;
  const double* eps = spin ? k->eps_b : k->eps_a;
  double h_el = h->T[a + i * n];
  const double* G_ov = h->G_red + a * n + i * n2;
  const double* V_ov = h->V_red + a * n + i * n2;
  for(int p = 0; p < no; ++p) h_el += G_ov[occ[p]];
  for(int p = 0; p < nt; ++p) h_el += V_ov[oth[p]];
  if(fabs(h_el) < tol) return 0;
  *ex = k->s;
  if(spin)
    ex->b ^= (1ULL << i) | (1ULL << a);
  else
    ex->a ^= (1ULL << i) | (1ULL << a);
  h_el *= orc_single_excitation_sign(ss, a, i);
  const double hd = orc_fast_diag_single(h, eps[i], eps[a], i, a, k->h_diag);
  *val = h_el / (E0 - hd);
  return 1;
}

/* One same-spin double (i, j) -> (a, b) (append_ss_doubles_asci_contributions,
 * :68-120). */
static int cand_ss(const orc_ham* h, const core_t* k, int spin, uint32_t i, uint32_t j,
                   uint32_t a, uint32_t b, double E0, double tol, det_t* ex, double* val) {
  const int64_t n = h->n, n2 = n * n;
  const uint64_t ss = spin ? k->s.b : k->s.a;
  const double* eps = spin ? k->eps_b : k->eps_a;
  const double G_aibj = h->G[(a + i * n) * n2 + b + j * n];
  if(fabs(G_aibj) < tol) return 0;
  const uint64_t fex = (1ULL << i) | (1ULL << j) | (1ULL << a) | (1ULL << b);
  uint32_t o1, v1, o2, v2;
  const double sign = orc_doubles_sign_indices(ss, ss ^ fex, fex, &o1, &v1, &o2, &v2);
  *ex = k->s;
  if(spin)
    ex->b ^= fex;
  else
    ex->a ^= fex;
  const double h_el = sign * G_aibj;
  const double hd =
      orc_fast_diag_ss_double(h, eps[i], eps[j], eps[a], eps[b], i, j, a, b, k->h_diag);
  *val = h_el / (E0 - hd);
  return 1;
}

/* One opposite-spin double alpha i -> a, beta j -> b
 * (append_os_doubles_asci_contributions, :122-160). */
static int cand_os(const orc_ham* h, const core_t* k, uint32_t i, uint32_t a, uint32_t j,
                   uint32_t b, double E0, double tol, det_t* ex, double* val) {
  const int64_t n = h->n, n2 = n * n;
  const double V_aibj = h->V[a + i * n + (b + j * n) * n2];
  if(fabs(V_aibj) < tol) return 0;
  const double sign =
      orc_single_excitation_sign(k->s.a, a, i) * orc_single_excitation_sign(k->s.b, b, j);
  *ex = k->s;
  ex->a ^= (1ULL << a) | (1ULL << i);
  ex->b ^= (1ULL << j) | (1ULL << b);
  const double h_el = sign * V_aibj;
  const double hd = orc_fast_diag_os_double(h, k->eps_a[i], k->eps_b[j], k->eps_a[a],
                                            k->eps_b[b], i, j, a, b, k->h_diag);
  *val = h_el / (E0 - hd);
  return 1;
}

/* asci_contributions_standard (determinant_search.hpp:53-141) with the
 * append_* kernels of determinant_contributions.hpp:25-160, in their loop
 * order. When `prune` is set, applies the pair_size_max pruning of
 * :120-137. */
static int asci_contribs(const orc_ham* h, int W, const orc_asci_settings* st,
                         const uint64_t* cdets, int64_t ncdets, double E0,
                         const double* C, int64_t norb, pair_buf* out,
                         int prune) {
  const double tol = st->h_el_tol;
  core_t k;
  k.eps_a = (double*)malloc(sizeof(double) * norb);
  k.eps_b = (double*)malloc(sizeof(double) * norb);
  det_t ex;
  double v;
  for(int64_t c = 0; c < ncdets; ++c) {
    core_load(h, W, cdets + W * c, norb, &k);
    const double coeff = C[c];
    /* singles AA then BB */
    for(int spin = 0; spin < 2; ++spin) {
      const uint32_t* occ = spin ? k.ob : k.oa;
      const uint32_t* vir = spin ? k.vb : k.va;
      const int no = spin ? k.nb : k.na, nv = spin ? k.nvb : k.nva;
      for(int ii = 0; ii < no; ++ii)
        for(int aa = 0; aa < nv; ++aa)
          if(cand_single(h, &k, spin, occ[ii], vir[aa], E0, tol, &ex, &v))
            if(pb_push(out, ex, coeff * v)) goto nomem;
    }
    if(!st->just_singles) {
      /* same-spin doubles AAAA then BBBB */
      for(int spin = 0; spin < 2; ++spin) {
        const uint32_t* occ = spin ? k.ob : k.oa;
        const uint32_t* vir = spin ? k.vb : k.va;
        const int no = spin ? k.nb : k.na, nv = spin ? k.nvb : k.nva;
        for(int ii = 0; ii < no; ++ii)
          for(int aa = 0; aa < nv; ++aa)
            for(int jj = ii + 1; jj < no; ++jj)
              for(int bb = aa + 1; bb < nv; ++bb)
                if(cand_ss(h, &k, spin, occ[ii], occ[jj], vir[aa], vir[bb], E0, tol, &ex, &v))
                  if(pb_push(out, ex, coeff * v)) goto nomem;
      }
      /* opposite-spin doubles AABB */
      for(int ii = 0; ii < k.na; ++ii)
        for(int aa = 0; aa < k.nva; ++aa)
          for(int jj = 0; jj < k.nb; ++jj)
            for(int bb = 0; bb < k.nvb; ++bb)
              if(cand_os(h, &k, k.oa[ii], k.va[aa], k.ob[jj], k.vb[bb], E0, tol, &ex, &v))
                if(pb_push(out, ex, coeff * v)) goto nomem;
    }
    if(prune && out->n > st->pair_size_max) { /* :120-137 */
      int64_t w = 0;
      for(int64_t i = 0; i < out->n; ++i)
        if(fabs(out->p[i].rv) > st->rv_prune_tol) out->p[w++] = out->p[i];
      out->n = w;
      if(out->n > st->pair_size_max) sort_and_accumulate(out);
      for(int64_t i = 0; i < out->n; ++i) out->p[i].idx = i;
    }
  }
  free(k.eps_a);
  free(k.eps_b);
  return ORC_OK;
nomem:
  free(k.eps_a);
  free(k.eps_b);
  return ORC_ERR_NOMEM;
}

int64_t orc_asci_contributions(const orc_ham* h, int W,
                               const orc_asci_settings* s,
                               const uint64_t* cdets, int64_t ncdets,
                               double E_ASCI, const double* C, int64_t norb,
                               uint64_t* states, double* rv) {
  pair_buf b = {NULL, 0, 0};
  if(asci_contribs(h, W, s, cdets, ncdets, E_ASCI, C, norb, &b, 0)) {
    free(b.p);
    return ORC_ERR_NOMEM;
  }
  if(states)
    for(int64_t i = 0; i < b.n; ++i) {
      store_det(W, b.p[i].s, states + W * i);
      rv[i] = b.p[i].rv;
    }
  int64_t n = b.n;
  free(b.p);
  return n;
}

/* The end of asci_search after the contributions (determinant_search.hpp:
 * 498-653): accumulate, rv <- -|rv|, core dets appended with +|C|, keep the
 * largest copy, keep rv < 0, top (ndets_max - ncdets) by |rv|, core last. */
static int asci_finish(int W, pair_buf* b, int64_t ndets_max, const uint64_t* cdets,
                       int64_t ncdets, const double* C, uint64_t* new_dets, int64_t* n_new,
                       double* rv_out) {
  sort_and_accumulate(b);                                     /* :498 */
  for(int64_t i = 0; i < b->n; ++i) b->p[i].rv = -fabs(b->p[i].rv); /* :530 */
  for(int64_t i = 0; i < ncdets; ++i)                          /* :533-536 */
    if(pb_push(b, load_det(W, cdets + W * i), fabs(C[i]))) return ORC_ERR_NOMEM;
  keep_largest(b); /* :540 */
  {                /* keep rv < 0 (:542-544) */
    int64_t w = 0;
    for(int64_t i = 0; i < b->n; ++i)
      if(b->p[i].rv < 0.0) b->p[w++] = b->p[i];
    b->n = w;
  }
  const int64_t top_k = ndets_max - ncdets; /* :547 */
  if(b->n > top_k) {                        /* :565-629 */
    qsort(b->p, b->n, sizeof(pair_t), cmp_absrv_desc);
    b->n = top_k;
  }
  for(int64_t i = 0; i < b->n; ++i) {
    store_det(W, b->p[i].s, new_dets + W * i);
    if(rv_out) rv_out[i] = b->p[i].rv;
  }
  memcpy(new_dets + W * b->n, cdets, sizeof(uint64_t) * W * ncdets); /* :653 */
  *n_new = b->n + ncdets;
  return ORC_OK;
}

int orc_asci_search(const orc_ham* h, int W, const orc_asci_settings* s,
                    int64_t ndets_max, const uint64_t* cdets, int64_t ncdets,
                    double E_ASCI, const double* C, int64_t norb,
                    uint64_t* new_dets, int64_t* n_new, double* rv_out) {
  /* asci_search (determinant_search.hpp:394-663), world_size == 1 */
  pair_buf b = {NULL, 0, 0};
  int rc = asci_contribs(h, W, s, cdets, ncdets, E_ASCI, C, norb, &b, 1);
  if(!rc) rc = asci_finish(W, &b, ndets_max, cdets, ncdets, C, new_dets, n_new, rv_out);
  free(b.p);
  return rc;
}

/* ---- excitation blocks (the block screen; no reference counterpart) ----
 * The candidates of a core det, grouped by hole set: [alpha single holes |
 * beta single holes | alpha-alpha hole pairs | beta-beta hole pairs |
 * alpha-beta hole pairs (alpha hole major)]; pairs (x < y) of the occupied
 * lists in lexicographic order; particles in the loop order of the
 * append_* kernels. */
static int64_t block_total(const core_t* k, int singles) {
  int64_t n = k->na + k->nb;
  if(!singles)
    n += (int64_t)k->na * (k->na - 1) / 2 + (int64_t)k->nb * (k->nb - 1) / 2 +
         (int64_t)k->na * k->nb;
  return n;
}

static void pair_of(int64_t q, int m, int* x, int* y) {
  int a = 0;
  while(q >= m - 1 - a) {
    q -= m - 1 - a;
    ++a;
  }
  *x = a;
  *y = a + 1 + (int)q;
}

/* Walks block q of core k: the block's determinant (holes cleared), the
 * largest |h / (E0 - H_kk)| of its screened-in candidates (-1 if none), and,
 * with `out`, pushes (candidate, coeff * h / (E0 - H_kk)) in block order. */
static int block_walk(const orc_ham* h, const core_t* k, int64_t q, double E0, double tol,
                      double coeff, pair_buf* out, double* maxabs, det_t* bdet) {
  det_t ex;
  double v, m = -1.0;
  *bdet = k->s;
  int x, y;
#define TAKE                                          \
  do {                                                \
    if(fabs(v) > m) m = fabs(v);                      \
    if(out && pb_push(out, ex, coeff * v)) return ORC_ERR_NOMEM; \
  } while(0)
  if(q < k->na) {
    bdet->a &= ~(1ULL << k->oa[q]);
    for(int aa = 0; aa < k->nva; ++aa)
      if(cand_single(h, k, 0, k->oa[q], k->va[aa], E0, tol, &ex, &v)) TAKE;
    *maxabs = m;
    return ORC_OK;
  }
  q -= k->na;
  if(q < k->nb) {
    bdet->b &= ~(1ULL << k->ob[q]);
    for(int bb = 0; bb < k->nvb; ++bb)
      if(cand_single(h, k, 1, k->ob[q], k->vb[bb], E0, tol, &ex, &v)) TAKE;
    *maxabs = m;
    return ORC_OK;
  }
  q -= k->nb;
  for(int spin = 0; spin < 2; ++spin) {
    const int no = spin ? k->nb : k->na, nv = spin ? k->nvb : k->nva;
    const uint32_t* occ = spin ? k->ob : k->oa;
    const uint32_t* vir = spin ? k->vb : k->va;
    const int64_t np = (int64_t)no * (no - 1) / 2;
    if(q < np) {
      pair_of(q, no, &x, &y);
      const uint64_t hb = (1ULL << occ[x]) | (1ULL << occ[y]);
      if(spin)
        bdet->b &= ~hb;
      else
        bdet->a &= ~hb;
      for(int aa = 0; aa < nv; ++aa)
        for(int bb = aa + 1; bb < nv; ++bb)
          if(cand_ss(h, k, spin, occ[x], occ[y], vir[aa], vir[bb], E0, tol, &ex, &v)) TAKE;
      *maxabs = m;
      return ORC_OK;
    }
    q -= np;
  }
  {
    const int ii = (int)(q / k->nb), jj = (int)(q % k->nb);
    bdet->a &= ~(1ULL << k->oa[ii]);
    bdet->b &= ~(1ULL << k->ob[jj]);
    for(int aa = 0; aa < k->nva; ++aa)
      for(int bb = 0; bb < k->nvb; ++bb)
        if(cand_os(h, k, k->oa[ii], k->va[aa], k->ob[jj], k->vb[bb], E0, tol, &ex, &v)) TAKE;
  }
#undef TAKE
  *maxabs = m;
  return ORC_OK;
}

int64_t orc_asci_blocks(const orc_ham* h, int W, const orc_asci_settings* s,
                        const uint64_t* cdets, int64_t ncdets, double E_ASCI, int64_t norb,
                        uint64_t* block_dets, double* target) {
  core_t k;
  k.eps_a = (double*)malloc(sizeof(double) * norb);
  k.eps_b = (double*)malloc(sizeof(double) * norb);
  int64_t g = 0;
  for(int64_t c = 0; c < ncdets; ++c) {
    core_load(h, W, cdets + W * c, norb, &k);
    const int64_t nb = block_total(&k, s->just_singles);
    if(block_dets || target)
      for(int64_t q = 0; q < nb; ++q) {
        double m;
        det_t bd;
        block_walk(h, &k, q, E_ASCI, s->h_el_tol, 1.0, NULL, &m, &bd);
        if(block_dets) store_det(W, bd, block_dets + W * (g + q));
        if(target) target[g + q] = m < 0.0 ? -INFINITY : log(m);
      }
    g += nb;
  }
  free(k.eps_a);
  free(k.eps_b);
  return g;
}

typedef struct {
  float s;
  int64_t i;
} bscore_t;
static int cmp_bscore(const void* x_, const void* y_) {
  const bscore_t *x = (const bscore_t*)x_, *y = (const bscore_t*)y_;
  if(x->s != y->s) return x->s > y->s ? -1 : 1;
  return (x->i > y->i) - (x->i < y->i);
}

int orc_asci_search_blocks(const orc_ham* h, int W, const orc_asci_settings* s,
                           const float* scores, double keep_fraction, int64_t ndets_max,
                           const uint64_t* cdets, int64_t ncdets, double E_ASCI, const double* C,
                           int64_t norb, uint64_t* new_dets, int64_t* n_new, double* rv_out) {
  const int64_t nblk = orc_asci_blocks(h, W, s, cdets, ncdets, E_ASCI, norb, NULL, NULL);
  int64_t nkeep = (int64_t)ceil(keep_fraction * (double)nblk);
  if(nkeep < 1) nkeep = 1;
  if(nkeep > nblk) nkeep = nblk;
  bscore_t* o = (bscore_t*)malloc(sizeof(bscore_t) * (nblk + 1));
  char* kept = (char*)calloc(nblk + 1, 1);
  for(int64_t g = 0; g < nblk; ++g) {
    float v = scores[g];
    o[g].s = v != v ? -INFINITY : v;
    o[g].i = g;
  }
  qsort(o, nblk, sizeof(bscore_t), cmp_bscore);
  for(int64_t t = 0; t < nkeep; ++t) kept[o[t].i] = 1;
  free(o);
  core_t k;
  k.eps_a = (double*)malloc(sizeof(double) * norb);
  k.eps_b = (double*)malloc(sizeof(double) * norb);
  pair_buf b = {NULL, 0, 0};
  int rc = ORC_OK;
  int64_t g = 0;
  for(int64_t c = 0; c < ncdets && !rc; ++c) {
    core_load(h, W, cdets + W * c, norb, &k);
    const int64_t nb = block_total(&k, s->just_singles);
    for(int64_t q = 0; q < nb && !rc; ++q)
      if(kept[g + q]) {
        double m;
        det_t bd;
        rc = block_walk(h, &k, q, E_ASCI, s->h_el_tol, C[c], &b, &m, &bd);
      }
    g += nb;
  }
  if(!rc) rc = asci_finish(W, &b, ndets_max, cdets, ncdets, C, new_dets, n_new, rv_out);
  free(b.p);
  free(kept);
  free(k.eps_a);
  free(k.eps_b);
  return rc;
}

typedef struct {
  double c;
  int64_t i;
} cidx_t;
static int cmp_absc_desc(const void* x_, const void* y_) {
  const cidx_t *x = (const cidx_t*)x_, *y = (const cidx_t*)y_;
  double ax = fabs(x->c), ay = fabs(y->c);
  if(ax != ay) return ax > ay ? -1 : 1;
  return (x->i > y->i) - (x->i < y->i);
}

int orc_asci_iter(const orc_ham* h, int W, const orc_asci_settings* s,
                  int64_t ci_max_subspace, double ci_res_tol,
                  double ci_matel_tol, int64_t ndets_max, double E0,
                  uint64_t* wfn, double* X, int64_t* nwfn, int64_t norb,
                  double* E_out) {
  /* asci_iter (iteration.hpp:16-67) */
  int64_t nw = *nwfn;
  if(nw > 1) { /* reorder_ci_on_coeff (determinant_sort.hpp:18-37) */
    cidx_t* o = (cidx_t*)malloc(sizeof(cidx_t) * nw);
    uint64_t* tw = (uint64_t*)malloc(sizeof(uint64_t) * W * nw);
    double* tx = (double*)malloc(sizeof(double) * nw);
    for(int64_t i = 0; i < nw; ++i) {
      o[i].c = X[i];
      o[i].i = i;
    }
    qsort(o, nw, sizeof(cidx_t), cmp_absc_desc);
    for(int64_t i = 0; i < nw; ++i) {
      tx[i] = X[o[i].i];
      memcpy(tw + W * i, wfn + W * o[i].i, sizeof(uint64_t) * W);
    }
    memcpy(X, tx, sizeof(double) * nw);
    memcpy(wfn, tw, sizeof(uint64_t) * W * nw);
    free(o);
    free(tw);
    free(tx);
  }
  const int64_t nkeep = s->ncdets_max < nw ? s->ncdets_max : nw;
  uint64_t* core = (uint64_t*)malloc(sizeof(uint64_t) * W * (nkeep + 1));
  double* Cc = (double*)malloc(sizeof(double) * (nkeep + 1));
  memcpy(core, wfn, sizeof(uint64_t) * W * nkeep);
  memcpy(Cc, X, sizeof(double) * nkeep);
  int64_t nn = 0;
  int rc = orc_asci_search(h, W, s, ndets_max, core, nkeep, E0, Cc, norb, wfn,
                           &nn, NULL);
  free(core);
  free(Cc);
  if(rc) return rc;
  *nwfn = nn;
  memset(X, 0, sizeof(double) * nn); /* fresh guess (iteration.hpp:33) */
  return orc_selected_ci_diag(h, W, wfn, nn, ci_matel_tol, ci_max_subspace,
                              ci_res_tol, X, E_out, NULL);
}

int orc_asci_grow(const orc_ham* h, int W, const orc_asci_settings* s,
                  int64_t ci_max_subspace, double ci_res_tol,
                  double ci_matel_tol, double* E0, uint64_t* wfn, double* X,
                  int64_t* nwfn, int64_t norb) {
  /* asci_grow (grow.hpp:16-173), grow_with_rot = false */
  while(*nwfn < s->ntdets_max) {
    int64_t nd = *nwfn * s->grow_factor;
    if(nd < s->ntdets_min) nd = s->ntdets_min;
    if(nd > s->ntdets_max) nd = s->ntdets_max;
    double E;
    int rc = orc_asci_iter(h, W, s, ci_max_subspace, ci_res_tol, ci_matel_tol,
                           nd, *E0, wfn, X, nwfn, norb, &E);
    if(rc) return rc;
    if(nd > *nwfn) return ORC_ERR_DID_NOT_GROW;
    *E0 = E;
  }
  return ORC_OK;
}

int orc_asci_refine(const orc_ham* h, int W, const orc_asci_settings* s,
                    int64_t ci_max_subspace, double ci_res_tol,
                    double ci_matel_tol, double* E0, uint64_t* wfn,
                    double* X, int64_t* nwfn, int64_t norb) {
  /* asci_refine (refine.hpp:14-68) */
  const int64_t ndets = *nwfn;
  for(int64_t it = 0; it < s->max_refine_iter; ++it) {
    double E;
    int rc = orc_asci_iter(h, W, s, ci_max_subspace, ci_res_tol, ci_matel_tol,
                           ndets, *E0, wfn, X, nwfn, norb, &E);
    if(rc) return rc;
    if(*nwfn != ndets) return ORC_ERR_REFINE_SIZE;
    const double dE = E - *E0;
    *E0 = E;
    if(fabs(dE) < s->refine_energy_tol) return ORC_OK;
  }
  return ORC_ERR_REFINE;
}

/* ===================================================================== */
/* Orbital transforms (transform.cxx:22-115), square case norb_new ==     */
/* norb_old = n, leading dimensions n; Y may alias X.                     */
/* ===================================================================== */
/* One quarter: out(c, p) = sum_i in(i, c) U(i, p) for c < ncols -- the
 * contracted index leaves the front and the new one joins at the back. */
static void quarter(int64_t n, int64_t ncols, const double* in, const double* U, double* out) {
  for(int64_t p = 0; p < n; ++p)
    for(int64_t c = 0; c < ncols; ++c) {
      double acc = 0.;
      for(int64_t i = 0; i < n; ++i) acc += in[i + n * c] * U[i + n * p];
      out[c + ncols * p] = acc;
    }
}

/* Y(p,q) = U(i,p) X(i,j) U(j,q) */
void orc_two_index_transform(int64_t n, const double* X, const double* U, double* Y) {
  double* t = (double*)malloc(sizeof(double) * n * n);
  double* y = (double*)malloc(sizeof(double) * n * n);
  quarter(n, n, X, U, t); /* t(j,p) */
  quarter(n, n, t, U, y); /* y(p,q) */
  memcpy(Y, y, sizeof(double) * n * n);
  free(t);
  free(y);
}

/* Y(p,q,r,s) = X(i,j,k,l) U(i,p) U(j,q) U(k,r) U(l,s): four quarters
 * (i,j,k,l) -> (j,k,l,p) -> (k,l,p,q) -> (l,p,q,r) -> (p,q,r,s) */
void orc_four_index_transform(int64_t n, const double* X, const double* U, double* Y) {
  const int64_t n3 = n * n * n, n4 = n3 * n;
  double* a = (double*)malloc(sizeof(double) * n4);
  double* b = (double*)malloc(sizeof(double) * n4);
  quarter(n, n3, X, U, a);
  quarter(n, n3, a, U, b);
  quarter(n, n3, b, U, a);
  quarter(n, n3, a, U, b);
  memcpy(Y, b, sizeof(double) * n4);
  free(a);
  free(b);
}

/* Natural-orbital rotation of grow.hpp:73-114: RDMs of (wfn, X), natural
 * orbitals = eigenvectors of -ordm (ascending, i.e. occupations
 * descending), T and V rotated in place, intermediates regenerated. */
int orc_natural_orbital_rotation(orc_ham* h, int W, const uint64_t* wfn, int64_t nwfn,
                                 const double* X, double* occupations) {
  const int64_t n = h->n, n2 = n * n, n4 = n2 * n2;
  double* ordm = (double*)calloc(n2, sizeof(double));
  double* trdm = (double*)calloc(n4, sizeof(double));
  double* ons = (double*)malloc(sizeof(double) * n);
  if(!ordm || !trdm || !ons) {
    free(ordm);
    free(trdm);
    free(ons);
    return ORC_ERR_NOMEM;
  }
  orc_form_rdms(n, W, wfn, nwfn, wfn, nwfn, X, ordm, trdm);
  for(int64_t i = 0; i < n2; ++i) ordm[i] *= -1.0;
  orc_syev_lower(n, ordm, n, ons);
  for(int64_t i = 0; i < n; ++i) ons[i] *= -1.0;
  if(occupations) memcpy(occupations, ons, sizeof(double) * n);
  orc_two_index_transform(n, h->T, ordm, h->T);
  orc_four_index_transform(n, h->V, ordm, h->V);
  ham_intermediates(h);
  free(ordm);
  free(trdm);
  free(ons);
  return ORC_OK;
}

/* asci_grow (grow.hpp:16-173) with grow_with_rot: after each asci_iter that
 * leaves at least rot_size_start determinants, rotate to natural orbitals
 * and re-diagonalise the same determinants; E0 takes the asci_iter energy
 * (:162). The rotated integrals stay in h. */
int orc_asci_grow_rot(orc_ham* h, int W, const orc_asci_settings* s, int64_t ci_max_subspace,
                      double ci_res_tol, double ci_matel_tol, double* E0, uint64_t* wfn,
                      double* X, int64_t* nwfn, int64_t norb, int64_t rot_size_start) {
  while(*nwfn < s->ntdets_max) {
    int64_t nd = *nwfn * s->grow_factor;
    if(nd < s->ntdets_min) nd = s->ntdets_min;
    if(nd > s->ntdets_max) nd = s->ntdets_max;
    double E;
    int rc = orc_asci_iter(h, W, s, ci_max_subspace, ci_res_tol, ci_matel_tol, nd, *E0, wfn, X,
                           nwfn, norb, &E);
    if(rc) return rc;
    if(nd > *nwfn) return ORC_ERR_DID_NOT_GROW;
    if(*nwfn >= rot_size_start) {
      rc = orc_natural_orbital_rotation(h, W, wfn, *nwfn, X, NULL);
      if(rc) return rc;
      double Er;
      rc = orc_selected_ci_diag(h, W, wfn, *nwfn, ci_matel_tol, ci_max_subspace, ci_res_tol, X,
                                &Er, NULL);
      if(rc) return rc;
    }
    *E0 = E;
  }
  return ORC_OK;
}

/* ===================================================================== */
/* RDMs: form_rdms (double_loop.hpp:116-165) + rdm_contributions          */
/* (util/rdms.hpp:15-132). ordm(p,q) = ordm[p + n q], trdm(p,q,r,s) =     */
/* trdm[p + n q + n^2 r + n^3 s] (column-major spans, types.hpp:24-33);   */
/* both are accumulated into (+=), as the reference does.                 */
/* ===================================================================== */
#define ORC_T4(p, q, r, s) trdm[(p) + n * ((q) + n * ((r) + n * (int64_t)(s)))]

/* rdm_contributions_4 (util/rdms.hpp:15-26) */
static void rdm_4(int64_t n, uint64_t bra, uint64_t ket, uint64_t ex, double val,
                  double* trdm) {
  uint32_t o1, v1, o2, v2;
  const double sign = orc_doubles_sign_indices(bra, ket, ex, &o1, &v1, &o2, &v2);
  val *= sign * 0.5;
  ORC_T4(v1, o1, v2, o2) += val;
  ORC_T4(v2, o1, v1, o2) -= val;
  ORC_T4(v1, o2, v2, o1) -= val;
  ORC_T4(v2, o2, v1, o1) += val;
}

/* single_excitation_sign_indices (sd_operations.hpp:516-525) */
static double single_sign_indices(uint64_t bra, uint64_t ket, uint64_t ex, uint32_t* o,
                                  uint32_t* v) {
  *o = ffs0(ket & ex);
  *v = ffs0(bra & ex);
  return orc_single_excitation_sign(ket, *v, *o);
}

/* rdm_contributions_22 (util/rdms.hpp:28-42) */
static void rdm_22(int64_t n, det_t bra, det_t ket, uint64_t exa, uint64_t exb,
                   double val, double* trdm) {
  uint32_t o1, v1, o2, v2;
  const double sa = single_sign_indices(bra.a, ket.a, exa, &o1, &v1);
  const double sb = single_sign_indices(bra.b, ket.b, exb, &o2, &v2);
  val *= sa * sb * 0.5;
  ORC_T4(v1, o1, v2, o2) += val;
  ORC_T4(v2, o2, v1, o1) += val;
}

/* rdm_contributions_2 (util/rdms.hpp:44-67): `same` / `othr` are the bra's
 * occupied orbitals of the excited spin and of the other spin. */
static void rdm_2(int64_t n, uint64_t bra, uint64_t ket, uint64_t ex, const uint32_t* same,
                  int nsame, const uint32_t* othr, int nothr, double val, double* ordm,
                  double* trdm) {
  uint32_t o1, v1;
  const double sign = single_sign_indices(bra, ket, ex, &o1, &v1);
  ordm[v1 + n * o1] += sign * val;
  if(trdm) {
    val *= sign * 0.5;
    for(int k = 0; k < nsame; ++k) {
      const uint32_t p = same[k];
      ORC_T4(v1, o1, p, p) += val;
      ORC_T4(p, p, v1, o1) += val;
      ORC_T4(v1, p, p, o1) -= val;
      ORC_T4(p, o1, v1, p) -= val;
    }
    for(int k = 0; k < nothr; ++k) {
      const uint32_t p = othr[k];
      ORC_T4(v1, o1, p, p) += val;
      ORC_T4(p, p, v1, o1) += val;
    }
  }
}

/* rdm_contributions_diag (util/rdms.hpp:69-101) */
static void rdm_diag(int64_t n, const uint32_t* oa, int na, const uint32_t* ob, int nb,
                     double val, double* ordm, double* trdm) {
  for(int k = 0; k < na; ++k) ordm[oa[k] * (n + 1)] += val;
  for(int k = 0; k < nb; ++k) ordm[ob[k] * (n + 1)] += val;
  if(!trdm) return;
  val *= 0.5;
  for(int kq = 0; kq < na; ++kq)
    for(int kp = 0; kp < na; ++kp) {
      const uint32_t p = oa[kp], q = oa[kq];
      ORC_T4(p, p, q, q) += val;
      ORC_T4(p, q, p, q) -= val;
    }
  for(int kq = 0; kq < nb; ++kq)
    for(int kp = 0; kp < nb; ++kp) {
      const uint32_t p = ob[kp], q = ob[kq];
      ORC_T4(p, p, q, q) += val;
      ORC_T4(p, q, p, q) -= val;
    }
  for(int kq = 0; kq < nb; ++kq)
    for(int kp = 0; kp < na; ++kp) {
      const uint32_t p = oa[kp], q = ob[kq];
      ORC_T4(p, p, q, q) += val;
      ORC_T4(q, q, p, p) += val;
    }
}

/* form_rdms (double_loop.hpp:116-165) with the dispatch of
 * rdm_contributions (util/rdms.hpp:103-130). C[i] weights bra i and C[j]
 * ket j (one array, as the reference indexes it); trdm may be NULL. */
void orc_form_rdms(int64_t norb, int W, const uint64_t* bra, int64_t nbra,
                   const uint64_t* ket, int64_t nket, const double* C, double* ordm,
                   double* trdm) {
  const int64_t n = norb;
  uint32_t oa[64], ob[64];
  for(int64_t i = 0; i < nbra; ++i) {
    const det_t b = load_det(W, bra + i * W);
    if(!(b.a | b.b)) continue;
    const int na = bits_to_indices(b.a, oa), nb = bits_to_indices(b.b, ob);
    for(int64_t j = 0; j < nket; ++j) {
      const det_t k = load_det(W, ket + j * W);
      if(!(k.a | k.b)) continue;
      const uint64_t exa = b.a ^ k.a, exb = b.b ^ k.b;
      const int ca = popc(exa), cb = popc(exb);
      if(ca + cb > 4) continue;
      const double val = C[i] * C[j];
      if(!(fabs(val) > 1e-16)) continue;
      if(ca == 4 && trdm)
        rdm_4(n, b.a, k.a, exa, val, trdm);
      else if(cb == 4 && trdm)
        rdm_4(n, b.b, k.b, exb, val, trdm);
      else if(ca == 2 && cb == 2 && trdm)
        rdm_22(n, b, k, exa, exb, val, trdm);
      else if(ca == 2)
        rdm_2(n, b.a, k.a, exa, oa, na, ob, nb, val, ordm, trdm);
      else if(cb == 2)
        rdm_2(n, b.b, k.b, exb, ob, nb, oa, na, val, ordm, trdm);
      else
        rdm_diag(n, oa, na, ob, nb, val, ordm, trdm);
    }
  }
}
#undef ORC_T4

/* ===================================================================== */
/* MLP pre-screen oracle (no reference counterpart)                        */
/* ===================================================================== */
/* Float32 -> bf16 -> float32 with round-to-nearest-even. */
static double bf16r(double x) {
  float f = (float)x;
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  memcpy(&f, &u, 4);
  return (double)f;
}
static double gelu_d(double z) {
  const double u = 0.7978845608028654 * (z + 0.044715 * z * z * z);
  return 0.5 * z * (1. + tanh(u));
}
static double gelu_grad_d(double z) {
  const double u = 0.7978845608028654 * (z + 0.044715 * z * z * z);
  const double t = tanh(u);
  return 0.5 * (1. + t) +
         0.5 * z * (1. - t * t) * 0.7978845608028654 * (1. + 3. * 0.044715 * z * z);
}

#define MLP_IN 128
#define MLP_H 512
int64_t orc_mlp_num_params(void) {
  return (int64_t)MLP_H * MLP_IN + MLP_H + (int64_t)MLP_H * MLP_H + MLP_H + MLP_H + 1;
}

double orc_mlp_loss_grad(const float* P, int W, const uint64_t* dets,
                         const float* target, int64_t n, int bf16_points,
                         double* grad, double* pred) {
  const int64_t oW1 = 0, ob1 = oW1 + MLP_H * MLP_IN, oW2 = ob1 + MLP_H,
                ob2 = oW2 + (int64_t)MLP_H * MLP_H, ow3 = ob2 + MLP_H,
                ob3 = ow3 + MLP_H;
  const int64_t np = orc_mlp_num_params();
#define RB(x) (bf16_points ? bf16r(x) : (x))
  double* Wq = (double*)malloc(sizeof(double) * np);
  for(int64_t i = 0; i < np; ++i) Wq[i] = P[i];
  for(int64_t i = oW1; i < ob1; ++i) Wq[i] = RB(P[i]);
  for(int64_t i = oW2; i < ob2; ++i) Wq[i] = RB(P[i]);
  if(grad) memset(grad, 0, sizeof(double) * np);
  double* x = (double*)malloc(sizeof(double) * MLP_IN);
  double *z1 = (double*)malloc(sizeof(double) * MLP_H), *a1 = (double*)malloc(sizeof(double) * MLP_H);
  double *z2 = (double*)malloc(sizeof(double) * MLP_H), *a2 = (double*)malloc(sizeof(double) * MLP_H);
  double *dz2 = (double*)malloc(sizeof(double) * MLP_H), *dz1 = (double*)malloc(sizeof(double) * MLP_H);
  double loss = 0.;
  for(int64_t b = 0; b < n; ++b) {
    for(int f = 0; f < MLP_IN; ++f)
      x[f] = (f < 64 * W) ? (double)((dets[b * W + (f >> 6)] >> (f & 63)) & 1ULL) : 0.;
    for(int j = 0; j < MLP_H; ++j) {
      double s = 0.;
      for(int f = 0; f < MLP_IN; ++f) s += Wq[oW1 + (int64_t)j * MLP_IN + f] * x[f];
      z1[j] = s + Wq[ob1 + j];
      a1[j] = RB(gelu_d(z1[j]));
    }
    for(int j = 0; j < MLP_H; ++j) {
      double s = 0.;
      for(int k = 0; k < MLP_H; ++k) s += Wq[oW2 + (int64_t)j * MLP_H + k] * a1[k];
      z2[j] = s + Wq[ob2 + j];
      a2[j] = RB(gelu_d(z2[j]));
    }
    double y = Wq[ob3];
    for(int j = 0; j < MLP_H; ++j) y += a2[j] * Wq[ow3 + j];
    if(pred) pred[b] = y;
    if(!target) continue;
    const double e = y - target[b];
    loss += e * e;
    if(!grad) continue;
    const double dy = 2. * e / (double)n;
    grad[ob3] += dy;
    for(int j = 0; j < MLP_H; ++j) {
      grad[ow3 + j] += a2[j] * dy;
      dz2[j] = RB(dy * Wq[ow3 + j] * gelu_grad_d(z2[j]));
    }
    for(int j = 0; j < MLP_H; ++j) {
      grad[ob2 + j] += dz2[j];
      for(int k = 0; k < MLP_H; ++k) grad[oW2 + (int64_t)j * MLP_H + k] += dz2[j] * a1[k];
    }
    for(int k = 0; k < MLP_H; ++k) {
      double s = 0.;
      for(int j = 0; j < MLP_H; ++j) s += dz2[j] * Wq[oW2 + (int64_t)j * MLP_H + k];
      dz1[k] = RB(s * gelu_grad_d(z1[k]));
    }
    for(int j = 0; j < MLP_H; ++j) {
      grad[ob1 + j] += dz1[j];
      for(int f = 0; f < MLP_IN; ++f) grad[oW1 + (int64_t)j * MLP_IN + f] += dz1[j] * x[f];
    }
  }
#undef RB
  free(Wq);
  free(x);
  free(z1);
  free(a1);
  free(z2);
  free(a2);
  free(dz2);
  free(dz1);
  return target ? loss / (double)n : 0.;
}
