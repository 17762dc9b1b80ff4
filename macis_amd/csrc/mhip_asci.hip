// K6-K8: ASCI search on the device (asci_search, world_size == 1,
// determinant_search.hpp:394-663; candidate kernels of
// determinant_contributions.hpp:25-160).
//
//   prep    : per core determinant, occupied/virtual lists, the orbital
//             energies single_orbital_ens (fast_diagonals.hpp:31-51) and the
//             diagonal element.
//   count   : every core determinant owns a flat candidate index space
//             [alpha singles | beta singles | aa doubles | bb doubles |
//             ab doubles]; one wavefront walks a 4096-candidate chunk, applies
//             the reference's screens (|h| < h_el_tol for singles,
//             |G| / |V| < h_el_tol for doubles) and counts survivors with a
//             ballot.
//   fill    : the chunk offsets (device scan) give every survivor a fixed
//             slot; the same walk writes (determinant key, rv) with
//             rv = c * sign * h / (E0 - H_kk^fast), bit-identical to the
//             reference's expression.
//   sort    : radix sort by the determinant key (bitset_less order: beta
//             word most significant), reduce-by-key sums duplicates
//             (sort_and_accumulate_asci_pairs, determinant_sort.hpp:39-83).
//   top-k   : drops core determinants and zero scores (:528-544), then a
//             stable descending radix sort on |rv| keeps the
//             ndets_max - ncdets largest (:563-629); ties keep determinant
//             order. The output is top-k followed by the core dets (:653).
// The count/fill decomposition is fixed, so the result is run-to-run
// deterministic. When the candidates outnumber pair_size_max, the
// reference's pruning (:120-137) is replayed on the sorted list before the
// accumulation (prune_* kernels).
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce_by_key.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "mhip_device.hpp"
#include "mhip_internal.hpp"

namespace mhip {
IntsView device_view(const Integrals& I);

namespace {

constexpr int kChunk = 4096;  // candidates per wavefront chunk
constexpr int64_t kStage2Blocks = 8192;  // grid cap of the pruning stage-2 passes
using u128 = rocprim::uint128_t;

struct CoreArrays {
  const uint64_t* __restrict__ sa;     // alpha strings [nc]
  const uint64_t* __restrict__ sb;     // beta strings [nc]
  const uint8_t* __restrict__ lists;   // [nc][4][64]: occ_a, vir_a, occ_b, vir_b
  const int32_t* __restrict__ counts;  // [nc][4]: na, nva, nb, nvb
  const double* __restrict__ eps;      // [nc][2][n]
  const double* __restrict__ hdiag;    // [nc]
  const double* __restrict__ coeff;    // [nc]
  const int64_t* __restrict__ chunk_start;  // [nc+1]
  int64_t nc;
  int n;
  int just_singles;
};

// ---- prep: thread per core det -------------------------------------------
template <int W>
__global__ void prep_kernel(IntsView I, const uint64_t* __restrict__ cdets, int64_t nc,
                            int norb, uint64_t* __restrict__ sa, uint64_t* __restrict__ sb,
                            uint8_t* __restrict__ lists, int32_t* __restrict__ counts,
                            double* __restrict__ eps, double* __restrict__ hdiag) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(c >= nc) return;
  const Spins<W> s = load_spins<W>(cdets, c);
  sa[c] = s.a;
  sb[c] = s.b;
  const int n = I.n;
  for(int spin = 0; spin < 2; ++spin) {
    const uint64_t st = spin ? s.b : s.a;
    uint8_t* occ = lists + (c * 4 + 2 * spin) * 64;
    uint8_t* vir = lists + (c * 4 + 2 * spin + 1) * 64;
    int no = 0;
    for(uint64_t m = st; m; m &= m - 1) occ[no++] = (uint8_t)ctz64(m);
    // bitset_to_occ_vir: the first norb - nocc clear bits (sd_operations.hpp:82-97)
    const int nv = norb - no;
    uint64_t t = ~st;
    for(int v = 0; v < nv; ++v) {
      vir[v] = (uint8_t)ctz64(t);
      t &= t - 1;
    }
    counts[c * 4 + 2 * spin] = no;
    counts[c * 4 + 2 * spin + 1] = nv;
    // single_orbital_ens(norb, ss_occ, os_occ)
    const uint64_t ss = st, os = spin ? s.a : s.b;
    double* e = eps + (c * 2 + spin) * n;
    for(int i = 0; i < norb; ++i) {
      double v = I.T[i * (n + 1)];
      for(uint64_t m = ss; m; m &= m - 1) {
        const int q = ctz64(m);
        v += I.G2_red[i + n * q] + I.G2_red[q + n * i];
      }
      v -= I.G2_red[i * (n + 1)];
      for(uint64_t m = os; m; m &= m - 1) v += I.V2_red[i + n * ctz64(m)];
      e[i] = v;
    }
  }
  hdiag[c] = me_diag(I, s.a, s.b);
}

// Pair (x < y) of the lexicographic enumeration of pairs of m items.
__device__ __forceinline__ void tri_decode(int64_t p, int m, int& x, int& y) {
  const double b = 2.0 * m - 1.0;
  int xx = (int)floor((b - sqrt(b * b - 8.0 * (double)p)) * 0.5);
  auto start = [&](int q) { return (int64_t)q * m - (int64_t)q * (q + 1) / 2; };
  if(xx < 0) xx = 0;
  while(xx > 0 && start(xx) > p) --xx;
  while(xx + 1 < m && start(xx + 1) <= p) ++xx;
  x = xx;
  y = (int)(p - start(xx)) + xx + 1;
}

__device__ __forceinline__ int64_t pairs(int m) { return (int64_t)m * (m - 1) / 2; }

struct Cand {
  bool keep;
  uint64_t a, b;  // excited alpha / beta strings
  double rv;
  // entries of the alpha / beta string tables of the core det's strings
  // (0: the string itself, 1 + single index, 1 + singles + double index)
  int32_t ga, gb;
};

// Evaluates flat candidate f of core det c. With FULL = false only the
// screen is evaluated (count pass).
template <bool FULL>
__device__ __forceinline__ Cand candidate(const IntsView& I, const CoreArrays& A, int64_t c,
                                          int64_t f, double E0, double tol) {
  Cand r;
  r.keep = false;
  r.a = r.b = 0;
  r.rv = 0.;
  r.ga = r.gb = 0;
  const int n = I.n;
  const int na = A.counts[c * 4 + 0], nva = A.counts[c * 4 + 1];
  const int nb = A.counts[c * 4 + 2], nvb = A.counts[c * 4 + 3];
  const uint8_t* L = A.lists + c * 4 * 64;
  const uint64_t sa = A.sa[c], sb = A.sb[c];
  const double hd0 = A.hdiag[c];
  const double coeff = A.coeff[c];
  const double* eps_a = A.eps + (c * 2 + 0) * n;
  const double* eps_b = A.eps + (c * 2 + 1) * n;
  const int64_t Sa = (int64_t)na * nva, Sb = (int64_t)nb * nvb;
  if(f < Sa + Sb) {
    // singles (append_singles_asci_contributions, :25-66)
    const int spin = f < Sa ? 0 : 1;
    const int64_t g = spin ? f - Sa : f;
    const int nv = spin ? nvb : nva;
    const int i = L[(2 * spin) * 64 + g / nv], a = L[(2 * spin + 1) * 64 + g % nv];
    const uint64_t same = spin ? sb : sa, othr = spin ? sa : sb;
    double h = I.T[a + i * n];
    const double* G_ov = I.G_red + a * n + (int64_t)i * n * n;
    const double* V_ov = I.V_red + a * n + (int64_t)i * n * n;
    for(uint64_t m = same; m; m &= m - 1) h += G_ov[ctz64(m)];
    for(uint64_t m = othr; m; m &= m - 1) h += V_ov[ctz64(m)];
    if(fabs(h) < tol) return r;
    r.keep = true;
    if constexpr(FULL) {
      const uint64_t fl = (1ull << i) | (1ull << a);
      r.a = spin ? sa : sa ^ fl;
      r.b = spin ? sb ^ fl : sb;
      (spin ? r.gb : r.ga) = (int32_t)(1 + g);
      h *= excitation_sign(same, a, i);
      const double* eps = spin ? eps_b : eps_a;
      const double hd = hd0 + eps[a] - eps[i] - I.G2_red[a + n * i] - I.G2_red[i + n * a];
      h /= (E0 - hd);
      r.rv = coeff * h;
    }
    return r;
  }
  if(A.just_singles) return r;
  int64_t g = f - Sa - Sb;
  const int64_t Da = pairs(na) * pairs(nva), Db = pairs(nb) * pairs(nvb);
  if(g < Da + Db) {
    // same-spin doubles (append_ss_doubles_asci_contributions, :68-120)
    const int spin = g < Da ? 0 : 1;
    if(spin) g -= Da;
    const int no = spin ? nb : na, nv = spin ? nvb : nva;
    const int64_t pv = pairs(nv);
    int ii, jj, aa, bb;
    tri_decode(g / pv, no, ii, jj);
    tri_decode(g % pv, nv, aa, bb);
    const int i = L[(2 * spin) * 64 + ii], j = L[(2 * spin) * 64 + jj];
    const int a = L[(2 * spin + 1) * 64 + aa], b = L[(2 * spin + 1) * 64 + bb];
    const double Gv = I.G[((int64_t)a + (int64_t)i * n) * n * n + b + (int64_t)j * n];
    if(fabs(Gv) < tol) return r;
    r.keep = true;
    if constexpr(FULL) {
      const uint64_t st = spin ? sb : sa;
      const uint64_t fex = (1ull << i) | (1ull << j) | (1ull << a) | (1ull << b);
      uint64_t ket = st ^ fex, ex = fex;
      // doubles_sign(state, excited, ex) (sd_operations.hpp:528-551)
      const uint32_t o1 = ctz64(ket & ex), v1 = ctz64(st & ex);
      double sign = excitation_sign(ket, v1, o1);
      const uint64_t f1 = (1ull << o1) | (1ull << v1);
      ket ^= f1;
      ex ^= f1;
      const uint32_t o2 = ctz64(ket & ex), v2 = ctz64(st & ex);
      sign *= excitation_sign(ket, v2, o2);
      r.a = spin ? sa : sa ^ fex;
      r.b = spin ? sb ^ fex : sb;
      (spin ? r.gb : r.ga) = (int32_t)(1 + (spin ? Sb : Sa) + g);
      double h = sign * Gv;
      const double* eps = spin ? eps_b : eps_a;
      const double* G2 = I.G2_red;
      // fast_diag_ss_double (fast_diagonals.hpp:72-85), left to right
      const double hd = hd0 + eps[a] + eps[b] - eps[i] - eps[j] + G2[i + n * j] +
                        G2[j + n * i] + G2[a + n * b] + G2[b + n * a] - G2[a + n * i] -
                        G2[i + n * a] - G2[b + n * i] - G2[i + n * b] - G2[a + n * j] -
                        G2[j + n * a] - G2[b + n * j] - G2[j + n * b];
      h /= (E0 - hd);
      r.rv = coeff * h;
    }
    return r;
  }
  // opposite-spin doubles (append_os_doubles_asci_contributions, :122-160)
  g -= Da + Db;
  const int64_t sbj = (int64_t)nb * nvb;
  const int64_t ia = g / sbj, jb = g % sbj;
  const int i = L[0 * 64 + ia / nva], a = L[1 * 64 + ia % nva];
  const int j = L[2 * 64 + jb / nvb], b = L[3 * 64 + jb % nvb];
  const double Vv = I.V[a + (int64_t)i * n + ((int64_t)b + (int64_t)j * n) * n * n];
  if(fabs(Vv) < tol) return r;
  r.keep = true;
  if constexpr(FULL) {
    const double sgn_a = excitation_sign(sa, a, i);
    const double sgn_b = excitation_sign(sb, b, j);
    const double sign = sgn_a * sgn_b;
    r.a = sa ^ ((1ull << a) | (1ull << i));
    r.b = sb ^ ((1ull << j) | (1ull << b));
    r.ga = (int32_t)(1 + ia);
    r.gb = (int32_t)(1 + jb);
    double h = sign * Vv;
    const double* G2 = I.G2_red;
    const double* V2 = I.V2_red;
    // fast_diag_os_double (fast_diagonals.hpp:101-112), left to right
    const double hd = hd0 + eps_a[a] + eps_b[b] - eps_a[i] - eps_b[j] + V2[i + n * j] +
                      V2[a + n * b] - G2[a + n * i] - G2[i + n * a] - G2[b + n * j] -
                      G2[j + n * b] - V2[a + n * j] - V2[i + n * b];
    h /= (E0 - hd);
    r.rv = coeff * h;
  }
  return r;
}

__device__ __forceinline__ int64_t find_core(const int64_t* __restrict__ start, int64_t nc,
                                             int64_t chunk) {
  int64_t lo = 0, hi = nc;  // last c with start[c] <= chunk
  while(hi - lo > 1) {
    const int64_t mid = (lo + hi) / 2;
    if(start[mid] <= chunk)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// ---- dense determinant keys ----------------------------------------------
// Every candidate string is the alpha (beta) string of a core det with at
// most two electrons moved. Per distinct core string a table lists, in the
// order `candidate` enumerates them, that string, its singles and its
// doubles, each replaced by its rank among all distinct strings of all
// tables. A candidate's key is rank_beta << bits_a | rank_alpha: equal
// determinants get equal keys and the key order is bitset_less order (beta
// string first), as with the full strings, so the sort, accumulation,
// pruning replay and top-k run unchanged on keys of bits_a + bits_b bits
// instead of 2 * 32 * W.
struct KeyMap {
  const uint32_t* __restrict__ ta;  // alpha table entries -> rank
  const uint32_t* __restrict__ tb;
  const int64_t* __restrict__ oa;   // per core det: its alpha string's table block
  const int64_t* __restrict__ ob;
  const uint64_t* __restrict__ ua;  // alpha strings by rank
  const uint64_t* __restrict__ ub;
  int bits_a;
};

__device__ __forceinline__ uint64_t rank_key(const KeyMap& M, int64_t c, int32_t ga, int32_t gb) {
  return ((uint64_t)M.tb[M.ob[c] + gb] << M.bits_a) | (uint64_t)M.ta[M.oa[c] + ga];
}

// Thread per distinct string: occupied orbitals ascending, then the first
// norb - nocc clear bits (the lists prep_kernel builds per core det).
__global__ void str_lists_kernel(const uint64_t* __restrict__ str, int64_t ns, int norb,
                                 uint8_t* __restrict__ lists, int32_t* __restrict__ cnt) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(k >= ns) return;
  const uint64_t st = str[k];
  uint8_t* occ = lists + k * 128;
  uint8_t* vir = occ + 64;
  int no = 0;
  for(uint64_t m = st; m; m &= m - 1) occ[no++] = (uint8_t)ctz64(m);
  const int nv = norb - no;
  uint64_t t = ~st;
  for(int v = 0; v < nv; ++v) {
    vir[v] = (uint8_t)ctz64(t);
    t &= t - 1;
  }
  cnt[2 * k] = no;
  cnt[2 * k + 1] = nv;
}

// Thread per table entry e (block k: boff[k] <= e < boff[k + 1]): the
// string of entry e - boff[k] in `candidate`'s order, and e as payload.
__global__ void str_table_kernel(const uint64_t* __restrict__ str, const int64_t* __restrict__ boff,
                                 int64_t ns, const uint8_t* __restrict__ lists,
                                 const int32_t* __restrict__ cnt, int64_t E,
                                 uint64_t* __restrict__ keys, uint32_t* __restrict__ slots) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(e >= E) return;
  int64_t lo = 0, hi = ns;  // last k with boff[k] <= e
  while(hi - lo > 1) {
    const int64_t mid = (lo + hi) / 2;
    if(boff[mid] <= e)
      lo = mid;
    else
      hi = mid;
  }
  const int64_t k = lo, g = e - boff[k];
  const uint8_t* occ = lists + k * 128;
  const uint8_t* vir = occ + 64;
  const int no = cnt[2 * k], nv = cnt[2 * k + 1];
  const int64_t S = (int64_t)no * nv;
  uint64_t x = str[k];
  if(g > 0 && g <= S) {
    const int64_t q = g - 1;
    x ^= (1ull << occ[q / nv]) | (1ull << vir[q % nv]);
  } else if(g > S) {
    const int64_t q = g - 1 - S, pv = pairs(nv);
    int ii, jj, aa, bb;
    tri_decode(q / pv, no, ii, jj);
    tri_decode(q % pv, nv, aa, bb);
    x ^= (1ull << occ[ii]) | (1ull << occ[jj]) | (1ull << vir[aa]) | (1ull << vir[bb]);
  }
  keys[e] = x;
  slots[e] = (uint32_t)e;
}

__global__ void str_head_kernel(const uint64_t* __restrict__ sk, int64_t n,
                                int32_t* __restrict__ head) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(p < n) head[p] = (p == 0 || sk[p] != sk[p - 1]) ? 1 : 0;
}

// rank1 = inclusive count of distinct strings: table entries get their
// string's rank, and each rank its string
__global__ void str_rank_kernel(const uint64_t* __restrict__ sk, const uint32_t* __restrict__ ss,
                                const int32_t* __restrict__ rank1, int64_t E,
                                uint32_t* __restrict__ table, uint64_t* __restrict__ uniq) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(p >= E) return;
  const int32_t r = rank1[p] - 1;
  table[ss[p]] = (uint32_t)r;
  if(p == 0 || rank1[p - 1] != rank1[p]) uniq[r] = sk[p];
}

// keys of the core dets themselves (entry 0 of their strings' tables)
__global__ void core_rank_keys_kernel(KeyMap M, int64_t nc, uint64_t* __restrict__ keys) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(c < nc) keys[c] = rank_key(M, c, 0, 0);
}

template <int W>
__global__ void rank_keys_to_dets_kernel(KeyMap M, const uint64_t* __restrict__ keys, int64_t n,
                                         uint64_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const uint64_t k = keys[i];
  const uint64_t a = M.ua[k & ((1ull << M.bits_a) - 1ull)], b = M.ub[k >> M.bits_a];
  if constexpr(W == 1) {
    out[i] = a | (b << 32);
  } else {
    out[2 * i] = a;
    out[2 * i + 1] = b;
  }
}

template <int W, bool RANK>
using KeyT = typename std::conditional<RANK || W == 1, uint64_t, u128>::type;

template <int W, bool RANK>
__device__ __forceinline__ KeyT<W, RANK> cand_key(const KeyMap& M, int64_t c, const Cand& cd) {
  if constexpr(RANK)
    return rank_key(M, c, cd.ga, cd.gb);
  else if constexpr(W == 1)
    return cd.a | (cd.b << 32);
  else
    return ((u128)cd.b << 64) | (u128)cd.a;
}

template <int W, bool FILL, bool RANK>
__global__ __launch_bounds__(256) void contrib_kernel(
    IntsView I, CoreArrays A, const int64_t* __restrict__ ncand, int64_t nchunks, double E0,
    double tol, int64_t* __restrict__ chunk_counts, const int64_t* __restrict__ chunk_off,
    KeyT<W, RANK>* __restrict__ keys, double* __restrict__ vals, KeyMap M) {
  const int64_t chunk = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if(chunk >= nchunks) return;
  const uint32_t lane = threadIdx.x & 63;
  const int64_t c = find_core(A.chunk_start, A.nc, chunk);
  const int64_t f0 = (chunk - A.chunk_start[c]) * kChunk;
  const int64_t f1 = min(f0 + (int64_t)kChunk, ncand[c]);
  int64_t out = FILL ? chunk_off[chunk] : 0;
  int64_t cnt = 0;
  for(int64_t fb = f0; fb < f1; fb += 64) {
    const int64_t f = fb + lane;
    Cand cd;
    cd.keep = false;
    if(f < f1) cd = candidate<FILL>(I, A, c, f, E0, tol);
    const uint64_t m = ballot(cd.keep);
    if constexpr(FILL) {
      if(cd.keep) {
        const int64_t p = out + lanes_below(m);
        keys[p] = cand_key<W, RANK>(M, c, cd);
        vals[p] = cd.rv;
      }
      out += popc64(m);
    } else {
      cnt += popc64(m);
    }
  }
  if constexpr(!FILL)
    if(lane == 0) chunk_counts[chunk] = cnt;
}

// ---- excitation blocks (learned screen) -----------------------------------
// The candidates of a core determinant fall into blocks, one per hole set,
// in the order [alpha single holes (na) | beta single holes (nb) | alpha-
// alpha hole pairs (na(na-1)/2) | beta-beta hole pairs | alpha-beta hole
// pairs (na*nb)] (singles only with just_singles); a block's candidates are
// its particle choices. A block maps onto a run (alpha-beta: a strided set)
// of the flat candidate indices of `candidate`, so a block is evaluated with
// the same code and in the same order within its core determinant.
__host__ __device__ __forceinline__ int64_t block_count(int64_t na, int64_t nb, bool singles) {
  int64_t k = na + nb;
  if(!singles) k += na * (na - 1) / 2 + nb * (nb - 1) / 2 + na * nb;
  return k;
}

struct Blk {
  int64_t base;       // flat index of the block's first candidate
  int64_t size;       // candidates in the block
  int64_t rowstride;  // alpha-beta: flat step per alpha particle (0: a contiguous run)
  int nvb;            // alpha-beta: beta particles per alpha particle
  uint64_t ha, hb;    // hole bits of the block (alpha, beta string)
};

__device__ __forceinline__ Blk block_decode(const CoreArrays& A, int64_t c, int64_t q) {
  const int na = A.counts[c * 4 + 0], nva = A.counts[c * 4 + 1];
  const int nb = A.counts[c * 4 + 2], nvb = A.counts[c * 4 + 3];
  const uint8_t* L = A.lists + c * 4 * 64;
  const int64_t Sa = (int64_t)na * nva, Sb = (int64_t)nb * nvb;
  const int64_t pva = pairs(nva), pvb = pairs(nvb);
  const int64_t Da = pairs(na) * pva, Db = pairs(nb) * pvb;
  Blk b;
  b.rowstride = 0;
  b.nvb = 1;
  b.ha = b.hb = 0;
  if(q < na) {
    b.base = q * nva;
    b.size = nva;
    b.ha = 1ull << L[q];
    return b;
  }
  q -= na;
  if(q < nb) {
    b.base = Sa + q * nvb;
    b.size = nvb;
    b.hb = 1ull << L[128 + q];
    return b;
  }
  q -= nb;
  int x, y;
  if(q < pairs(na)) {
    tri_decode(q, na, x, y);
    b.base = Sa + Sb + q * pva;
    b.size = pva;
    b.ha = (1ull << L[x]) | (1ull << L[y]);
    return b;
  }
  q -= pairs(na);
  if(q < pairs(nb)) {
    tri_decode(q, nb, x, y);
    b.base = Sa + Sb + Da + q * pvb;
    b.size = pvb;
    b.hb = (1ull << L[128 + x]) | (1ull << L[128 + y]);
    return b;
  }
  q -= pairs(nb);
  const int ii = (int)(q / nb), jj = (int)(q % nb);
  b.base = Sa + Sb + Da + Db + (int64_t)ii * nva * Sb + (int64_t)jj * nvb;
  b.size = (int64_t)nva * nvb;
  b.rowstride = Sb;
  b.nvb = nvb;
  b.ha = 1ull << L[ii];
  b.hb = 1ull << L[128 + jj];
  return b;
}

__device__ __forceinline__ int64_t block_flat(const Blk& b, int64_t l) {
  return b.rowstride ? b.base + (l / b.nvb) * b.rowstride + l % b.nvb : b.base + l;
}

// Thread per block: its determinant (the core determinant with the block's
// holes cleared), in the layout of the input determinants.
template <int W>
__global__ void block_dets_kernel(CoreArrays A, const int64_t* __restrict__ bstart, int64_t nblk,
                                  uint64_t* __restrict__ out) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(g >= nblk) return;
  const int64_t c = find_core(bstart, A.nc, g);
  const Blk b = block_decode(A, c, g - bstart[c]);
  const uint64_t sa = A.sa[c] & ~b.ha, sb = A.sb[c] & ~b.hb;
  if constexpr(W == 1) {
    out[g] = sa | (sb << 32);
  } else {
    out[2 * g] = sa;
    out[2 * g + 1] = sb;
  }
}

// Wavefront per block: log of the largest |h / (E0 - H_kk)| among the
// block's candidates that pass the h_el_tol screen (-inf if none). A.coeff
// must hold ones, so `candidate` returns h / (E0 - H_kk) itself.
__global__ __launch_bounds__(256) void block_target_kernel(IntsView I, CoreArrays A,
                                                           const int64_t* __restrict__ bstart,
                                                           int64_t nblk, double E0, double tol,
                                                           float* __restrict__ target) {
  const int64_t g = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if(g >= nblk) return;
  const uint32_t lane = threadIdx.x & 63;
  const int64_t c = find_core(bstart, A.nc, g);
  const Blk b = block_decode(A, c, g - bstart[c]);
  double m = -1.0;
  for(int64_t l0 = 0; l0 < b.size; l0 += 64) {
    const int64_t l = l0 + lane;
    if(l < b.size) {
      const Cand cd = candidate<true>(I, A, c, block_flat(b, l), E0, tol);
      if(cd.keep) m = fmax(m, fabs(cd.rv));
    }
  }
#pragma unroll
  for(int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if(lane == 0) target[g] = m < 0. ? -INFINITY : (float)log(m);
}

// Block score -> descending-sortable key (NaN ranks last, -0 as +0) and the
// block index as payload. With `pred`, score = pred * scale + shift +
// log|C_c| (the network predicts the block's target); otherwise the
// caller's scores are used as given.
__global__ void block_key_kernel(const float* __restrict__ pred, float scale, float shift,
                                 const double* __restrict__ coeff,
                                 const int64_t* __restrict__ bstart, int64_t nc,
                                 const float* __restrict__ given, int64_t nblk,
                                 uint32_t* __restrict__ key, int64_t* __restrict__ idx) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(g >= nblk) return;
  float s;
  if(pred) {
    const int64_t c = find_core(bstart, nc, g);
    s = fmaf(pred[g], scale, shift) + (float)log(fabs(coeff[c]));
  } else {
    s = given[g];
  }
  if(s != s) s = -INFINITY;
  s += 0.0f;
  const uint32_t u = __float_as_uint(s);
  key[g] = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  idx[g] = g;
}

// Candidate count (FILL = false) and fill of the kept blocks: wavefront per
// kept block (kept ids ascending, so the fill runs in core-det order).
template <int W, bool FILL, bool RANK>
__global__ __launch_bounds__(256) void block_contrib_kernel(
    IntsView I, CoreArrays A, const int64_t* __restrict__ bstart, const int64_t* __restrict__ kept,
    int64_t nkept, double E0, double tol, int64_t* __restrict__ counts,
    const int64_t* __restrict__ off, KeyT<W, RANK>* __restrict__ keys,
    double* __restrict__ vals, KeyMap M) {
  const int64_t k = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if(k >= nkept) return;
  const uint32_t lane = threadIdx.x & 63;
  const int64_t g = kept[k];
  const int64_t c = find_core(bstart, A.nc, g);
  const Blk b = block_decode(A, c, g - bstart[c]);
  int64_t out = FILL ? off[k] : 0;
  int64_t cnt = 0;
  for(int64_t l0 = 0; l0 < b.size; l0 += 64) {
    const int64_t l = l0 + lane;
    Cand cd;
    cd.keep = false;
    if(l < b.size) cd = candidate<FILL>(I, A, c, block_flat(b, l), E0, tol);
    const uint64_t m = ballot(cd.keep);
    if constexpr(FILL) {
      if(cd.keep) {
        const int64_t p = out + lanes_below(m);
        keys[p] = cand_key<W, RANK>(M, c, cd);
        vals[p] = cd.rv;
      }
      out += popc64(m);
    } else {
      cnt += popc64(m);
    }
  }
  if constexpr(!FILL)
    if(lane == 0) counts[k] = cnt;
}

__global__ void fill_ones_kernel(double* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) p[i] = 1.0;
}

// ---- pair_size_max pruning (asci_contributions_standard, :120-137) -------
// The reference prunes its growing pair list after the core determinant at
// which it first exceeds pair_size_max: pairs with |rv| <= rv_prune_tol are
// erased (stage 1) and, if the list is still too long, it is sorted and
// accumulated (stage 2); later prunes then test accumulated scores. Here the
// whole candidate list exists at once, sorted stably by determinant with
// each pair's slot (slots run in core-det order, so "core det <= q" is
// "slot < first slot of det q+1"). The host walks per-core-det counts to
// place the events (stage 1 is count arithmetic); a stage 2 at q folds only
// the pairs of dets (q0, q] into their segments' accumulated scores (a left
// fold in core-det order continued from the score of q0: the reference's
// order after its stable sort) and first erases the accumulated entries
// that the first prune after q0 found small. Erased pairs get score 0
// before reduce-by-key; a determinant whose pairs are all erased sums to 0
// and leaves with the zero scores.
struct RvSlot {
  double rv;
  int64_t slot;
};

// Per chunk (one wavefront): (rv, slot) of its slots and the number of
// slots with |rv| <= tol.
__global__ __launch_bounds__(256) void prune_prep_kernel(int64_t nchunks,
                                                         const int64_t* __restrict__ chunk_off,
                                                         const double* __restrict__ vals, double tol,
                                                         RvSlot* __restrict__ rs,
                                                         int64_t* __restrict__ chunk_tiny) {
  const int64_t chunk = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if(chunk >= nchunks) return;
  const uint32_t lane = threadIdx.x & 63;
  int64_t tiny = 0;
  for(int64_t p = chunk_off[chunk] + lane; p < chunk_off[chunk + 1]; p += 64) {
    const double v = vals[p];
    rs[p] = RvSlot{v, p};
    tiny += fabs(v) <= tol ? 1 : 0;
  }
  tiny = wave_sum(tiny);
  if(lane == 0) chunk_tiny[chunk] = tiny;
}

// pos[slot] = sorted position
__global__ void prune_pos_kernel(const RvSlot* __restrict__ rs, int64_t P,
                                 int64_t* __restrict__ pos) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if(j < P) pos[rs[j].slot] = j;
}

__device__ __forceinline__ void add_counts(unsigned long long a, unsigned long long b,
                                           unsigned long long* __restrict__ cnt) {
  __shared__ unsigned long long red[2][4];
  const unsigned long long s0 = wave_sum(a), s1 = wave_sum(b);
  if((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = s0;
    red[1][threadIdx.x >> 6] = s1;
  }
  __syncthreads();
  if(threadIdx.x == 0) {
    const unsigned long long t0 = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    const unsigned long long t1 = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    if(t0) atomicAdd(cnt, t0);
    if(t1) atomicAdd(cnt + 1, t1);
  }
}

// The small accumulated entries listed by the last stage 2 (accf == 2) are
// erased: their pairs (slots < lim0, the dets folded into them) die.
template <typename K>
__global__ __launch_bounds__(256) void prune_drop_acc_kernel(
    const K* __restrict__ keys, const RvSlot* __restrict__ rs, int64_t P, int64_t lim0,
    const int64_t* __restrict__ list, int64_t nlist, uint8_t* __restrict__ accf,
    uint8_t* __restrict__ alive) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if(i >= nlist) return;
  const int64_t h = list[i];
  accf[h] = 0;
  for(int64_t k = h; k < P && keys[k] == keys[h]; ++k)
    if(rs[k].slot < lim0) alive[k] = 0;
}

// Stage 2 at core det q: the pairs of slots [lim0, lim1) (dets (q0, q])
// are folded into their segments, erasing those with |rv| <= tol. One
// thread per segment touched: the first such pair of the segment walks
// back to the segment head for its accumulated score and forward over the
// run. cnt[0] += segments gaining an entry; new entries with |acc| <= tol
// are marked accf = 2 and listed (cnt[1] of them) for the next prune.
// hd[j] = j where sorted position j starts a determinant's segment, else 0
// (an inclusive max-scan then gives every position its segment head)
template <typename K>
__global__ void seg_head_kernel(const K* __restrict__ keys, int64_t P, int64_t* __restrict__ hd) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if(j < P) hd[j] = (j == 0 || keys[j - 1] != keys[j]) ? j : 0;
}

// SEQ: threads run over the sorted positions and take those whose slot is
// in range (a streaming pass, cheaper for an event that covers a large
// share of the pairs); otherwise over the slots, each finding its sorted
// position through pos. Same leaders, same result.
template <typename K, bool SEQ>
__global__ __launch_bounds__(256) void prune_stage2_kernel(
    const K* __restrict__ keys, const RvSlot* __restrict__ rs, int64_t P,
    const int64_t* __restrict__ pos, const int64_t* __restrict__ head, int64_t lim0,
    int64_t lim1, double tol, double* __restrict__ acc, uint8_t* __restrict__ accf,
    uint8_t* __restrict__ alive, unsigned long long* __restrict__ cnt,
    int64_t* __restrict__ list) {
  // grid-stride over the items (sorted positions with SEQ, else slots
  // lim0...), at most a few thousand blocks: each block adds its count with
  // one atomic on a single counter, which at one block per 256 items had
  // serialised the large events
  const int64_t N = SEQ ? P : lim1 - lim0;
  unsigned long long gained = 0;
  for(int64_t b0 = (int64_t)blockIdx.x * 256; b0 < N; b0 += (int64_t)gridDim.x * 256) {
    const int64_t t = b0 + threadIdx.x;
    bool push = false;
    int64_t hp = 0;
    bool in;
    int64_t j = 0;
    if constexpr(SEQ) {
      j = t;
      in = t < P;
      if(in) {
        const int64_t sl = rs[t].slot;
        in = sl >= lim0 && sl < lim1;
      }
    } else {
      in = t < N;
      if(in) j = pos[lim0 + t];
    }
    if(in) {
      const K kj = keys[j];
      // leader: the first pair of the segment inside [lim0, lim1) (the
      // segment's pairs are in slot order); it folds for the segment, whose
      // accumulated score sits at the segment head
      if(j == 0 || keys[j - 1] != kj || rs[j - 1].slot < lim0) {
        const int64_t h = head[j];
        const uint8_t had = accf[h];
        bool any = had != 0;
        double a = any ? acc[h] : 0.;
        // the segment's pairs from j while their slot is below lim1, folded
        // in order; four pairs are loaded per step (a popular determinant
        // has hundreds)
        bool go = true;
        for(int64_t k0 = j; go; k0 += 4) {
          K kk[4];
          RvSlot rr[4];
#pragma unroll
          for(int u = 0; u < 4; ++u) {
            const int64_t k = k0 + u < P ? k0 + u : P - 1;
            kk[u] = keys[k];
            rr[u] = rs[k];
          }
#pragma unroll
          for(int u = 0; u < 4; ++u) {
            go = go && k0 + u < P && kk[u] == kj && rr[u].slot < lim1;
            if(go) {
              const double v = rr[u].rv;
              if(fabs(v) <= tol) {
                alive[k0 + u] = 0;
              } else {
                a = any ? a + v : v;
                any = true;
              }
            }
          }
        }
        if(any) {
          acc[h] = a;
          accf[h] = fabs(a) <= tol ? 2 : 1;
          gained += had ? 0ull : 1ull;
          push = fabs(a) <= tol;
          hp = h;
        }
      }
    }
    // small entries appended with one atomic per wavefront; list order does
    // not matter
    const uint64_t m = ballot(push);
    if(m) {
      const uint32_t lane = threadIdx.x & 63, first = (uint32_t)ctz64(m);
      unsigned long long base = 0;
      if(lane == first) base = atomicAdd(cnt + 1, (unsigned long long)popc64(m));
      base = __shfl(base, (int)first, 64);
      if(push) list[base + lanes_below(m)] = hp;
    }
  }
  add_counts(gained, 0ull, cnt);
}

// Scores for reduce-by-key after the last event: erased pairs and pairs of
// slots [lim0, lim1) (dets after the last stage 2 up to the last event)
// with |rv| <= tol give 0.
__global__ void prune_final_kernel(const RvSlot* __restrict__ rs, const uint8_t* __restrict__ alive,
                                   int64_t P, int64_t lim0, int64_t lim1, double tol,
                                   double* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if(j >= P) return;
  const RvSlot e = rs[j];
  const bool dead = !alive[j] || (e.slot >= lim0 && e.slot < lim1 && fabs(e.rv) <= tol);
  out[j] = dead ? 0. : e.rv;
}

// keep = not a core det and |rv| > 0 (the -|rv| < 0 partition, :528-544);
// absval receives |rv| for the top-k key.
template <typename K>
__global__ void flag_kernel(const K* __restrict__ ukeys, const double* __restrict__ usum,
                            const int64_t* __restrict__ nu_dev, const K* __restrict__ core,
                            int64_t nc, uint8_t* __restrict__ flags,
                            uint64_t* __restrict__ absbits) {
  const int64_t nu = *nu_dev;
  for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nu;
      i += (int64_t)gridDim.x * blockDim.x) {
    const K k = ukeys[i];
    int64_t lo = 0, hi = nc;
    while(lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if(core[mid] < k)
        lo = mid + 1;
      else
        hi = mid;
    }
    const bool in_core = lo < nc && core[lo] == k;
    const double v = fabs(usum[i]);
    flags[i] = (!in_core && -v < 0.0) ? 1 : 0;
    absbits[i] = __double_as_longlong(v);
  }
}

template <int W, typename K>
__global__ void keys_to_dets_kernel(const K* __restrict__ keys, int64_t n,
                                    uint64_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  if constexpr(W == 1)
    out[i] = (uint64_t)keys[i];
  else {
    out[2 * i] = (uint64_t)keys[i];
    out[2 * i + 1] = (uint64_t)(keys[i] >> 64);
  }
}

template <int W, typename K>
__global__ void dets_to_keys_kernel(const uint64_t* __restrict__ d, int64_t n,
                                    K* __restrict__ keys) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  if constexpr(W == 1)
    keys[i] = d[i];
  else
    keys[i] = ((K)d[2 * i + 1] << 64) | (K)d[2 * i];
}

inline int64_t host_pairs(int64_t m) { return m * (m - 1) / 2; }

// Candidates equal to a core determinant are not new (the reference keeps
// the core copy with +|C| and drops rv >= 0, determinant_search.hpp:533-544):
// their accumulated score is zeroed, which removes them with the zero scores.
template <typename K>
__global__ void core_zero_kernel(const K* __restrict__ ukeys, int64_t nu,
                                 const K* __restrict__ core, int64_t nc,
                                 double* __restrict__ usum) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(c >= nc) return;
  const K k = core[c];
  int64_t lo = 0, hi = nu;
  while(lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if(ukeys[mid] < k)
      lo = mid + 1;
    else
      hi = mid;
  }
  if(lo < nu && ukeys[lo] == k) usum[lo] = 0.;
}

// Digit histogram of |score| bit patterns straight from the accumulated
// scores (zero scores skipped).
__global__ __launch_bounds__(256) void score_hist_kernel(const double* __restrict__ usum,
                                                         int64_t n, int shift, uint64_t prefix,
                                                         uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[4096];
  for(int i = threadIdx.x; i < 4096; i += 256) h[i] = 0;
  __syncthreads();
  for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
      i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = (uint64_t)__double_as_longlong(fabs(usum[i]));
    const bool match = k != 0 && (shift + 12 >= 64 || (k >> (shift + 12)) == prefix);
    if(match) atomicAdd(&h[(k >> shift) & 4095u], 1u);
  }
  __syncthreads();
  for(int i = threadIdx.x; i < 4096; i += 256)
    if(h[i]) atomicAdd(&hist[i], h[i]);
}

// Candidates with |score| bits at or above the 24-bit prefix thr24 (and a
// nonzero score), compacted in one pass: keys and scores written in any order (one atomic per wavefront for the output
// position; entries past cap are counted, not written). The caller sorts
// them by key, which restores determinant order.
template <typename K>
__global__ __launch_bounds__(256) void select_above_kernel(const K* __restrict__ ukeys,
                                                           const double* __restrict__ usum,
                                                           int64_t n, uint64_t thr24, int64_t cap,
                                                           K* __restrict__ sk,
                                                           double* __restrict__ sv,
                                                           unsigned long long* __restrict__ cnt) {
  const uint32_t lane = threadIdx.x & 63;
  for(int64_t b0 = (int64_t)blockIdx.x * 256; b0 < n; b0 += (int64_t)gridDim.x * 256) {
    const int64_t i = b0 + threadIdx.x;
    double v = 0.;
    bool take = false;
    if(i < n) {
      v = usum[i];
      const uint64_t k = (uint64_t)__double_as_longlong(fabs(v));
      take = k != 0 && (k >> 40) >= thr24;
    }
    const uint64_t m = ballot(take);
    if(m) {
      const uint32_t first = (uint32_t)ctz64(m);
      unsigned long long base = 0;
      if(lane == first) base = atomicAdd(cnt, (unsigned long long)popc64(m));
      base = __shfl(base, (int)first, 64);
      const int64_t p = (int64_t)(base + lanes_below(m));
      if(take && p < cap) {
        sk[p] = ukeys[i];
        sv[p] = v;
      }
    }
  }
}

__global__ void abs_bits_kernel(const double* __restrict__ v, int64_t n,
                                uint64_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) out[i] = (uint64_t)__double_as_longlong(fabs(v[i]));
}

__global__ void bits_to_double_kernel(const uint64_t* __restrict__ b, int64_t n,
                                      double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) out[i] = __longlong_as_double((long long)b[i]);
}

// Optional stages of the search: export of the scored candidates, or the
// block screen (learned or with the caller's block scores).
struct SearchMode {
  // export: unique non-core candidates with their accumulated |rv|
  uint64_t* exp_dets = nullptr;
  double* exp_abs = nullptr;
  int64_t exp_cap = 0;
  int64_t* exp_n = nullptr;
  // block screen: only the keep_fraction best-scored excitation blocks are
  // generated; their exact contributions are accumulated and ranked as in
  // the plain search. Scores are network predictions (pred * scale + shift
  // + log|C_c|) or the caller's.
  mhip_mlp* mlp = nullptr;
  float pred_scale = 1.f, pred_shift = 0.f;
  const float* block_scores = nullptr;
  double keep_fraction = 1.0;
  bool screened() const { return mlp || block_scores; }
};

// Host side of the core determinants: validates them against norb and
// returns per-det alpha / beta electron counts.
template <int W>
void core_counts(mhip_context* ctx, const uint64_t* d_cdets, int64_t nc, int64_t norb,
                 std::vector<int>& na, std::vector<int>& nb) {
  std::vector<uint64_t> hc(nc * W);
  d2h(ctx, hc.data(), d_cdets, 8 * W * nc);
  na.resize(nc);
  nb.resize(nc);
  for(int64_t c = 0; c < nc; ++c) {
    uint64_t a, b;
    if(W == 1) {
      a = hc[c] & 0xffffffffull;
      b = hc[c] >> 32;
    } else {
      a = hc[2 * c];
      b = hc[2 * c + 1];
    }
    MHIP_REQUIRE(!((a | b) & (norb >= 64 ? 0ull : ~((1ull << norb) - 1ull))),
                 MHIP_ERR_INVALID_ARG, "core determinant occupies an orbital >= norb");
    na[c] = __builtin_popcountll(a);
    nb[c] = __builtin_popcountll(b);
  }
}

// Per-core-det arrays of the candidate kernels (prep_kernel's outputs).
struct CorePrep {
  DevBuf sa, sb, lists, counts, eps, hdiag;
  CoreArrays A;
};

template <int W>
void prepare_core(mhip_context* ctx, const IntsView& I, const uint64_t* d_cdets, int64_t nc,
                  int64_t norb, const double* d_C, int just_singles, CorePrep& P) {
  const int n = (int)ctx->ints.norb;
  P.sa.alloc(8 * nc);
  P.sb.alloc(8 * nc);
  P.lists.alloc(4 * 64 * nc);
  P.counts.alloc(16 * nc);
  P.eps.alloc(16 * n * nc);
  P.hdiag.alloc(8 * nc);
  hipLaunchKernelGGL(prep_kernel<W>, dim3((nc + 63) / 64), dim3(64), 0, ctx->stream, I, d_cdets,
                     nc, (int)norb, P.sa.as<uint64_t>(), P.sb.as<uint64_t>(),
                     P.lists.as<uint8_t>(), P.counts.as<int32_t>(), P.eps.as<double>(),
                     P.hdiag.as<double>());
  MHIP_HIP_CHECK(hipGetLastError());
  CoreArrays& A = P.A;
  A.sa = P.sa.as<uint64_t>();
  A.sb = P.sb.as<uint64_t>();
  A.lists = P.lists.as<uint8_t>();
  A.counts = P.counts.as<int32_t>();
  A.eps = P.eps.as<double>();
  A.hdiag = P.hdiag.as<double>();
  A.coeff = d_C;
  A.chunk_start = nullptr;
  A.nc = nc;
  A.n = n;
  A.just_singles = just_singles;
}

// Dense keys (KeyMap): the string tables of both spins. False (full-string
// keys) when MHIP_ASCI_KEYS=bits, when the tables would hold more entries
// than there are candidates to generate (unless MHIP_ASCI_KEYS=rank), or
// when the two rank widths exceed 64 bits.
struct RankPrep {
  DevBuf ta, tb, oa, ob, ua, ub;
  KeyMap M{};
  int bits = 0;  // bits_a + bits_b
  double ms = 0.;
};

template <int W>
bool prepare_ranks(mhip_context* ctx, const uint64_t* d_cdets, int64_t nc, int64_t norb,
                   bool singles_only, double gen_fraction, RankPrep& R) {
  const char* env = std::getenv("MHIP_ASCI_KEYS");
  const bool force = env && std::string(env) == "rank";
  if(env && std::string(env) == "bits") return false;
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t st = ctx->stream;
  std::vector<uint64_t> hc(nc * W);
  d2h(ctx, hc.data(), d_cdets, 8 * W * nc);
  struct Side {
    std::vector<uint64_t> u;
    std::vector<int64_t> boff, off_c;
    int64_t E = 0;
  } side[2];
  int64_t raw = 0;
  for(int sp = 0; sp < 2; ++sp) {
    Side& S = side[sp];
    std::vector<uint64_t> str(nc);
    for(int64_t c = 0; c < nc; ++c)
      str[c] = W == 1 ? (sp ? hc[c] >> 32 : hc[c] & 0xffffffffull) : hc[2 * c + sp];
    S.u = str;
    std::sort(S.u.begin(), S.u.end());
    S.u.erase(std::unique(S.u.begin(), S.u.end()), S.u.end());
    S.boff.assign(S.u.size() + 1, 0);
    for(size_t k = 0; k < S.u.size(); ++k) {
      const int64_t no = __builtin_popcountll(S.u[k]), nv = norb - no;
      int64_t len = 1 + no * nv;
      if(!singles_only) len += host_pairs(no) * host_pairs(nv);
      S.boff[k + 1] = S.boff[k] + len;
    }
    S.E = S.boff.back();
    S.off_c.resize(nc);
    for(int64_t c = 0; c < nc; ++c)
      S.off_c[c] = S.boff[std::lower_bound(S.u.begin(), S.u.end(), str[c]) - S.u.begin()];
  }
  for(int64_t c = 0; c < nc; ++c) {
    const int64_t na = __builtin_popcountll(W == 1 ? hc[c] & 0xffffffffull : hc[2 * c]);
    const int64_t nb = __builtin_popcountll(W == 1 ? hc[c] >> 32 : hc[2 * c + 1]);
    const int64_t Sa = na * (norb - na), Sb = nb * (norb - nb);
    raw += Sa + Sb;
    if(!singles_only)
      raw += host_pairs(na) * host_pairs(norb - na) + host_pairs(nb) * host_pairs(norb - nb) +
             Sa * Sb;
  }
  const int64_t E = side[0].E + side[1].E;
  if(side[0].E >= (1ll << 31) || side[1].E >= (1ll << 31)) return false;
  // gen_fraction: the share of the raw candidates that will be generated
  // (the block screen's keep fraction)
  // (a table entry costs about as much to build as the dense key saves per
  // candidate in the sort)
  if(!force && (double)E > gen_fraction * (double)raw) return false;
  int bits[2] = {0, 0};
  for(int sp = 0; sp < 2; ++sp) {
    Side& S = side[sp];
    const int64_t ns = (int64_t)S.u.size(), En = S.E;
    DevBuf du(8 * ns), dboff(8 * (ns + 1)), lists(128 * ns), cnt(8 * ns);
    MHIP_HIP_CHECK(hipMemcpyAsync(du.p, S.u.data(), 8 * ns, hipMemcpyHostToDevice, st));
    MHIP_HIP_CHECK(hipMemcpyAsync(dboff.p, S.boff.data(), 8 * (ns + 1), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(str_lists_kernel, dim3((ns + 255) / 256), dim3(256), 0, st,
                       du.as<uint64_t>(), ns, (int)norb, lists.as<uint8_t>(), cnt.as<int32_t>());
    DevBuf keys(8 * En), slots(4 * En), skeys(8 * En), sslots(4 * En), head(4 * En), rank1(4 * En);
    hipLaunchKernelGGL(str_table_kernel, dim3((unsigned)((En + 255) / 256)), dim3(256), 0, st,
                       du.as<uint64_t>(), dboff.as<int64_t>(), ns, lists.as<uint8_t>(),
                       cnt.as<int32_t>(), En, keys.as<uint64_t>(), slots.as<uint32_t>());
    MHIP_HIP_CHECK(hipGetLastError());
    size_t tmp = 0;
    const unsigned end_bit = W == 1 ? 32u : 64u;
    MHIP_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tmp, keys.as<uint64_t>(),
                                             skeys.as<uint64_t>(), slots.as<uint32_t>(),
                                             sslots.as<uint32_t>(), (size_t)En, 0u, end_bit, st));
    ctx->scratch_c.reserve(tmp + 16);
    MHIP_HIP_CHECK(rocprim::radix_sort_pairs(ctx->scratch_c.p, tmp, keys.as<uint64_t>(),
                                             skeys.as<uint64_t>(), slots.as<uint32_t>(),
                                             sslots.as<uint32_t>(), (size_t)En, 0u, end_bit, st));
    hipLaunchKernelGGL(str_head_kernel, dim3((unsigned)((En + 255) / 256)), dim3(256), 0, st,
                       skeys.as<uint64_t>(), En, head.as<int32_t>());
    tmp = 0;
    MHIP_HIP_CHECK(rocprim::inclusive_scan(nullptr, tmp, head.as<int32_t>(), rank1.as<int32_t>(),
                                           (size_t)En, rocprim::plus<int32_t>(), st));
    ctx->scratch_c.reserve(tmp + 16);
    MHIP_HIP_CHECK(rocprim::inclusive_scan(ctx->scratch_c.p, tmp, head.as<int32_t>(),
                                           rank1.as<int32_t>(), (size_t)En,
                                           rocprim::plus<int32_t>(), st));
    int32_t nu = 0;
    d2h(ctx, &nu, rank1.as<int32_t>() + En - 1, 4);
    DevBuf& table = sp ? R.tb : R.ta;
    DevBuf& uniq = sp ? R.ub : R.ua;
    DevBuf& offc = sp ? R.ob : R.oa;
    table.alloc(4 * En);
    uniq.alloc(8 * (int64_t)nu);
    hipLaunchKernelGGL(str_rank_kernel, dim3((unsigned)((En + 255) / 256)), dim3(256), 0, st,
                       skeys.as<uint64_t>(), sslots.as<uint32_t>(), rank1.as<int32_t>(), En,
                       table.as<uint32_t>(), uniq.as<uint64_t>());
    MHIP_HIP_CHECK(hipGetLastError());
    offc.alloc(8 * nc);
    MHIP_HIP_CHECK(hipMemcpyAsync(offc.p, S.off_c.data(), 8 * nc, hipMemcpyHostToDevice, st));
    int b = 1;
    while(b < 63 && (1ll << b) < (int64_t)nu) ++b;
    bits[sp] = b;
    // the host vectors and the temporaries above go out of scope: wait
    MHIP_HIP_CHECK(hipStreamSynchronize(st));
  }
  if(bits[0] + bits[1] > 64) return false;
  R.bits = bits[0] + bits[1];
  R.M.ta = R.ta.as<uint32_t>();
  R.M.tb = R.tb.as<uint32_t>();
  R.M.oa = R.oa.as<int64_t>();
  R.M.ob = R.ob.as<int64_t>();
  R.M.ua = R.ua.as<uint64_t>();
  R.M.ub = R.ub.as<uint64_t>();
  R.M.bits_a = bits[0];
  R.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return true;
}

// First block of every core det (bstart[nc] = number of blocks).
inline std::vector<int64_t> block_starts(const std::vector<int>& na, const std::vector<int>& nb,
                                         bool singles) {
  std::vector<int64_t> bs(na.size() + 1, 0);
  for(size_t c = 0; c < na.size(); ++c) bs[c + 1] = bs[c] + block_count(na[c], nb[c], singles);
  return bs;
}

template <int W, bool RANK>
int64_t search_impl(mhip_context* ctx, const mhip_asci_settings* s, int64_t ndets_max,
                    const uint64_t* d_cdets, int64_t nc, double E0, const double* d_C,
                    int64_t norb, uint64_t* d_out, const SearchMode& mode, const RankPrep* R) {
  using K = KeyT<W, RANK>;
  const unsigned kbits = RANK ? (unsigned)R->bits : 32u * W * 2u;
  const KeyMap M = RANK ? R->M : KeyMap{};
  hipStream_t st = ctx->stream;
  mhip_timings& tm = ctx->timings;
  Timer t_all(st), t(st);
  t_all.start();
  t.start();
  const IntsView I = device_view(ctx->ints);
  const bool screened = mode.screened();

  // candidate counts per core det (closed form of the loop bounds)
  std::vector<int> hna, hnb;
  core_counts<W>(ctx, d_cdets, nc, norb, hna, hnb);
  std::vector<int64_t> ncand(nc), cstart(nc + 1, 0);
  for(int64_t c = 0; c < nc; ++c) {
    const int64_t na = hna[c], nb = hnb[c];
    const int64_t nva = norb - na, nvb = norb - nb;
    int64_t k = na * nva + nb * nvb;
    if(!s->just_singles)
      k += host_pairs(na) * host_pairs(nva) + host_pairs(nb) * host_pairs(nvb) +
           na * nva * nb * nvb;
    ncand[c] = k;
    cstart[c + 1] = cstart[c] + (k + kChunk - 1) / kChunk;
  }
  const int64_t nchunks = screened ? 0 : cstart[nc];

  CorePrep core;
  prepare_core<W>(ctx, I, d_cdets, nc, norb, d_C, s->just_singles, core);
  CoreArrays& A = core.A;
  DevBuf dn(8 * nc), dstart(8 * (nc + 1));
  MHIP_HIP_CHECK(hipMemcpyAsync(dn.p, ncand.data(), 8 * nc, hipMemcpyHostToDevice, st));
  MHIP_HIP_CHECK(hipMemcpyAsync(dstart.p, cstart.data(), 8 * (nc + 1), hipMemcpyHostToDevice, st));
  A.chunk_start = dstart.as<int64_t>();

  DevBuf ccount, coff, keys, vals;
  size_t tmp = 0;
  int64_t P = 0;
  const int64_t grid = (nchunks + 3) / 4;
  tm.asci_nn_ms = 0.;
  tm.asci_nsurvivors = 0;
  tm.asci_nblocks = tm.asci_nblocks_kept = 0;
  if(!screened) {
    ccount.alloc(8 * (nchunks + 1));
    coff.alloc(8 * (nchunks + 1));
    MHIP_HIP_CHECK(hipMemsetAsync(ccount.p, 0, 8 * (nchunks + 1), st));
    MHIP_REQUIRE(grid < (1ll << 31), MHIP_ERR_INVALID_ARG, "too many candidates");
    hipLaunchKernelGGL((contrib_kernel<W, false, RANK>), dim3(grid), dim3(256), 0, st, I, A,
                       dn.as<int64_t>(), nchunks, E0, s->h_el_tol, ccount.as<int64_t>(), nullptr,
                       nullptr, nullptr, M);
    MHIP_HIP_CHECK(hipGetLastError());
    MHIP_HIP_CHECK(rocprim::exclusive_scan(nullptr, tmp, ccount.as<int64_t>(), coff.as<int64_t>(),
                                           (int64_t)0, (size_t)(nchunks + 1),
                                           rocprim::plus<int64_t>(), st));
    ctx->scratch_c.reserve(tmp + 16);
    MHIP_HIP_CHECK(rocprim::exclusive_scan(ctx->scratch_c.p, tmp, ccount.as<int64_t>(),
                                           coff.as<int64_t>(), (int64_t)0, (size_t)(nchunks + 1),
                                           rocprim::plus<int64_t>(), st));
    d2h(ctx, &P, coff.as<int64_t>() + nchunks, 8);
    keys.alloc(sizeof(K) * std::max<int64_t>(P, 1));
    vals.alloc(8 * std::max<int64_t>(P, 1));
    hipLaunchKernelGGL((contrib_kernel<W, true, RANK>), dim3(grid), dim3(256), 0, st, I, A,
                       dn.as<int64_t>(), nchunks, E0, s->h_el_tol, nullptr, coff.as<int64_t>(),
                       keys.as<K>(), vals.as<double>(), M);
    MHIP_HIP_CHECK(hipGetLastError());
  } else {
    // ---- block screen: score every excitation block, keep the best
    // keep_fraction of them (ties: lower block index first), generate only
    // their candidates
    Timer tn(st);
    tn.start();
    const std::vector<int64_t> bs = block_starts(hna, hnb, s->just_singles != 0);
    const int64_t nblk = bs[nc];
    DevBuf dbs(8 * (nc + 1));
    MHIP_HIP_CHECK(hipMemcpyAsync(dbs.p, bs.data(), 8 * (nc + 1), hipMemcpyHostToDevice, st));
    DevBuf bkey(4 * std::max<int64_t>(nblk, 1)), bidx(8 * std::max<int64_t>(nblk, 1));
    const unsigned bgrid = (unsigned)((nblk + 255) / 256);
    if(mode.mlp) {
      DevBuf bdets(8 * W * std::max<int64_t>(nblk, 1)), pred(4 * std::max<int64_t>(nblk, 1));
      if(nblk)
        hipLaunchKernelGGL(block_dets_kernel<W>, dim3(bgrid), dim3(256), 0, st, A,
                           dbs.as<int64_t>(), nblk, bdets.as<uint64_t>());
      MHIP_HIP_CHECK(hipGetLastError());
      const int prc =
          mhip_mlp_predict(ctx, mode.mlp, W, bdets.as<uint64_t>(), nblk, pred.as<float>());
      MHIP_REQUIRE(prc == MHIP_SUCCESS, prc, "block prediction failed");
      if(nblk)
        hipLaunchKernelGGL(block_key_kernel, dim3(bgrid), dim3(256), 0, st, pred.as<float>(),
                           mode.pred_scale, mode.pred_shift, d_C, dbs.as<int64_t>(), nc, nullptr,
                           nblk, bkey.as<uint32_t>(), bidx.as<int64_t>());
    } else if(nblk) {
      hipLaunchKernelGGL(block_key_kernel, dim3(bgrid), dim3(256), 0, st, nullptr, 1.f, 0.f,
                         d_C, dbs.as<int64_t>(), nc, mode.block_scores, nblk,
                         bkey.as<uint32_t>(), bidx.as<int64_t>());
    }
    MHIP_HIP_CHECK(hipGetLastError());
    const int64_t nkept = std::min<int64_t>(
        nblk, std::max<int64_t>(1, (int64_t)std::ceil(mode.keep_fraction * (double)nblk)));
    DevBuf kept(8 * std::max<int64_t>(nkept, 1));
    if(nblk) {
      DevBuf bkey2(4 * nblk), bidx2(8 * nblk);
      tmp = 0;
      MHIP_HIP_CHECK(rocprim::radix_sort_pairs_desc(nullptr, tmp, bkey.as<uint32_t>(),
                                                    bkey2.as<uint32_t>(), bidx.as<int64_t>(),
                                                    bidx2.as<int64_t>(), (size_t)nblk, 0, 32, st));
      ctx->scratch_c.reserve(tmp + 16);
      MHIP_HIP_CHECK(rocprim::radix_sort_pairs_desc(ctx->scratch_c.p, tmp, bkey.as<uint32_t>(),
                                                    bkey2.as<uint32_t>(), bidx.as<int64_t>(),
                                                    bidx2.as<int64_t>(), (size_t)nblk, 0, 32, st));
      // kept ids back in ascending order: the fill then runs in core-det order
      int bits = 1;
      while(bits < 63 && (int64_t(1) << bits) < nblk) ++bits;
      tmp = 0;
      MHIP_HIP_CHECK(rocprim::radix_sort_keys(nullptr, tmp, bidx2.as<int64_t>(),
                                              kept.as<int64_t>(), (size_t)nkept, 0, bits, st));
      ctx->scratch_c.reserve(tmp + 16);
      MHIP_HIP_CHECK(rocprim::radix_sort_keys(ctx->scratch_c.p, tmp, bidx2.as<int64_t>(),
                                              kept.as<int64_t>(), (size_t)nkept, 0, bits, st));
    }
    tn.stop();
    tm.asci_nn_ms = tn.ms();
    const int64_t kgrid = (nkept + 3) / 4;
    MHIP_REQUIRE(kgrid < (1ll << 31), MHIP_ERR_INVALID_ARG, "too many blocks");
    ccount.alloc(8 * (nkept + 1));
    coff.alloc(8 * (nkept + 1));
    MHIP_HIP_CHECK(hipMemsetAsync(ccount.p, 0, 8 * (nkept + 1), st));
    if(nblk)
      hipLaunchKernelGGL((block_contrib_kernel<W, false, RANK>), dim3(kgrid), dim3(256), 0, st,
                         I, A, dbs.as<int64_t>(), kept.as<int64_t>(), nkept, E0, s->h_el_tol,
                         ccount.as<int64_t>(), nullptr, nullptr, nullptr, M);
    MHIP_HIP_CHECK(hipGetLastError());
    tmp = 0;
    MHIP_HIP_CHECK(rocprim::exclusive_scan(nullptr, tmp, ccount.as<int64_t>(), coff.as<int64_t>(),
                                           (int64_t)0, (size_t)(nkept + 1),
                                           rocprim::plus<int64_t>(), st));
    ctx->scratch_c.reserve(tmp + 16);
    MHIP_HIP_CHECK(rocprim::exclusive_scan(ctx->scratch_c.p, tmp, ccount.as<int64_t>(),
                                           coff.as<int64_t>(), (int64_t)0, (size_t)(nkept + 1),
                                           rocprim::plus<int64_t>(), st));
    d2h(ctx, &P, coff.as<int64_t>() + nkept, 8);
    keys.alloc(sizeof(K) * std::max<int64_t>(P, 1));
    vals.alloc(8 * std::max<int64_t>(P, 1));
    if(nblk)
      hipLaunchKernelGGL((block_contrib_kernel<W, true, RANK>), dim3(kgrid), dim3(256), 0, st,
                         I, A, dbs.as<int64_t>(), kept.as<int64_t>(), nkept, E0, s->h_el_tol,
                         nullptr, coff.as<int64_t>(), keys.as<K>(), vals.as<double>(), M);
    MHIP_HIP_CHECK(hipGetLastError());
    tm.asci_nblocks = nblk;
    tm.asci_nblocks_kept = nkept;
  }
  t.stop();
  tm.asci_contrib_ms = t.ms() + (RANK ? R->ms : 0.);
  tm.asci_npairs = P;

  // ---- sort + accumulate
  t.start();
  DevBuf keys2(sizeof(K) * std::max<int64_t>(P, 1)), vals2(8 * std::max<int64_t>(P, 1));
  DevBuf ukeys(sizeof(K) * std::max<int64_t>(P, 1)), usum(8 * std::max<int64_t>(P, 1));
  DevBuf nu_dev(16);
  MHIP_HIP_CHECK(hipMemsetAsync(nu_dev.p, 0, 16, st));
  int64_t nu = 0;
  if(P > s->pair_size_max && nc > 0 && !screened) {
    // the reference's list would exceed pair_size_max: replay its prunes
    DevBuf rs(sizeof(RvSlot) * P), rs2(sizeof(RvSlot) * P), tiny(8 * nchunks);
    hipLaunchKernelGGL(prune_prep_kernel, dim3(grid), dim3(256), 0, st, nchunks,
                       coff.as<int64_t>(), vals.as<double>(), s->rv_prune_tol, rs.as<RvSlot>(),
                       tiny.as<int64_t>());
    MHIP_HIP_CHECK(hipGetLastError());
    tmp = 0;
    MHIP_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tmp, keys.as<K>(), keys2.as<K>(),
                                             rs.as<RvSlot>(), rs2.as<RvSlot>(), (size_t)P, 0,
                                             kbits, st));
    ctx->scratch_c.reserve(tmp + 16);
    MHIP_HIP_CHECK(rocprim::radix_sort_pairs(ctx->scratch_c.p, tmp, keys.as<K>(), keys2.as<K>(),
                                             rs.as<RvSlot>(), rs2.as<RvSlot>(), (size_t)P, 0,
                                             kbits, st));
    rs.release();
    // per core det: first slot, pairs and pairs with |rv| <= rv_prune_tol
    std::vector<int64_t> hcnt(nchunks), htiny(nchunks), dslot(nc + 1, 0), tdet(nc, 0);
    d2h(ctx, hcnt.data(), ccount.p, 8 * nchunks);
    d2h(ctx, htiny.data(), tiny.p, 8 * nchunks);
    for(int64_t c = 0; c < nc; ++c) {
      dslot[c + 1] = dslot[c];
      for(int64_t ch = cstart[c]; ch < cstart[c + 1]; ++ch) {
        dslot[c + 1] += hcnt[ch];
        tdet[c] += htiny[ch];
      }
    }
    const unsigned pgrid = (unsigned)((P + 255) / 256);
    DevBuf pos(8 * P), acc(8 * P), accf(P), alive(P), pcnt(16), slist(8 * P);
    hipLaunchKernelGGL(prune_pos_kernel, dim3(pgrid), dim3(256), 0, st, rs2.as<RvSlot>(), P,
                       pos.as<int64_t>());
    // segment head of every sorted position (slist serves as the scan input
    // before the events use it)
    DevBuf shead(8 * P);
    hipLaunchKernelGGL((seg_head_kernel<K>), dim3(pgrid), dim3(256), 0, st, keys2.as<K>(), P,
                       slist.as<int64_t>());
    tmp = 0;
    MHIP_HIP_CHECK(rocprim::inclusive_scan(nullptr, tmp, slist.as<int64_t>(), shead.as<int64_t>(),
                                           (size_t)P, rocprim::maximum<int64_t>(), st));
    ctx->scratch_c.reserve(tmp + 16);
    MHIP_HIP_CHECK(rocprim::inclusive_scan(ctx->scratch_c.p, tmp, slist.as<int64_t>(),
                                           shead.as<int64_t>(), (size_t)P,
                                           rocprim::maximum<int64_t>(), st));
    MHIP_HIP_CHECK(hipMemsetAsync(accf.p, 0, P, st));
    MHIP_HIP_CHECK(hipMemsetAsync(alive.p, 1, P, st));
    // S: the reference's list length; nseg: accumulated entries (stage 2);
    // pending_small: small accumulated entries the next stage 1 erases;
    // marks / pending_list: those entries (listed in slist) not yet erased
    // on the device
    int64_t S = 0, tiny_since = 0, pending_small = 0, nseg = 0, pending_list = 0;
    int64_t last_event = -1, last_s2 = -1;
    bool marks = false;
    auto drop_marked = [&] {
      hipLaunchKernelGGL((prune_drop_acc_kernel<K>), dim3((unsigned)((pending_list + 255) / 256)),
                         dim3(256), 0, st, keys2.as<K>(), rs2.as<RvSlot>(), P, dslot[last_s2 + 1],
                         slist.as<int64_t>(), pending_list, accf.as<uint8_t>(),
                         alive.as<uint8_t>());
      MHIP_HIP_CHECK(hipGetLastError());
      nseg -= pending_list;
      pending_list = 0;
      marks = false;
    };
    const char* stream_env = std::getenv("MHIP_PRUNE_STREAM");
    for(int64_t c = 0; c < nc; ++c) {
      S += dslot[c + 1] - dslot[c];
      tiny_since += tdet[c];
      if(S <= s->pair_size_max) continue;
      S -= tiny_since + pending_small;  // stage 1 (:123-128)
      tiny_since = pending_small = 0;
      if(S > s->pair_size_max) {  // stage 2 (:131-135)
        if(marks) drop_marked();
        const int64_t lim0 = dslot[last_s2 + 1], lim1 = dslot[c + 1];
        MHIP_HIP_CHECK(hipMemsetAsync(pcnt.p, 0, 16, st));
        // an event over more than 1/16 of the pairs streams the sorted list
        // (24 bytes a pair) instead of three random lines a slot
        // (MHIP_PRUNE_STREAM=0 / 1: never / always, for tests)
        const bool stream = stream_env ? stream_env[0] == '1' : 16 * (lim1 - lim0) > P;
        if(lim1 > lim0 && stream)
          hipLaunchKernelGGL((prune_stage2_kernel<K, true>),
                             dim3((unsigned)std::min<int64_t>(kStage2Blocks, (P + 255) / 256)),
                             dim3(256), 0, st, keys2.as<K>(), rs2.as<RvSlot>(), P,
                             pos.as<int64_t>(), shead.as<int64_t>(), lim0, lim1,
                             s->rv_prune_tol, acc.as<double>(),
                             accf.as<uint8_t>(), alive.as<uint8_t>(),
                             pcnt.as<unsigned long long>(), slist.as<int64_t>());
        else if(lim1 > lim0)
          hipLaunchKernelGGL((prune_stage2_kernel<K, false>),
                             dim3((unsigned)std::min<int64_t>(kStage2Blocks,
                                                              (lim1 - lim0 + 255) / 256)),
                             dim3(256), 0, st,
                             keys2.as<K>(), rs2.as<RvSlot>(), P, pos.as<int64_t>(),
                             shead.as<int64_t>(), lim0, lim1, s->rv_prune_tol,
                             acc.as<double>(), accf.as<uint8_t>(),
                             alive.as<uint8_t>(), pcnt.as<unsigned long long>(),
                             slist.as<int64_t>());
        MHIP_HIP_CHECK(hipGetLastError());
        int64_t h2[2] = {0, 0};
        d2h(ctx, h2, pcnt.p, 16);
        nseg += h2[0];
        S = nseg;
        pending_small = pending_list = h2[1];
        marks = h2[1] > 0;
        last_s2 = c;
      }
      last_event = c;
    }
    // marks of the last stage 2 are erased only if a prune followed it
    if(marks && last_event > last_s2) drop_marked();
    hipLaunchKernelGGL(prune_final_kernel, dim3(pgrid), dim3(256), 0, st, rs2.as<RvSlot>(),
                       alive.as<uint8_t>(), P, dslot[last_s2 + 1], dslot[last_event + 1],
                       s->rv_prune_tol, vals2.as<double>());
    MHIP_HIP_CHECK(hipGetLastError());
  } else if(P) {
    tmp = 0;
    MHIP_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tmp, keys.as<K>(), keys2.as<K>(),
                                             vals.as<double>(), vals2.as<double>(), (size_t)P,
                                             0, kbits, st));
    ctx->scratch_c.reserve(tmp + 16);
    MHIP_HIP_CHECK(rocprim::radix_sort_pairs(ctx->scratch_c.p, tmp, keys.as<K>(), keys2.as<K>(),
                                             vals.as<double>(), vals2.as<double>(), (size_t)P,
                                             0, kbits, st));
  }
  if(P) {
    tmp = 0;
    MHIP_HIP_CHECK(rocprim::reduce_by_key(nullptr, tmp, keys2.as<K>(), vals2.as<double>(),
                                          (size_t)P, ukeys.as<K>(), usum.as<double>(),
                                          nu_dev.as<int64_t>(), rocprim::plus<double>(),
                                          rocprim::equal_to<K>(), st));
    ctx->scratch_c.reserve(tmp + 16);
    MHIP_HIP_CHECK(rocprim::reduce_by_key(ctx->scratch_c.p, tmp, keys2.as<K>(),
                                          vals2.as<double>(), (size_t)P, ukeys.as<K>(),
                                          usum.as<double>(), nu_dev.as<int64_t>(),
                                          rocprim::plus<double>(), rocprim::equal_to<K>(), st));
    d2h(ctx, &nu, nu_dev.p, 8);
  }
  t.stop();
  tm.asci_sort_ms = t.ms();
  tm.asci_nunique = nu;
  keys.release();
  vals.release();

  // ---- exclude core dets and zero scores, top-k
  t.start();
  DevBuf core_keys(sizeof(K) * std::max<int64_t>(nc, 1)), core_sorted(sizeof(K) * std::max<int64_t>(nc, 1));
  if constexpr(RANK)
    hipLaunchKernelGGL(core_rank_keys_kernel, dim3((nc + 255) / 256), dim3(256), 0, st, M, nc,
                       core_keys.as<K>());
  else
    hipLaunchKernelGGL((dets_to_keys_kernel<W, K>), dim3((nc + 255) / 256), dim3(256), 0, st,
                       d_cdets, nc, core_keys.as<K>());
  tmp = 0;
  MHIP_HIP_CHECK(rocprim::radix_sort_keys(nullptr, tmp, core_keys.as<K>(), core_sorted.as<K>(),
                                          (size_t)nc, 0, kbits, st));
  ctx->scratch_c.reserve(tmp + 16);
  MHIP_HIP_CHECK(rocprim::radix_sort_keys(ctx->scratch_c.p, tmp, core_keys.as<K>(),
                                          core_sorted.as<K>(), (size_t)nc, 0, kbits, st));
  const int64_t top_k = ndets_max - nc;
  if(!mode.exp_dets) {
    // Plain and block-screened search: exclusion and top-k straight on the accumulated arrays.
    // The k-th largest |score| is bracketed by two 12-bit digit histograms
    // (24-bit prefix); only candidates at or above that prefix (in
    // determinant order) are selected, and a stable descending sort of them
    // gives the same first k as sorting all candidates (ties keep
    // determinant order, as the reference's comparison does).
    double* us = usum.as<double>();
    if(nu && nc)
      hipLaunchKernelGGL((core_zero_kernel<K>), dim3((nc + 255) / 256), dim3(256), 0, st,
                         ukeys.as<K>(), nu, core_sorted.as<K>(), nc, us);
    DevBuf hist(sizeof(uint32_t) * 4096);
    std::vector<uint32_t> hh(4096);
    const unsigned hblocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(1024, (nu + 255) / 256));
    uint64_t prefix = 0;
    int64_t above = 0, in_bucket = 0, eligible = 0;
    const int npass = top_k > 0 ? 2 : 0;
    for(int pass = 0; pass < std::max(npass, 1); ++pass) {
      const int shift = pass == 0 ? 52 : 40;
      MHIP_HIP_CHECK(hipMemsetAsync(hist.p, 0, sizeof(uint32_t) * 4096, st));
      if(nu)
        hipLaunchKernelGGL(score_hist_kernel, dim3(hblocks), dim3(256), 0, st, us, nu, shift,
                           prefix, hist.as<uint32_t>());
      MHIP_HIP_CHECK(hipGetLastError());
      d2h(ctx, hh.data(), hist.p, sizeof(uint32_t) * 4096);
      if(pass == 0)
        for(uint32_t c : hh) eligible += c;
      if(eligible <= top_k || top_k <= 0) break;  // everything fits (or nothing is asked)
      int d = 4095;
      while(d > 0 && above + (int64_t)hh[d] < top_k) above += hh[d--];
      prefix = (prefix << 12) | (uint64_t)d;
      in_bucket = hh[d];
    }
    tm.asci_nsurvivors = eligible;
    int64_t nout = 0;
    DevBuf cnt(16), sk, sv, ab, skeys, sabs;
    if(top_k > 0 && eligible > 0) {
      const bool all = eligible <= top_k;
      const uint64_t thr = all ? 0 : prefix;
      const int64_t nsel_expected = all ? eligible : above + in_bucket;
      // every candidate at or above the prefix, compacted unordered, then
      // sorted by key (determinant order) with its score
      DevBuf sku(sizeof(K) * std::max<int64_t>(nsel_expected, 1));
      DevBuf svu(8 * std::max<int64_t>(nsel_expected, 1));
      sk.alloc(sizeof(K) * std::max<int64_t>(nsel_expected, 1));
      sv.alloc(8 * std::max<int64_t>(nsel_expected, 1));
      MHIP_HIP_CHECK(hipMemsetAsync(cnt.p, 0, 16, st));
      hipLaunchKernelGGL((select_above_kernel<K>),
                         dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(4096, (nu + 255) / 256))),
                         dim3(256), 0, st, ukeys.as<K>(), us, nu, thr, nsel_expected, sku.as<K>(),
                         svu.as<double>(), cnt.as<unsigned long long>());
      MHIP_HIP_CHECK(hipGetLastError());
      int64_t nsel = 0;
      d2h(ctx, &nsel, cnt.p, 8);
      MHIP_REQUIRE(nsel == nsel_expected, MHIP_ERR_HIP, "top-k selection count mismatch");
      tmp = 0;
      MHIP_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tmp, sku.as<K>(), sk.as<K>(),
                                               svu.as<double>(), sv.as<double>(), (size_t)nsel, 0,
                                               kbits, st));
      ctx->scratch_c.reserve(tmp + 16);
      MHIP_HIP_CHECK(rocprim::radix_sort_pairs(ctx->scratch_c.p, tmp, sku.as<K>(), sk.as<K>(),
                                               svu.as<double>(), sv.as<double>(), (size_t)nsel, 0,
                                               kbits, st));
      const K* outk = sk.as<K>();
      nout = nsel;
      if(!all) {
        ab.alloc(8 * nsel);
        hipLaunchKernelGGL(abs_bits_kernel, dim3((nsel + 255) / 256), dim3(256), 0, st,
                           sv.as<double>(), nsel, ab.as<uint64_t>());
        skeys.alloc(sizeof(K) * nsel);
        sabs.alloc(8 * nsel);
        tmp = 0;
        MHIP_HIP_CHECK(rocprim::radix_sort_pairs_desc(nullptr, tmp, ab.as<uint64_t>(),
                                                      sabs.as<uint64_t>(), sk.as<K>(),
                                                      skeys.as<K>(), (size_t)nsel, 0, 64, st));
        ctx->scratch_c.reserve(tmp + 16);
        MHIP_HIP_CHECK(rocprim::radix_sort_pairs_desc(ctx->scratch_c.p, tmp, ab.as<uint64_t>(),
                                                      sabs.as<uint64_t>(), sk.as<K>(),
                                                      skeys.as<K>(), (size_t)nsel, 0, 64, st));
        outk = skeys.as<K>();
        nout = top_k;
      }
      if constexpr(RANK)
        hipLaunchKernelGGL(rank_keys_to_dets_kernel<W>, dim3((nout + 255) / 256), dim3(256), 0,
                           st, M, outk, nout, d_out);
      else
        hipLaunchKernelGGL((keys_to_dets_kernel<W, K>), dim3((nout + 255) / 256), dim3(256), 0,
                           st, outk, nout, d_out);
    }
    if(nc)
      MHIP_HIP_CHECK(hipMemcpyAsync(d_out + W * nout, d_cdets, 8 * W * nc,
                                    hipMemcpyDeviceToDevice, st));
    MHIP_HIP_CHECK(hipGetLastError());
    t.stop();
    tm.asci_topk_ms = t.ms();
    t_all.stop();
    tm.asci_total_ms = t_all.ms() + (RANK ? R->ms : 0.);
    return nout + nc;
  }
  int64_t nkeep = 0;
  DevBuf flags(std::max<int64_t>(nu, 1)), absbits(8 * std::max<int64_t>(nu, 1));
  DevBuf kkeys(sizeof(K) * std::max<int64_t>(nu, 1)), kabs(8 * std::max<int64_t>(nu, 1));
  if(nu) {
    hipLaunchKernelGGL((flag_kernel<K>), dim3(std::min<int64_t>(4096, (nu + 255) / 256)),
                       dim3(256), 0, st, ukeys.as<K>(), usum.as<double>(), nu_dev.as<int64_t>(),
                       core_sorted.as<K>(), nc, flags.as<uint8_t>(), absbits.as<uint64_t>());
    tmp = 0;
    MHIP_HIP_CHECK(rocprim::select(nullptr, tmp, ukeys.as<K>(), flags.as<uint8_t>(), kkeys.as<K>(),
                                   nu_dev.as<int64_t>() + 1, (size_t)nu, st));
    ctx->scratch_c.reserve(tmp + 16);
    MHIP_HIP_CHECK(rocprim::select(ctx->scratch_c.p, tmp, ukeys.as<K>(), flags.as<uint8_t>(),
                                   kkeys.as<K>(), nu_dev.as<int64_t>() + 1, (size_t)nu, st));
    MHIP_HIP_CHECK(rocprim::select(ctx->scratch_c.p, tmp, absbits.as<uint64_t>(),
                                   flags.as<uint8_t>(), kabs.as<uint64_t>(),
                                   nu_dev.as<int64_t>() + 1, (size_t)nu, st));
    d2h(ctx, &nkeep, nu_dev.as<int64_t>() + 1, 8);
  }
  // export of the scored candidates
  tm.asci_nsurvivors = nkeep;
  MHIP_REQUIRE(nkeep <= mode.exp_cap, MHIP_ERR_INVALID_ARG, "candidate export capacity too small");
  if(nkeep) {
    if constexpr(RANK)
      hipLaunchKernelGGL(rank_keys_to_dets_kernel<W>, dim3((nkeep + 255) / 256), dim3(256), 0, st,
                         M, kkeys.as<K>(), nkeep, mode.exp_dets);
    else
      hipLaunchKernelGGL((keys_to_dets_kernel<W, K>), dim3((nkeep + 255) / 256), dim3(256), 0,
                         st, kkeys.as<K>(), nkeep, mode.exp_dets);
    hipLaunchKernelGGL(bits_to_double_kernel, dim3((nkeep + 255) / 256), dim3(256), 0, st,
                       kabs.as<uint64_t>(), nkeep, mode.exp_abs);
  }
  MHIP_HIP_CHECK(hipGetLastError());
  MHIP_HIP_CHECK(hipStreamSynchronize(st));
  *mode.exp_n = nkeep;
  t.stop();
  tm.asci_topk_ms = t.ms();
  t_all.stop();
  tm.asci_total_ms = t_all.ms() + (RANK ? R->ms : 0.);
  return 0;
}

// Dense keys when their tables are small against the candidates, else
// keys of the full strings (same result).
template <int W>
int64_t search(mhip_context* ctx, const mhip_asci_settings* s, int64_t ndets_max,
               const uint64_t* d_cdets, int64_t nc, double E0, const double* d_C, int64_t norb,
               uint64_t* d_out, const SearchMode& mode) {
  RankPrep R;
  bool ranks = false;
  try {
    ranks = prepare_ranks<W>(ctx, d_cdets, nc, norb, s->just_singles != 0,
                             mode.screened() ? std::min(1.0, mode.keep_fraction) : 1.0, R);
  } catch(const Error& e) {  // out of memory for the tables: full-string keys
    if(e.code != MHIP_ERR_OUT_OF_MEMORY) throw;
    (void)hipGetLastError();
    R = RankPrep();
  }
  if(ranks)
    return search_impl<W, true>(ctx, s, ndets_max, d_cdets, nc, E0, d_C, norb, d_out, mode, &R);
  return search_impl<W, false>(ctx, s, ndets_max, d_cdets, nc, E0, d_C, norb, d_out, mode,
                               nullptr);
}

}  // namespace
}  // namespace mhip

using namespace mhip;

extern "C" int mhip_asci_search(mhip_context* ctx, int words_per_det,
                                const mhip_asci_settings* settings,
                                int64_t ndets_max, const uint64_t* d_cdets,
                                int64_t ncdets, double E_ASCI, const double* d_C,
                                int64_t norb, uint64_t* d_new_dets,
                                int64_t* n_new) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(settings && n_new, MHIP_ERR_INVALID_ARG, "null argument");
    MHIP_REQUIRE(words_per_det == 1 || words_per_det == 2, MHIP_ERR_UNSUPPORTED,
                 "words_per_det must be 1 or 2");
    MHIP_REQUIRE(ctx->ints.norb > 0, MHIP_ERR_NO_INTEGRALS, "no integrals set");
    MHIP_REQUIRE(norb > 0 && norb <= ctx->ints.norb && norb <= 32 * words_per_det,
                 MHIP_ERR_INVALID_ARG, "bad norb");
    MHIP_REQUIRE(ncdets >= 0 && ndets_max >= ncdets, MHIP_ERR_INVALID_ARG,
                 "ndets_max must be >= ncdets");
    MHIP_REQUIRE(ncdets == 0 || (d_cdets && d_C && d_new_dets), MHIP_ERR_INVALID_ARG,
                 "null pointer");
    *n_new = 0;
    if(!ncdets) return;
    if(words_per_det == 1)
      *n_new = search<1>(ctx, settings, ndets_max, d_cdets, ncdets, E_ASCI, d_C, norb,
                              d_new_dets, SearchMode{});
    else
      *n_new = search<2>(ctx, settings, ndets_max, d_cdets, ncdets, E_ASCI, d_C, norb,
                              d_new_dets, SearchMode{});
  });
}

namespace {
int asci_entry(mhip_context* ctx, int W, const mhip_asci_settings* settings, int64_t ndets_max,
               const uint64_t* d_cdets, int64_t ncdets, double E_ASCI, const double* d_C,
               int64_t norb, uint64_t* d_out, int64_t* n_new, const SearchMode& mode) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(settings && n_new, MHIP_ERR_INVALID_ARG, "null argument");
    MHIP_REQUIRE(W == 1 || W == 2, MHIP_ERR_UNSUPPORTED, "words_per_det must be 1 or 2");
    MHIP_REQUIRE(ctx->ints.norb > 0, MHIP_ERR_NO_INTEGRALS, "no integrals set");
    MHIP_REQUIRE(norb > 0 && norb <= ctx->ints.norb && norb <= 32 * W, MHIP_ERR_INVALID_ARG,
                 "bad norb");
    MHIP_REQUIRE(ncdets > 0 && ndets_max >= ncdets && d_cdets && d_C, MHIP_ERR_INVALID_ARG,
                 "bad core determinants");
    *n_new = 0;
    if(W == 1)
      *n_new = search<1>(ctx, settings, ndets_max, d_cdets, ncdets, E_ASCI, d_C, norb, d_out,
                              mode);
    else
      *n_new = search<2>(ctx, settings, ndets_max, d_cdets, ncdets, E_ASCI, d_C, norb, d_out,
                              mode);
  });
}
}  // namespace

extern "C" int mhip_asci_candidates(mhip_context* ctx, int words_per_det,
                                    const mhip_asci_settings* settings, const uint64_t* d_cdets,
                                    int64_t ncdets, double E_ASCI, const double* d_C,
                                    int64_t norb, uint64_t* d_dets_out, double* d_absrv_out,
                                    int64_t capacity, int64_t* n_out) {
  SearchMode m;
  m.exp_dets = d_dets_out;
  m.exp_abs = d_absrv_out;
  m.exp_cap = capacity;
  m.exp_n = n_out;
  int64_t dummy = 0;
  if(!n_out || !d_dets_out || !d_absrv_out) {
    set_last_error("null output");
    return MHIP_ERR_INVALID_ARG;
  }
  return asci_entry(ctx, words_per_det, settings, ncdets, d_cdets, ncdets, E_ASCI, d_C, norb,
                    nullptr, &dummy, m);
}

extern "C" int mhip_asci_search_screened(mhip_context* ctx, int words_per_det,
                                         const mhip_asci_settings* settings, mhip_mlp* mlp,
                                         float pred_scale, float pred_shift, double keep_fraction,
                                         int64_t ndets_max, const uint64_t* d_cdets,
                                         int64_t ncdets, double E_ASCI, const double* d_C,
                                         int64_t norb, uint64_t* d_new_dets, int64_t* n_new) {
  SearchMode m;
  m.mlp = mlp;
  m.pred_scale = pred_scale;
  m.pred_shift = pred_shift;
  m.keep_fraction = keep_fraction;
  if(!mlp || !(keep_fraction > 0. && keep_fraction <= 1.) || !d_new_dets ||
     !std::isfinite(pred_scale) || !std::isfinite(pred_shift)) {
    set_last_error(
        "mhip_asci_search_screened: need a network, finite scale/shift, 0 < keep_fraction <= 1, "
        "an output");
    return MHIP_ERR_INVALID_ARG;
  }
  return asci_entry(ctx, words_per_det, settings, ndets_max, d_cdets, ncdets, E_ASCI, d_C, norb,
                    d_new_dets, n_new, m);
}

extern "C" int mhip_asci_search_blocks(mhip_context* ctx, int words_per_det,
                                       const mhip_asci_settings* settings,
                                       const float* d_block_scores, double keep_fraction,
                                       int64_t ndets_max, const uint64_t* d_cdets, int64_t ncdets,
                                       double E_ASCI, const double* d_C, int64_t norb,
                                       uint64_t* d_new_dets, int64_t* n_new) {
  SearchMode m;
  m.block_scores = d_block_scores;
  m.keep_fraction = keep_fraction;
  if(!d_block_scores || !(keep_fraction > 0. && keep_fraction <= 1.) || !d_new_dets) {
    set_last_error("mhip_asci_search_blocks: need block scores, 0 < keep_fraction <= 1, an output");
    return MHIP_ERR_INVALID_ARG;
  }
  return asci_entry(ctx, words_per_det, settings, ndets_max, d_cdets, ncdets, E_ASCI, d_C, norb,
                    d_new_dets, n_new, m);
}

namespace {
template <int W>
int64_t blocks_impl(mhip_context* ctx, const mhip_asci_settings* s, const uint64_t* d_cdets,
                    int64_t nc, double E0, int64_t norb, uint64_t* d_dets, float* d_target,
                    int64_t capacity, int64_t* n_out) {
  hipStream_t st = ctx->stream;
  const IntsView I = device_view(ctx->ints);
  std::vector<int> hna, hnb;
  core_counts<W>(ctx, d_cdets, nc, norb, hna, hnb);
  const std::vector<int64_t> bs = block_starts(hna, hnb, s->just_singles != 0);
  const int64_t nblk = bs[nc];
  *n_out = nblk;  // reported also when the capacity check below fails
  if(!d_dets && !d_target) return nblk;
  MHIP_REQUIRE(nblk <= capacity, MHIP_ERR_INVALID_ARG, "block export capacity too small");
  if(!nblk) return 0;
  DevBuf ones(8 * nc);
  hipLaunchKernelGGL(fill_ones_kernel, dim3((nc + 255) / 256), dim3(256), 0, st,
                     ones.as<double>(), nc);
  CorePrep core;
  prepare_core<W>(ctx, I, d_cdets, nc, norb, ones.as<double>(), s->just_singles, core);
  DevBuf dbs(8 * (nc + 1));
  MHIP_HIP_CHECK(hipMemcpyAsync(dbs.p, bs.data(), 8 * (nc + 1), hipMemcpyHostToDevice, st));
  if(d_dets)
    hipLaunchKernelGGL(block_dets_kernel<W>, dim3((unsigned)((nblk + 255) / 256)), dim3(256), 0,
                       st, core.A, dbs.as<int64_t>(), nblk, d_dets);
  const int64_t grid = (nblk + 3) / 4;
  MHIP_REQUIRE(grid < (1ll << 31), MHIP_ERR_INVALID_ARG, "too many blocks");
  if(d_target)
    hipLaunchKernelGGL(block_target_kernel, dim3((unsigned)grid), dim3(256), 0, st, I, core.A,
                       dbs.as<int64_t>(), nblk, E0, s->h_el_tol, d_target);
  MHIP_HIP_CHECK(hipGetLastError());
  MHIP_HIP_CHECK(hipStreamSynchronize(st));
  return nblk;
}
}  // namespace

extern "C" int mhip_asci_blocks(mhip_context* ctx, int words_per_det,
                                const mhip_asci_settings* settings, const uint64_t* d_cdets,
                                int64_t ncdets, double E_ASCI, int64_t norb,
                                uint64_t* d_block_dets, float* d_target, int64_t capacity,
                                int64_t* n_blocks) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(settings && n_blocks, MHIP_ERR_INVALID_ARG, "null argument");
    MHIP_REQUIRE(words_per_det == 1 || words_per_det == 2, MHIP_ERR_UNSUPPORTED,
                 "words_per_det must be 1 or 2");
    MHIP_REQUIRE(ctx->ints.norb > 0, MHIP_ERR_NO_INTEGRALS, "no integrals set");
    MHIP_REQUIRE(norb > 0 && norb <= ctx->ints.norb && norb <= 32 * words_per_det,
                 MHIP_ERR_INVALID_ARG, "bad norb");
    MHIP_REQUIRE(ncdets >= 0 && (ncdets == 0 || d_cdets), MHIP_ERR_INVALID_ARG,
                 "bad core determinants");
    *n_blocks = 0;
    if(!ncdets) return;
    if(words_per_det == 1)
      *n_blocks = blocks_impl<1>(ctx, settings, d_cdets, ncdets, E_ASCI, norb, d_block_dets,
                                 d_target, capacity, n_blocks);
    else
      *n_blocks = blocks_impl<2>(ctx, settings, d_cdets, ncdets, E_ASCI, norb, d_block_dets,
                                 d_target, capacity, n_blocks);
  });
}
