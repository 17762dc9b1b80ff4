// Structure passes of the Hamiltonian build for product-space determinant
// lists (complete active spaces).
//
// The reference builds every CSR with the same serial double loop over
// (bra, ket) pairs (double_loop.hpp:45-91): a ket is a column when
// popcount(bra ^ ket) <= 4. A list produced by generate_hilbert_space
// (sd_operations.hpp:460-476), or the same list sorted with bitset_less, is
// a product of spin strings: det r*L + k carries outer string O[r] and inner
// string I[k], every run of L consecutive determinants shares its outer
// string and all runs carry the same inner-string sequence. Then
//   popcount(bra ^ ket) = popcount(O[r] ^ O[r']) + popcount(I[k] ^ I[k'])
// and the columns of row r*L + k are, for every outer neighbour r' at
// excitation level e = popcount(O[r] ^ O[r']) / 2 <= 2 (r' ascending), the
// inner neighbours k' of k at level <= 2 - e (k' ascending), at column
// r'*L + k'. The count pass is closed form; the fill pass writes those
// segments and evaluates their Slater-Condon elements with the same device
// function and diagonal as the tiled path's element pass, and counts
// |h| > thresh for the shared compaction, so both paths give the identical
// CSR.
//
// Recognition: the first run length L is read from the head of the list, a
// device pass checks outer(d[i]) == O[i / L] and inner(d[i]) == I[i % L] for
// every i, and the host checks that the O and I strings are distinct (a
// list with repeated determinants goes to the tiled path). The neighbour
// tables come from XOR/popcount over the string pairs on the host.
#include <algorithm>
#include <array>
#include <memory>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "mhip_device.hpp"
#include "mhip_internal.hpp"

namespace mhip {
namespace {

constexpr int kWavesPerBlock = 4;
constexpr int64_t kMaxHead = 1 << 16;      // determinants read to find L
constexpr int kMaxSeg = 512;  // outer neighbours of a run staged in LDS (more: tiled path)
constexpr double kMaxPairTests = 4.0e8;    // host neighbour-table budget

template <int W>
__device__ __forceinline__ void split(const uint64_t* d, int64_t i, uint64_t& a, uint64_t& b) {
  if constexpr(W == 1) {
    const uint64_t w = d[i];
    a = w & 0xffffffffull;
    b = w >> 32;
  } else {
    a = d[2 * i];
    b = d[2 * i + 1];
  }
}

}  // namespace

// Device view of one product space with its element tables and integrals.
// A single build passes one by value; a batched build keeps one per
// instance (row0 = the instance's first row, columns offset by it).
struct PSView {
  IntsView I;
  int64_t row0, L, nruns;
  int outer_alpha;
  const uint64_t *ostr, *istr;
  const int32_t *ocnt, *onb_ptr, *onb, *ptr1, *idx1, *ptr2, *idx2;
  double *oval, *ival2;
  int32_t *oexc, *iexc1, *iexc2;
};

namespace {

// Instance of global item g for a prefix array off[0..ninst] (off[0] = 0).
__device__ __forceinline__ int inst_of(const int64_t* __restrict__ off, int ninst, int64_t g) {
  int lo = 0, hi = ninst;  // last i with off[i] <= g
  while(hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if(off[mid] <= g)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Flags any determinant that breaks the product layout.
template <int W>
__global__ __launch_bounds__(256) void product_check_kernel(const uint64_t* __restrict__ d,
                                                            int64_t n, int64_t L,
                                                            bool outer_alpha,
                                                            int* __restrict__ bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  uint64_t a, b, a0, b0, a1, b1;
  split<W>(d, i, a, b);
  split<W>(d, (i / L) * L, a0, b0);  // run head: outer string of the run
  split<W>(d, i % L, a1, b1);        // first run: inner string of the column
  const uint64_t o = outer_alpha ? a : b, in = outer_alpha ? b : a;
  const uint64_t o_ref = outer_alpha ? a0 : b0, in_ref = outer_alpha ? b1 : a1;
  if(o != o_ref || in != in_ref) atomicOr(bad, 1);
}

// Outer string of every run.
template <int W>
__global__ void run_heads_kernel(const uint64_t* __restrict__ d, int64_t nruns, int64_t L,
                                 bool outer_alpha, uint64_t* __restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(r >= nruns) return;
  uint64_t a, b;
  split<W>(d, r * L, a, b);
  out[r] = outer_alpha ? a : b;
}

// counts[r*L+k] = n0(r) m2(k) + n1(r) m1(k) + n2(r)   (m0 = n0 = 1)
// (batched: the row's instance view; rows of one instance are contiguous)
template <bool BATCH>
__global__ void product_count_kernel(PSView one, const PSView* __restrict__ many,
                                     const int32_t* __restrict__ row_inst, int64_t n,
                                     uint32_t* __restrict__ counts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const PSView& P = BATCH ? many[row_inst[i]] : one;
  const int64_t li = i - P.row0;
  const int64_t r = li / P.L, k = li % P.L;
  const uint32_t m1 = (uint32_t)(P.ptr1[k + 1] - P.ptr1[k]);
  const uint32_t m2 = (uint32_t)(P.ptr2[k + 1] - P.ptr2[k]);
  counts[i] = (uint32_t)P.ocnt[3 * r] * m2 + (uint32_t)P.ocnt[3 * r + 1] * m1 +
              (uint32_t)P.ocnt[3 * r + 2];
}

// Excitation of a single (o, v, sign) packed in an int32: o | v << 8 |
// (sign < 0) << 16, with the hole/particle/sign of me_2 (ket & ex, bra & ex).
__device__ __forceinline__ int32_t pack_single(uint64_t bra, uint64_t ket) {
  const uint64_t ex = bra ^ ket;
  const uint32_t o = ctz64(ket & ex), v = ctz64(bra & ex);
  const double sg = excitation_sign(ket, v, o);
  return (int32_t)(o | (v << 8) | (sg < 0. ? (1u << 16) : 0u));
}

// T(v,o) + sum over the bra's same-spin occupied p of G_red(p,v,o): the
// first two terms of me_2 in its order (matrix_elements.hpp:86-105); the
// other-spin V_red terms are added later by the fill pass in the same order.
__device__ __forceinline__ double single_partial(const IntsView& I, uint64_t bra_same,
                                                 int32_t code) {
  const int n = I.n;
  const uint32_t o = code & 0xff, v = (code >> 8) & 0xff;
  double h = I.T[v + n * o];
  const double* G_ov = I.G_red + n * (v + n * o);
  for(uint64_t m = bra_same; m; m &= m - 1) h += G_ov[ctz64(m)];
  return h;
}

// Per outer-neighbour entry: level 2 -> the same-spin double element
// (me_4 of the outer strings); level 1 -> packed single and its partial sum.
// (batched: wavefront g over the runs of all instances, run_off = prefix of
// the instances' run counts)
template <bool BATCH>
__global__ __launch_bounds__(256) void outer_tables_kernel(PSView one,
                                                           const PSView* __restrict__ many,
                                                           const int64_t* __restrict__ run_off,
                                                           int ninst, int64_t total) {
  const int64_t g = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if(g >= total) return;
  const int inst = BATCH ? inst_of(run_off, ninst, g) : 0;
  const PSView& P = BATCH ? many[inst] : one;
  const int64_t r = BATCH ? g - run_off[inst] : g;
  const IntsView& I = P.I;
  const uint64_t* __restrict__ ostr = P.ostr;
  const int32_t* __restrict__ onb_ptr = P.onb_ptr;
  const int32_t* __restrict__ onb = P.onb;
  double* __restrict__ oval = P.oval;
  int32_t* __restrict__ oexc = P.oexc;
  const int lane = threadIdx.x & 63;
  const uint64_t b = ostr[r];
  for(int32_t j = onb_ptr[r] + lane; j < onb_ptr[r + 1]; j += 64) {
    const int32_t packed = onb[j];
    const uint64_t kt = ostr[packed >> 2];
    const int e = packed & 3;
    double val = 0.;
    int32_t code = 0;
    if(e == 2) {
      val = me_4(I, b, kt, b ^ kt);
    } else if(e == 1) {
      code = pack_single(b, kt);
      val = single_partial(I, b, code);
    }
    oval[j] = val;
    oexc[j] = code;
  }
}

// Per inner-list entry of string k: the level-<=2 list gets the same-spin
// double element (level 2) or the packed single and its partial sum
// (level 1), with the level in bits 24+; the level-<=1 list gets the packed
// single (level 1) or -1 (k itself).
template <bool BATCH>
__global__ __launch_bounds__(256) void inner_tables_kernel(PSView one,
                                                           const PSView* __restrict__ many,
                                                           const int64_t* __restrict__ inner_off,
                                                           int ninst, int64_t total) {
  const int64_t g = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if(g >= total) return;
  const int inst = BATCH ? inst_of(inner_off, ninst, g) : 0;
  const PSView& P = BATCH ? many[inst] : one;
  const int64_t k = BATCH ? g - inner_off[inst] : g;
  const IntsView& I = P.I;
  const uint64_t* __restrict__ istr = P.istr;
  const int32_t* __restrict__ ptr1 = P.ptr1;
  const int32_t* __restrict__ idx1 = P.idx1;
  const int32_t* __restrict__ ptr2 = P.ptr2;
  const int32_t* __restrict__ idx2 = P.idx2;
  int32_t* __restrict__ iexc1 = P.iexc1;
  double* __restrict__ ival2 = P.ival2;
  int32_t* __restrict__ iexc2 = P.iexc2;
  const int lane = threadIdx.x & 63;
  const uint64_t b = istr[k];
  for(int32_t t = ptr2[k] + lane; t < ptr2[k + 1]; t += 64) {
    const uint64_t kt = istr[idx2[t]];
    const int e = popc64(b ^ kt) >> 1;
    double val = 0.;
    int32_t code = 0;
    if(e == 2) {
      val = me_4(I, b, kt, b ^ kt);
    } else if(e == 1) {
      code = pack_single(b, kt);
      val = single_partial(I, b, code);
    }
    ival2[t] = val;
    iexc2[t] = code | (e << 24);
  }
  for(int32_t t = ptr1[k] + lane; t < ptr1[k + 1]; t += 64) {
    const uint64_t kt = istr[idx1[t]];
    iexc1[t] = kt == b ? -1 : pack_single(b, kt);
  }
}

__device__ __forceinline__ double code_sign(int32_t code) {
  return (code >> 16) & 1 ? -1. : 1.;
}

// Fill and elements in one pass, one wavefront per row. Lanes take
// consecutive output positions of the row (coalesced column and value
// stores). The row's segments (one per outer neighbour: 1 entry at level 2,
// the level-<=1 inner list at level 1, the level-<=2 list at level 0) are
// staged in LDS with their start (prefix sums of the lengths), packed
// neighbour, excitation code and a value: at level 2 the same-spin double
// element, at level 1 the element of the segment's entry with the row's own
// inner string (me_2 with that string as spectator), so the window loop
// reads it from LDS instead of summing integrals for one or two lanes of a
// window. A lane finds its segment from a 64-bit mask of the segment starts
// inside its 64-position window; the level-1 inner list of the row (at most
// 64 entries: registers, read across lanes) gives the inner neighbour and
// the inner single of an alpha-beta double. Elements are the Slater-Condon
// values of the tiled path's element pass (the same operations in the same
// order), so the CSR is bit-identical; |h| > thresh is counted per row for
// the shared compaction. LDS per wavefront: 18 bytes per outer neighbour
// (max_seg of the launch) + 260, so that at CAS(12e,12o) (262 outer
// neighbours) the 64-VGPR kernel keeps 8 wavefronts per SIMD.
__host__ __device__ constexpr size_t fill_lds_per_wave(int max_seg) {
  return ((size_t)18 * max_seg + 4 + 128 + 128 + 7) / 8 * 8;
}

// The alpha-beta double of an outer single x inner single is
// sign_out * sign_in * V(vi) with vi = P_in + n^2 P_out when the outer
// string is beta or the integrals are pair-symmetric (the pair-swapped
// entry), else P_out + n^2 P_in, where P = v + n o of the single (packed
// as o | v << 8 | (sign < 0) << 16). The fill stages the outer part as
// P_out | sign << 15 (16 bits, P < n^2 <= 4096) and holds the inner part
// of the row's level-1 list in registers as its term of vi | sign << 30
// (-1: the row's own inner string), so a lane forms vi with one
// multiply-add and the sign by one bit flip of the gathered integral
// (the same double as the product with +-1.0).
__device__ __forceinline__ uint16_t outer_pair(int32_t c, uint32_t n) {
  return (uint16_t)(((c & 0xff) * n + ((c >> 8) & 0xff)) | (((c >> 16) & 1) << 15));
}
__device__ __forceinline__ int32_t inner_pair(int32_t c, uint32_t n, uint32_t mult_in) {
  if(c < 0) return -1;
  return (int32_t)((((c & 0xff) * n + ((c >> 8) & 0xff)) * mult_in) | (((c >> 16) & 1) << 30));
}

template <bool BATCH>
__global__ __launch_bounds__(256) void product_fill_elements_kernel(
    PSView one, const PSView* __restrict__ many, const int32_t* __restrict__ row_inst,
    int64_t n, int max_seg, const double* __restrict__ diag, const int64_t* __restrict__ rowptr,
    int32_t* __restrict__ colind, double* __restrict__ nzval, double thresh,
    int64_t* __restrict__ kept, int* __restrict__ dropped) {
  extern __shared__ double fill_lds[];
  // the row, and what is derived from it, is uniform over the wavefront
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t row = (int64_t)blockIdx.x * kWavesPerBlock + wv;
  if(row >= n) return;  // whole wavefront leaves together
  const int lane = threadIdx.x & 63;
  // this wavefront's staging: values, then starts, packed neighbours, codes
  double* seg_val = fill_lds + wv * (fill_lds_per_wave(max_seg) / 8);
  int32_t* seg_start = reinterpret_cast<int32_t*>(seg_val + max_seg);
  int32_t* seg_code = seg_start + max_seg + 1;
  uint16_t* seg_exc = reinterpret_cast<uint16_t*>(seg_code + max_seg);
  int16_t* lv1 = reinterpret_cast<int16_t*>(seg_exc + max_seg);  // level-1 segments (<= 64)
  uint16_t* tag = reinterpret_cast<uint16_t*>(lv1 + 64);  // window start stamps (64)
  const PSView& P = BATCH ? many[row_inst[row]] : one;
  const IntsView& I = P.I;
  const int64_t L = P.L, row0 = P.row0;
  const int outer_alpha = P.outer_alpha;
  const uint64_t* __restrict__ ostr = P.ostr;
  const uint64_t* __restrict__ istr = P.istr;
  const int32_t* __restrict__ onb_ptr = P.onb_ptr;
  const int32_t* __restrict__ onb = P.onb;
  const int32_t* __restrict__ ptr1 = P.ptr1;
  const int32_t* __restrict__ idx1 = P.idx1;
  const int32_t* __restrict__ ptr2 = P.ptr2;
  const int32_t* __restrict__ idx2 = P.idx2;
  const double* __restrict__ oval = P.oval;
  const int32_t* __restrict__ oexc = P.oexc;
  const int32_t* __restrict__ iexc1 = P.iexc1;
  const double* __restrict__ ival2 = P.ival2;
  const int32_t* __restrict__ iexc2 = P.iexc2;
  const int64_t r = (row - row0) / L, k = (row - row0) % L;
  const uint64_t bo = ostr[r], bi = istr[k];
  const double d = diag[row];
  const int64_t st = rowptr[row];
  const int32_t s1 = ptr1[k], len1 = ptr1[k + 1] - s1;
  const int32_t s2 = ptr2[k], len2 = ptr2[k + 1] - s2;
  const int32_t j0 = onb_ptr[r], nseg = onb_ptr[r + 1] - j0;
  // alpha-beta integral index: vi = inner term + P_out * mult_out
  const uint32_t nn = (uint32_t)I.n;
  const bool swap_pairs = !outer_alpha || I.pair_sym;
  const uint32_t mult_out = swap_pairs ? nn * nn : 1u, mult_in = swap_pairs ? 1u : nn * nn;
  // the level-<=1 inner list in registers (lane t: entry t) when it fits
  const bool inreg = len1 <= 64;
  int32_t reg_k1 = 0, reg_c1 = 0;
  if(inreg && lane < len1) {
    reg_k1 = idx1[s1 + lane];
    reg_c1 = inner_pair(iexc1[s1 + lane], nn, mult_in);
  }
  // segment starts (exclusive prefix of lengths), neighbours, codes, values;
  // with at most 64 level-1 segments their own-inner-string elements are
  // formed afterwards in one round (a round per 64 segments would run the
  // integral sum for the few level-1 lanes of each)
  const int32_t n1 = P.ocnt[3 * r + 1];
  const bool compact1 = n1 <= 64;
  int32_t seen1 = 0;
  int32_t carry = 0;
  for(int32_t b = 0; b < nseg; b += 64) {
    const int32_t j = b + lane;
    const bool valid = j < nseg;
    const int32_t packed = valid ? onb[j0 + j] : 0;
    const int e = packed & 3;
    const int32_t len = !valid ? 0 : (e == 2 ? 1 : (e == 1 ? len1 : len2));
    int32_t incl = len;
#pragma unroll
    for(int off = 1; off < 64; off <<= 1) {
      const int32_t v = __shfl_up(incl, off, 64);
      if(lane >= off) incl += v;
    }
    if(valid) {
      seg_start[j] = carry + incl - len;
      seg_code[j] = packed;
      int32_t oc = 0;
      double val = 0.;
      if(e == 2) {  // outer double, same inner string: me_4 of the outer pair
        val = oval[j0 + j];
      } else if(e == 1 && !compact1) {  // outer single, same inner string: me_2
        oc = oexc[j0 + j];
        const uint32_t o = oc & 0xff, v = (oc >> 8) & 0xff;
        const double* V_ov = I.V_red + I.n * (v + I.n * o);
        double t = oval[j0 + j];
        for(uint64_t m = bi; m; m &= m - 1) t += V_ov[ctz64(m)];
        val = code_sign(oc) * t;
      }
      seg_exc[j] = outer_pair(oc, nn);
      seg_val[j] = val;
    }
    if(compact1) {
      const uint64_t m1 = ballot(valid && e == 1);
      if(valid && e == 1) lv1[seen1 + (int32_t)lanes_below(m1)] = (int16_t)j;
      seen1 += popc64(m1);
    }
    carry += __shfl(incl, 63, 64);
  }
  if(lane == 0) seg_start[nseg] = carry;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  if(compact1 && lane < n1) {  // outer single, same inner string: me_2, other spin = inner
    const int32_t j = lv1[lane];
    const int32_t oc = oexc[j0 + j];
    const uint32_t o = oc & 0xff, v = (oc >> 8) & 0xff;
    const double* V_ov = I.V_red + I.n * (v + I.n * o);
    double t = oval[j0 + j];
    for(uint64_t m = bi; m; m &= m - 1) t += V_ov[ctz64(m)];
    seg_exc[j] = outer_pair(oc, nn);
    seg_val[j] = code_sign(oc) * t;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  const int32_t total = carry;
  int64_t nk = 0;
  // Windows of 64 positions: the segments starting inside the window are
  // the next ones in order (jlo, jlo + 1, ...); their start offsets form a
  // 64-bit mask, and a lane's segment is jlo - 1 + the starts at or below it.
  int32_t jlo = 0;
  // lane's segment in the window at p0; advances jl past the window's starts
  uint32_t stamp = 0;
  tag[lane] = 0xffffu;
  auto locate = [&](int32_t p0, int32_t& jl) -> int32_t {
    const int32_t js = jl + lane;
    const int32_t sstart = js < nseg ? seg_start[js] : INT32_MAX;
    const bool inwin = sstart < p0 + 64;
    // each segment starting in the window stamps its start offset, each
    // lane reads its own offset's stamp: one LDS write and one read (an
    // OR-reduction of the mask through DPP ran 24 VALU operations)
    stamp = (stamp + 1u) & 0x7fffu;
    if(inwin) tag[sstart - p0] = (uint16_t)stamp;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const uint64_t startmask = ballot(tag[lane] == stamp);
    const int32_t lo = jl - 1 + (int32_t)popc64(startmask & lowmask((uint32_t)lane + 1u));
    jl += (int32_t)popc64(ballot(inwin));
    return lo;
  };
  // Column and element of position p (segment lo), without branches for
  // the level-2 and level-1 segments: `gather` issues the alpha-beta
  // integral load of every lane (index 0 where it is not needed), `finish`
  // forms the element, so that two windows' loads are in flight together.
  // Level-0 entries (the row's own outer string, a few windows per row)
  // are formed by `level0` afterwards. e: 3 past the row end.
  const double* __restrict__ Vg = I.V;
  const int32_t colrow = (int32_t)(row0 + r * L);  // first column of the row's own run
  struct Win {
    int32_t col, off;
    int e;
    bool ab;
    uint32_t neg;
    double sv, vv;
  };
  auto gather = [&](auto in_regs, int32_t p, int32_t lo, Win& w) {
    const bool act = p < total;
    const int32_t l = act ? lo : 0;
    const int32_t packed = seg_code[l];
    w.off = p - seg_start[l];
    w.e = act ? (packed & 3) : 3;
    w.sv = seg_val[l];
    const uint32_t xo = seg_exc[l];
    // the lane's entry of the level-<=1 inner list (meaningful at level 1)
    int32_t k1 = __shfl(reg_k1, w.off & 63, 64);
    int32_t c1 = __shfl(reg_c1, w.off & 63, 64);
    if constexpr(!decltype(in_regs)::value) {  // every lane reads an entry of the row's list
      const int32_t t1 = s1 + (w.off < len1 ? w.off : len1 - 1);
      k1 = idx1[t1];
      c1 = inner_pair(iexc1[t1], nn, mult_in);
    }
    // outer single x inner single: me_22 (alpha excitation first); outer
    // alpha with pair symmetry reads the swapped entry V(v2,o2,v1,o1), which
    // keeps a segment's lanes (fixed (o1, v1), varying (o2, v2)) inside one
    // n*n block
    w.ab = w.e == 1 && c1 >= 0;
    const uint32_t vi = (uint32_t)(c1 & 0xffffff) + (xo & 0xfffu) * mult_out;
    w.neg = ((uint32_t)c1 >> 30 ^ xo >> 15) & 1u;
    w.vv = Vg[w.ab ? vi : 0u];
    w.col = (int32_t)(row0 + (packed >> 2) * (int32_t)L) + (w.e == 1 ? k1 : (int32_t)k);
  };
  // level 2: me_4 of the outer pair; level 1 with the row's own inner
  // string: me_2 (both staged); level 0: the diagonal until level0();
  // alpha-beta: the integral with the product of the two signs
  auto finish = [&](const Win& w) -> double {
    const double sgn = __hiloint2double(__double2hiint(w.vv) ^ (int)(w.neg << 31),
                                        __double2loint(w.vv));
    return w.ab ? sgn : (w.e == 1 || w.e == 2) ? w.sv : d;
  };
  auto level0 = [&](int32_t off, int32_t& col, double& hv) {
    const int32_t t2 = s2 + off;
    col = colrow + idx2[t2];
    const int32_t ic = iexc2[t2];
    const int lvl = ic >> 24;
    if(lvl == 2) {  // inner double: me_4 of the inner pair
      hv = ival2[t2];
    } else if(lvl == 1) {  // inner single: me_2, other spin = outer
      const uint32_t o = ic & 0xff, v = (ic >> 8) & 0xff;
      const double* V_ov = I.V_red + I.n * (v + I.n * o);
      double t = ival2[t2];
      for(uint64_t m = bo; m; m &= m - 1) t += V_ov[ctz64(m)];
      hv = code_sign(ic) * t;
    }
  };
  // two windows per iteration: their element chains (cross-lane reads, the
  // integral gather) are independent and overlap
  auto windows = [&](auto in_regs) {
    for(int32_t p0 = 0; p0 < total; p0 += 128) {
      const int32_t loA = locate(p0, jlo);
      const int32_t loB = locate(p0 + 64, jlo);
      const int32_t pA = p0 + lane, pB = p0 + 64 + lane;
      Win wA, wB;
      gather(in_regs, pA, loA, wA);
      gather(in_regs, pB, loB, wB);
      // both integral loads are issued before either is waited for (the
      // empty statement consumes them here, so neither load sinks into the
      // branch that uses it)
      asm volatile("" : "+v"(wA.vv), "+v"(wB.vv));
      int32_t cA = wA.col, cB = wB.col;
      double hA = finish(wA), hB = finish(wB);
      if(ballot(wA.e == 0) | ballot(wB.e == 0)) {
        if(wA.e == 0) level0(wA.off, cA, hA);
        if(wB.e == 0) level0(wB.off, cB, hB);
      }
      // streamed out once: non-temporal, so the 19 GB of stores of a large
      // build do not evict the element tables and integrals from L2
      if(pA < total) {
        __builtin_nontemporal_store(cA, colind + st + pA);
        __builtin_nontemporal_store(hA, nzval + st + pA);
        nk += fabs(hA) > thresh ? 1 : 0;
      }
      if(pB < total) {
        __builtin_nontemporal_store(cB, colind + st + pB);
        __builtin_nontemporal_store(hB, nzval + st + pB);
        nk += fabs(hB) > thresh ? 1 : 0;
      }
    }
  };
  if(inreg)
    windows(std::integral_constant<bool, true>());
  else
    windows(std::integral_constant<bool, false>());
  nk = wave_sum(nk);
  if(lane == 0) {
    kept[row] = nk;
    if(nk != total) atomicOr(dropped, 1);
  }
}

// Neighbour counts of string i per excitation level 0/1/2 (wavefront per i).
__global__ __launch_bounds__(256) void nb_count_kernel(const uint64_t* __restrict__ s, int64_t m,
                                                       int32_t* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if(i >= m) return;
  const int lane = threadIdx.x & 63;
  const uint64_t si = s[i];
  int32_t c0 = 0, c1 = 0, c2 = 0;
  for(int64_t j0 = 0; j0 < m; j0 += 64) {
    const int64_t j = j0 + lane;
    const int e = j < m ? popc64(si ^ s[j]) : 99;
    c0 += popc64(ballot(e == 0));
    c1 += popc64(ballot(e == 2));
    c2 += popc64(ballot(e == 4));
  }
  if(lane == 0) {
    cnt[3 * i] = c0;
    cnt[3 * i + 1] = c1;
    cnt[3 * i + 2] = c2;
  }
}

// Ascending neighbour list of string i at levels <= maxlvl, written at
// ptr[i]; packed = j << 2 | level, or j.
__global__ __launch_bounds__(256) void nb_fill_kernel(const uint64_t* __restrict__ s, int64_t m,
                                                      int maxlvl, int packed,
                                                      const int32_t* __restrict__ ptr,
                                                      int32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if(i >= m) return;
  const int lane = threadIdx.x & 63;
  const uint64_t si = s[i];
  int64_t pos = ptr[i];
  for(int64_t j0 = 0; j0 < m; j0 += 64) {
    const int64_t j = j0 + lane;
    const int e = j < m ? popc64(si ^ s[j]) : 99;
    const bool hit = e <= 2 * maxlvl;
    const uint64_t mask = ballot(hit);
    if(hit) out[pos + lanes_below(mask)] = packed ? (int32_t)((j << 2) | (e >> 1)) : (int32_t)j;
    pos += popc64(mask);
  }
}

int popc(uint64_t x) { return __builtin_popcountll(x); }

bool distinct_equal_weight(std::vector<uint64_t> v) {
  for(uint64_t x : v)
    if(popc(x) != popc(v[0])) return false;
  std::sort(v.begin(), v.end());
  return std::adjacent_find(v.begin(), v.end()) == v.end();
}

template <int W>
bool detect(mhip_context* ctx, const uint64_t* d, int64_t n, ProductSpace& ps) {
  if(n < 2) return false;
  hipStream_t st = ctx->stream;
  // head of the list: outer spin and first run length
  std::vector<uint64_t> head;
  bool outer_alpha = true;
  int64_t L = 0, nh = 0;
  for(int64_t want = std::min<int64_t>(n, 4096);; want = std::min<int64_t>(n, kMaxHead)) {
    nh = want;
    head.resize(nh * W);
    d2h(ctx, head.data(), d, 8 * W * nh);
    auto alpha = [&](int64_t i) { return W == 1 ? head[i] & 0xffffffffull : head[2 * i]; };
    auto beta = [&](int64_t i) { return W == 1 ? head[i] >> 32 : head[2 * i + 1]; };
    if(alpha(0) == alpha(1)) outer_alpha = true;
    else if(beta(0) == beta(1)) outer_alpha = false;
    else return false;
    auto outer = [&](int64_t i) { return outer_alpha ? alpha(i) : beta(i); };
    L = 1;
    while(L < nh && outer(L) == outer(0)) ++L;
    if(L < nh || nh == n) break;           // run end found (or the list is one run)
    if(nh == std::min<int64_t>(n, kMaxHead)) return false;  // first run too long
  }
  if(n % L) return false;
  const int64_t nruns = n / L;
  if((double)nruns * nruns + (double)L * L > kMaxPairTests) return false;
  std::vector<uint64_t> I(L);
  for(int64_t k = 0; k < L; ++k)
    I[k] = outer_alpha ? (W == 1 ? head[k] >> 32 : head[2 * k + 1])
                       : (W == 1 ? head[k] & 0xffffffffull : head[2 * k]);
  if(!distinct_equal_weight(I)) return false;

  // arena 1: strings, per-level counts, row pointers, flag
  const size_t o_ostr = 0, o_istr = o_ostr + 8 * nruns, o_ocnt = o_istr + 8 * L,
               o_icnt = o_ocnt + 4 * 3 * nruns, o_optr = o_icnt + 4 * 3 * L,
               o_p1 = o_optr + 4 * (nruns + 1), o_p2 = o_p1 + 4 * (L + 1),
               o_flag = o_p2 + 4 * (L + 1), a1 = o_flag + 16;
  ps.arena.alloc(a1);
  char* A = ps.arena.as<char>();
  uint64_t* ostr = reinterpret_cast<uint64_t*>(A + o_ostr);
  uint64_t* istr = reinterpret_cast<uint64_t*>(A + o_istr);
  int32_t* ocnt = reinterpret_cast<int32_t*>(A + o_ocnt);
  int32_t* icnt = reinterpret_cast<int32_t*>(A + o_icnt);
  int* flag = reinterpret_cast<int*>(A + o_flag);
  MHIP_HIP_CHECK(hipMemcpyAsync(istr, I.data(), 8 * L, hipMemcpyHostToDevice, st));
  MHIP_HIP_CHECK(hipMemsetAsync(flag, 0, sizeof(int), st));
  hipLaunchKernelGGL(product_check_kernel<W>, dim3((n + 255) / 256), dim3(256), 0, st, d, n, L,
                     outer_alpha, flag);
  hipLaunchKernelGGL(run_heads_kernel<W>, dim3((nruns + 255) / 256), dim3(256), 0, st, d, nruns,
                     L, outer_alpha, ostr);
  hipLaunchKernelGGL(nb_count_kernel, dim3((nruns + kWavesPerBlock - 1) / kWavesPerBlock),
                     dim3(256), 0, st, ostr, nruns, ocnt);
  hipLaunchKernelGGL(nb_count_kernel, dim3((L + kWavesPerBlock - 1) / kWavesPerBlock), dim3(256),
                     0, st, istr, L, icnt);
  MHIP_HIP_CHECK(hipGetLastError());
  // ostr, istr, counts and the flag are contiguous: one copy back
  std::vector<char> h(o_flag + sizeof(int) - o_ostr);
  d2h(ctx, h.data(), A, h.size());
  int bad = 0;
  std::memcpy(&bad, h.data() + o_flag, sizeof(int));
  if(bad) return false;
  std::vector<uint64_t> O(nruns);
  std::memcpy(O.data(), h.data() + o_ostr, 8 * nruns);
  if(!distinct_equal_weight(O)) return false;
  const int32_t* hoc = reinterpret_cast<const int32_t*>(h.data() + o_ocnt);
  const int32_t* hic = reinterpret_cast<const int32_t*>(h.data() + o_icnt);
  std::vector<int32_t> ptrs(nruns + 1 + 2 * (L + 1));
  int32_t* optr = ptrs.data();
  int32_t* p1 = optr + nruns + 1;
  int32_t* p2 = p1 + L + 1;
  optr[0] = p1[0] = p2[0] = 0;
  int64_t max_seg = 0;
  for(int64_t r = 0; r < nruns; ++r) {
    const int64_t c = (int64_t)hoc[3 * r] + hoc[3 * r + 1] + hoc[3 * r + 2];
    MHIP_REQUIRE(optr[r] + c < (1ll << 31), MHIP_ERR_INDEX_OVERFLOW, "neighbour table too large");
    optr[r + 1] = optr[r] + (int32_t)c;
    max_seg = std::max(max_seg, c);
  }
  if(max_seg > kMaxSeg) return false;  // the fill stages a run's neighbours in LDS
  for(int64_t k = 0; k < L; ++k) {
    p1[k + 1] = p1[k] + hic[3 * k] + hic[3 * k + 1];
    p2[k + 1] = p2[k] + hic[3 * k] + hic[3 * k + 1] + hic[3 * k + 2];
  }
  MHIP_HIP_CHECK(hipMemcpyAsync(A + o_optr, ptrs.data(), 4 * ptrs.size(), hipMemcpyHostToDevice, st));
  // arena 2: the neighbour lists
  const size_t n_onb = optr[nruns], n1 = p1[L], n2 = p2[L];
  // lists (int32) then the element tables: oval, ival2 (double), oexc,
  // iexc1, iexc2 (int32)
  ps.lists.alloc(8 * (n_onb + n2) + 4 * (2 * n_onb + 2 * n1 + 2 * n2) + 16);
  double* oval = ps.lists.as<double>();
  double* ival2 = oval + n_onb;
  int32_t* onb = reinterpret_cast<int32_t*>(ival2 + n2);
  int32_t* idx1 = onb + n_onb;
  int32_t* idx2 = idx1 + n1;
  int32_t* oexc = idx2 + n2;
  int32_t* iexc1 = oexc + n_onb;
  int32_t* iexc2 = iexc1 + n1;
  hipLaunchKernelGGL(nb_fill_kernel, dim3((nruns + kWavesPerBlock - 1) / kWavesPerBlock),
                     dim3(256), 0, st, ostr, nruns, 2, 1,
                     reinterpret_cast<const int32_t*>(A + o_optr), onb);
  hipLaunchKernelGGL(nb_fill_kernel, dim3((L + kWavesPerBlock - 1) / kWavesPerBlock), dim3(256),
                     0, st, istr, L, 1, 0, reinterpret_cast<const int32_t*>(A + o_p1), idx1);
  hipLaunchKernelGGL(nb_fill_kernel, dim3((L + kWavesPerBlock - 1) / kWavesPerBlock), dim3(256),
                     0, st, istr, L, 2, 0, reinterpret_cast<const int32_t*>(A + o_p2), idx2);
  MHIP_HIP_CHECK(hipGetLastError());
  ps.L = L;
  ps.nruns = nruns;
  ps.max_seg = max_seg;
  ps.outer_alpha = outer_alpha;
  ps.ostr = ostr;
  ps.istr = istr;
  ps.ocnt = ocnt;
  ps.onb_ptr = reinterpret_cast<const int32_t*>(A + o_optr);
  ps.ptr1 = reinterpret_cast<const int32_t*>(A + o_p1);
  ps.ptr2 = reinterpret_cast<const int32_t*>(A + o_p2);
  ps.onb = onb;
  ps.idx1 = idx1;
  ps.idx2 = idx2;
  ps.oval = oval;
  ps.ival2 = ival2;
  ps.oexc = oexc;
  ps.iexc1 = iexc1;
  ps.iexc2 = iexc2;
  return true;
}

}  // namespace

bool detect_product_space(mhip_context* ctx, int W, const uint64_t* d, int64_t n,
                          ProductSpace& ps) {
  return W == 1 ? detect<1>(ctx, d, n, ps) : detect<2>(ctx, d, n, ps);
}

namespace {
PSView make_view(const ProductSpace& ps, const IntsView& I, int64_t row0) {
  PSView v;
  v.I = I;
  v.row0 = row0;
  v.L = ps.L;
  v.nruns = ps.nruns;
  v.outer_alpha = ps.outer_alpha ? 1 : 0;
  v.ostr = ps.ostr;
  v.istr = ps.istr;
  v.ocnt = ps.ocnt;
  v.onb_ptr = ps.onb_ptr;
  v.onb = ps.onb;
  v.ptr1 = ps.ptr1;
  v.idx1 = ps.idx1;
  v.ptr2 = ps.ptr2;
  v.idx2 = ps.idx2;
  v.oval = ps.oval;
  v.ival2 = ps.ival2;
  v.oexc = ps.oexc;
  v.iexc1 = ps.iexc1;
  v.iexc2 = ps.iexc2;
  return v;
}
}  // namespace

void product_count(mhip_context* ctx, const ProductSpace& ps, int64_t n, uint32_t* counts) {
  hipLaunchKernelGGL(product_count_kernel<false>, dim3((n + 255) / 256), dim3(256), 0,
                     ctx->stream, make_view(ps, IntsView{}, 0), nullptr, nullptr, n, counts);
  MHIP_HIP_CHECK(hipGetLastError());
}

void product_fill_elements(mhip_context* ctx, int W, const ProductSpace& ps, const IntsView& I,
                           int64_t n, const double* diag, const int64_t* rowptr, int32_t* colind,
                           double* nzval, double thresh, int64_t* kept, int* dropped) {
  (void)W;  // the strings are stored split; the kernel does not read determinants
  MHIP_REQUIRE(ps.max_seg <= kMaxSeg, MHIP_ERR_INVALID_ARG, "product space outside LDS staging");
  const PSView v = make_view(ps, I, 0);
  // element tables of the neighbour lists (they depend on the integrals)
  hipLaunchKernelGGL(outer_tables_kernel<false>,
                     dim3((ps.nruns + kWavesPerBlock - 1) / kWavesPerBlock), dim3(256), 0,
                     ctx->stream, v, nullptr, nullptr, 1, ps.nruns);
  hipLaunchKernelGGL(inner_tables_kernel<false>, dim3((ps.L + kWavesPerBlock - 1) / kWavesPerBlock),
                     dim3(256), 0, ctx->stream, v, nullptr, nullptr, 1, ps.L);
  const int ms = (int)std::max<int64_t>(1, ps.max_seg);
  hipLaunchKernelGGL((product_fill_elements_kernel<false>),
                     dim3((n + kWavesPerBlock - 1) / kWavesPerBlock), dim3(256),
                     kWavesPerBlock * fill_lds_per_wave(ms), ctx->stream, v, nullptr, nullptr, n,
                     ms, diag, rowptr, colind, nzval, thresh, kept, dropped);
  MHIP_HIP_CHECK(hipGetLastError());
}

// ---- batched instances -----------------------------------------------------
// Instances with the same determinant list (list_id) share one structure;
// every instance has its own element tables (its integrals) and view.
bool prepare_batch_product(mhip_context* ctx, int W, const uint64_t* d_dets,
                           const std::vector<int64_t>& off, const std::vector<int>& list_id,
                           const IntsView* d_views, ProductBatch& pb) {
  const int ninst = (int)list_id.size();
  pb.spaces.clear();
  pb.space_of.assign(ninst, -1);
  std::vector<int> first_of;  // structure index -> first instance with it
  for(int i = 0; i < ninst; ++i) {
    const int first = list_id[i];
    if(first != i) {  // same list as an earlier instance
      MHIP_REQUIRE(first >= 0 && first < i && pb.space_of[first] >= 0, MHIP_ERR_INVALID_ARG,
                   "bad list id");
      pb.space_of[i] = pb.space_of[first];
      continue;
    }
    std::unique_ptr<ProductSpace> ps(new ProductSpace());
    if(!detect_product_space(ctx, W, d_dets + W * off[i], off[i + 1] - off[i], *ps)) {
      pb.spaces.clear();
      return false;
    }
    pb.space_of[i] = (int)pb.spaces.size();
    pb.spaces.push_back(std::move(ps));
  }
  // per-instance element tables: oval, ival2 (double), oexc, iexc1, iexc2
  std::vector<size_t> tab_off(ninst + 1, 0);
  std::vector<int64_t> run_off(ninst + 1, 0), inner_off(ninst + 1, 0);
  for(int i = 0; i < ninst; ++i) {
    const ProductSpace& ps = *pb.spaces[pb.space_of[i]];
    run_off[i + 1] = run_off[i] + ps.nruns;
    inner_off[i + 1] = inner_off[i] + ps.L;
  }
  // table sizes need the list lengths (device row pointers): read them once
  // per structure
  std::vector<std::array<int64_t, 3>> sizes(pb.spaces.size());
  for(size_t q = 0; q < pb.spaces.size(); ++q) {
    const ProductSpace& ps = *pb.spaces[q];
    int32_t a = 0, b = 0, c = 0;
    d2h(ctx, &a, ps.onb_ptr + ps.nruns, 4);
    d2h(ctx, &b, ps.ptr1 + ps.L, 4);
    d2h(ctx, &c, ps.ptr2 + ps.L, 4);
    sizes[q] = {a, b, c};
  }
  for(int i = 0; i < ninst; ++i) {
    const auto& z = sizes[pb.space_of[i]];
    tab_off[i + 1] = tab_off[i] + 8 * (size_t)(z[0] + z[2]) + 4 * (size_t)(z[0] + z[1] + z[2]);
    tab_off[i + 1] = (tab_off[i + 1] + 15) / 16 * 16;
  }
  pb.tables.alloc(std::max<size_t>(tab_off[ninst], 16));
  std::vector<IntsView> hv(ninst);
  d2h(ctx, hv.data(), d_views, sizeof(IntsView) * ninst);
  std::vector<PSView> views(ninst);
  int64_t max_seg = 0;
  for(int i = 0; i < ninst; ++i) {
    const ProductSpace& ps = *pb.spaces[pb.space_of[i]];
    max_seg = std::max(max_seg, ps.max_seg);
    PSView v = make_view(ps, hv[i], off[i]);
    const auto& z = sizes[pb.space_of[i]];
    char* base = pb.tables.as<char>() + tab_off[i];
    v.oval = reinterpret_cast<double*>(base);
    v.ival2 = v.oval + z[0];
    v.oexc = reinterpret_cast<int32_t*>(v.ival2 + z[2]);
    v.iexc1 = v.oexc + z[0];
    v.iexc2 = v.iexc1 + z[1];
    views[i] = v;
  }
  MHIP_REQUIRE(max_seg <= kMaxSeg, MHIP_ERR_INVALID_ARG, "product space outside LDS staging");
  pb.views.alloc(sizeof(PSView) * ninst);
  pb.run_off.alloc(8 * (ninst + 1));
  pb.inner_off.alloc(8 * (ninst + 1));
  hipStream_t st = ctx->stream;
  MHIP_HIP_CHECK(hipMemcpyAsync(pb.views.p, views.data(), sizeof(PSView) * ninst,
                                hipMemcpyHostToDevice, st));
  MHIP_HIP_CHECK(hipMemcpyAsync(pb.run_off.p, run_off.data(), 8 * (ninst + 1),
                                hipMemcpyHostToDevice, st));
  MHIP_HIP_CHECK(hipMemcpyAsync(pb.inner_off.p, inner_off.data(), 8 * (ninst + 1),
                                hipMemcpyHostToDevice, st));
  MHIP_HIP_CHECK(hipStreamSynchronize(st));
  pb.ninst = ninst;
  pb.max_seg = max_seg;
  pb.total_runs = run_off[ninst];
  pb.total_inner = inner_off[ninst];
  return true;
}

void batch_product_count(mhip_context* ctx, const ProductBatch& pb, const int32_t* row_inst,
                         int64_t n, uint32_t* counts) {
  hipLaunchKernelGGL(product_count_kernel<true>, dim3((n + 255) / 256), dim3(256), 0,
                     ctx->stream, PSView{}, pb.views.as<PSView>(), row_inst, n, counts);
  MHIP_HIP_CHECK(hipGetLastError());
}

void batch_product_fill_elements(mhip_context* ctx, const ProductBatch& pb,
                                 const int32_t* row_inst, int64_t n, const double* diag,
                                 const int64_t* rowptr, int32_t* colind, double* nzval,
                                 double thresh, int64_t* kept, int* dropped) {
  hipStream_t st = ctx->stream;
  const PSView* v = pb.views.as<PSView>();
  hipLaunchKernelGGL(outer_tables_kernel<true>,
                     dim3((pb.total_runs + kWavesPerBlock - 1) / kWavesPerBlock), dim3(256), 0, st,
                     PSView{}, v, pb.run_off.as<int64_t>(), pb.ninst, pb.total_runs);
  hipLaunchKernelGGL(inner_tables_kernel<true>,
                     dim3((pb.total_inner + kWavesPerBlock - 1) / kWavesPerBlock), dim3(256), 0,
                     st, PSView{}, v, pb.inner_off.as<int64_t>(), pb.ninst, pb.total_inner);
  const int ms = (int)std::max<int64_t>(1, pb.max_seg);
  hipLaunchKernelGGL((product_fill_elements_kernel<true>),
                     dim3((n + kWavesPerBlock - 1) / kWavesPerBlock), dim3(256),
                     kWavesPerBlock * fill_lds_per_wave(ms), st, PSView{}, v, row_inst, n, ms, diag,
                     rowptr, colind, nzval, thresh, kept, dropped);
  MHIP_HIP_CHECK(hipGetLastError());
}

}  // namespace mhip
