// Device-side determinant primitives and Slater-Condon matrix elements.
//
// Determinants are W = 1 or 2 uint64 words (see macis_hip.h). All integer
// excitation / sign logic is done in registers; the floating-point sums
// follow the reference's source order term by term and the library is
// compiled with -ffp-contract=off, so elements are bit-identical to the
// reference's serial evaluation.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mhip {

struct IntsView {
  int n;
  const double* __restrict__ T;
  const double* __restrict__ V;
  const double* __restrict__ G;
  const double* __restrict__ G_red;
  const double* __restrict__ V_red;
  const double* __restrict__ G2_red;
  const double* __restrict__ V2_red;
  int pair_sym;  // V(pq|rs) == V(rs|pq) bit for bit
};

__device__ __forceinline__ int popc64(uint64_t x) { return __popcll(x); }
__device__ __forceinline__ uint32_t ctz64(uint64_t x) {
  return (uint32_t)__builtin_ctzll(x);
}
// full_mask<N>(i): bits [0, i)   (bitset_operations.hpp:225-230)
__device__ __forceinline__ uint64_t lowmask(uint32_t i) {
  return i >= 64 ? ~0ull : ((1ull << i) - 1ull);
}

// single_excitation_sign (sd_operations.hpp:503-513): parity of the occupied
// orbitals of `state` strictly between p and q.
__device__ __forceinline__ double excitation_sign(uint64_t state, uint32_t p,
                                                  uint32_t q) {
  const uint32_t lo = (p < q ? p : q) + 1u, hi = p < q ? q : p;
  const uint64_t m = state & (lowmask(hi) ^ lowmask(lo));
  return (popc64(m) & 1) ? -1. : 1.;
}

// Spin strings of a determinant.
template <int W>
struct Spins {
  uint64_t a, b;
};

template <int W>
__device__ __forceinline__ Spins<W> load_spins(const uint64_t* __restrict__ p,
                                               int64_t i) {
  Spins<W> s;
  if constexpr(W == 1) {
    const uint64_t w = p[i];
    s.a = w & 0xffffffffull;
    s.b = w >> 32;
  } else {
    const ulonglong2 v = reinterpret_cast<const ulonglong2*>(p)[i];
    s.a = v.x;
    s.b = v.y;
  }
  return s;
}

template <int W>
__device__ __forceinline__ void store_spins(uint64_t* __restrict__ p, int64_t i,
                                            uint64_t a, uint64_t b) {
  if constexpr(W == 1) {
    p[i] = a | (b << 32);
  } else {
    ulonglong2 v;
    v.x = a;
    v.y = b;
    reinterpret_cast<ulonglong2*>(p)[i] = v;
  }
}

// matrix_element_4 (matrix_elements.hpp:65-71) with doubles_sign_indices
// (sd_operations.hpp:528-543).
__device__ __forceinline__ double me_4(const IntsView& I, uint64_t bra,
                                       uint64_t ket, uint64_t ex) {
  const int n = I.n;
  const uint32_t o1 = ctz64(ket & ex), v1 = ctz64(bra & ex);
  double sign = excitation_sign(ket, v1, o1);
  const uint64_t f = (1ull << o1) | (1ull << v1);
  ket ^= f;
  ex ^= f;
  const uint32_t o2 = ctz64(ket & ex), v2 = ctz64(bra & ex);
  sign *= excitation_sign(ket, v2, o2);
  return sign * I.G[v1 + n * (o1 + n * (v2 + n * o2))];
}

// matrix_element_22 (matrix_elements.hpp:73-84)
__device__ __forceinline__ double me_22(const IntsView& I, uint64_t bra_a,
                                        uint64_t ket_a, uint64_t ex_a,
                                        uint64_t bra_b, uint64_t ket_b,
                                        uint64_t ex_b) {
  const int n = I.n;
  const uint32_t o1 = ctz64(ket_a & ex_a), v1 = ctz64(bra_a & ex_a);
  const double sa = excitation_sign(ket_a, v1, o1);
  const uint32_t o2 = ctz64(ket_b & ex_b), v2 = ctz64(bra_b & ex_b);
  const double sb = excitation_sign(ket_b, v2, o2);
  const double sign = sa * sb;
  return sign * I.V[v1 + n * (o1 + n * (v2 + n * o2))];
}

// h + row[k1] + row[k2] + ... over the set bits k1 < k2 < ... of m, added
// one at a time in that order; the loads are issued eight at a time (a
// lane's terms come from one integral row, so a plain loop waits one load
// latency per term).
__device__ __forceinline__ double add_row_terms(double h, const double* __restrict__ row,
                                                uint64_t m) {
  while(m) {
    double t[8];
    uint64_t mm = m;
#pragma unroll
    for(int u = 0; u < 8; ++u) {
      t[u] = row[mm ? ctz64(mm) : 0u];
      mm &= mm - 1;
    }
#pragma unroll
    for(int u = 0; u < 8; ++u) {
      if(m) h += t[u];
      m &= m - 1;
    }
  }
  return h;
}

// matrix_element_2 (matrix_elements.hpp:86-106): the sums run over the
// bra's occupied orbitals in ascending order, same spin then other spin.
__device__ __forceinline__ double me_2(const IntsView& I, uint64_t bra,
                                       uint64_t ket, uint64_t ex,
                                       uint64_t bra_same, uint64_t bra_othr) {
  const int n = I.n;
  const uint32_t o1 = ctz64(ket & ex), v1 = ctz64(bra & ex);
  const double sign = excitation_sign(ket, v1, o1);
  double h = I.T[v1 + n * o1];
  h = add_row_terms(h, I.G_red + n * (v1 + n * o1), bra_same);
  h = add_row_terms(h, I.V_red + n * (v1 + n * o1), bra_othr);
  return sign * h;
}

// matrix_element_diag (matrix_elements.hpp:108-135), exact term order.
__device__ __forceinline__ double me_diag(const IntsView& I, uint64_t oa,
                                          uint64_t ob) {
  const int n = I.n;
  double h = 0.;
  for(uint64_t m = oa; m; m &= m - 1) h += I.T[ctz64(m) * (n + 1)];
  for(uint64_t m = ob; m; m &= m - 1) h += I.T[ctz64(m) * (n + 1)];
  for(uint64_t mq = oa; mq; mq &= mq - 1) {
    const double* col = I.G2_red + n * ctz64(mq);
    for(uint64_t mp = oa; mp; mp &= mp - 1) h += col[ctz64(mp)];
  }
  for(uint64_t mq = ob; mq; mq &= mq - 1) {
    const double* col = I.G2_red + n * ctz64(mq);
    for(uint64_t mp = ob; mp; mp &= mp - 1) h += col[ctz64(mp)];
  }
  for(uint64_t mq = ob; mq; mq &= mq - 1) {
    const double* col = I.V2_red + n * ctz64(mq);
    for(uint64_t mp = oa; mp; mp &= mp - 1) h += col[ctz64(mp)];
  }
  return h;
}

// matrix_element dispatch (matrix_elements.hpp:32-63). `diag` is the
// precomputed me_diag of the bra.
__device__ __forceinline__ double matrix_element(const IntsView& I,
                                                 uint64_t ba, uint64_t bb,
                                                 uint64_t ka, uint64_t kb,
                                                 double diag) {
  const uint64_t exa = ba ^ ka, exb = bb ^ kb;
  const int ca = popc64(exa), cb = popc64(exb);
  if(ca + cb > 4) return 0.;
  if(ca == 4) return me_4(I, ba, ka, exa);
  if(cb == 4) return me_4(I, bb, kb, exb);
  if(ca == 2 && cb == 2) return me_22(I, ba, ka, exa, bb, kb, exb);
  if(ca == 2) return me_2(I, ba, ka, exa, ba, bb);
  if(cb == 2) return me_2(I, bb, kb, exb, bb, ba);
  return diag;
}

// Wave-level helpers (wave64).
__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
// Number of set bits of `mask` in lanes below this one.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for(int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// OR of a 64-bit value over the wavefront (every lane active), uniform
// result: DPP row shifts by 1, 2, 4, 8 leave each 16-lane row's OR in its
// lane 15, two row broadcasts carry rows 0-2 into lane 63, which is read.
// Both halves go through VALU DPP operations; no LDS instruction.
__device__ __forceinline__ uint64_t wave_or64(uint64_t v) {
  uint32_t a = (uint32_t)v, b = (uint32_t)(v >> 32);
#define MHIP_OR_DPP(ctrl, rows, bound)                                                  \
  a |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)a, ctrl, rows, 0xf, bound);       \
  b |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)b, ctrl, rows, 0xf, bound);
  MHIP_OR_DPP(0x111, 0xf, true)   // row_shr:1
  MHIP_OR_DPP(0x112, 0xf, true)   // row_shr:2
  MHIP_OR_DPP(0x114, 0xf, true)   // row_shr:4
  MHIP_OR_DPP(0x118, 0xf, true)   // row_shr:8
  MHIP_OR_DPP(0x142, 0xa, false)  // row_bcast:15 into rows 1 and 3
  MHIP_OR_DPP(0x143, 0xc, false)  // row_bcast:31 into rows 2 and 3
#undef MHIP_OR_DPP
  // readlane returns a signed int: widen through uint32_t (no sign extension)
  return (uint64_t)(uint32_t)__builtin_amdgcn_readlane(a, 63) |
         ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(b, 63) << 32);
}

__device__ __forceinline__ int64_t uniform(int64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

}  // namespace mhip
