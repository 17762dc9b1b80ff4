// K1, spin-string group structure pass (selected-CI lists).
//
// Produces the CSR pattern of DoubleLoopHamiltonianGenerator's loop
// (double_loop.hpp:45-91: for every bra row, every ket j with
// popcount(bra ^ ket) <= 4, columns ascending) for a square list whose
// determinants all carry the same alpha and the same beta electron count.
// Then popcount(a ^ a') and popcount(b ^ b') are even, and a connected ket
// falls in exactly one of three classes:
//   B: the row's beta string, alpha distance <= 4 (the diagonal included);
//   A: the row's alpha string, beta distance 2 or 4;
//   C: alpha distance 2 and beta distance 2.
// Determinants are grouped by alpha string and by beta string (stable radix
// sorts, so a group lists its members in ascending index). For each group
// the groups of the same spin one single excitation away are found (one
// lookup per hole/particle pair of the group string in a hash table of the
// distinct strings), and a flat list is written per group: its own members,
// then the members of those neighbour groups, each entry holding the
// member's other-spin string and its index. A row scans two contiguous
// lists, 64 entries per step:
//   its beta group's list — own part: class B (alpha distance <= 4);
//     neighbour part, when class C goes through beta: alpha distance == 2;
//   its alpha group's list — own part: class A (beta distance 2 or 4);
//     neighbour part, when class C goes through alpha: beta distance == 2;
// class C takes the side whose neighbour part is shorter. Hits are counted
// (count pass) or written in scan order (fill pass); a segmented radix sort
// then orders every row's columns. The CSR equals the tiled loop's: the
// classes partition {j : popcount(bra ^ ket_j) <= 4}.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <vector>

#include "mhip_device.hpp"
#include "mhip_internal.hpp"

namespace mhip {
namespace {

constexpr int kWaves = 4;  // wavefronts per 256-thread block

template <int W>
__device__ __forceinline__ uint64_t alpha_word(const uint64_t* __restrict__ d, int64_t i) {
  return W == 1 ? (d[i] & 0xffffffffull) : d[2 * i];
}
template <int W>
__device__ __forceinline__ uint64_t beta_word(const uint64_t* __restrict__ d, int64_t i) {
  return W == 1 ? (d[i] >> 32) : d[2 * i + 1];
}

// ctl: [0] OR of the alpha strings, [1] OR of the beta strings;
// cnt: min / max alpha count, min / max beta count.
__global__ void init_ctl_kernel(unsigned long long* ctl, int* cnt) {
  if(threadIdx.x == 0) {
    ctl[0] = ctl[1] = 0ull;
    cnt[0] = cnt[2] = INT_MAX;
    cnt[1] = cnt[3] = -1;
  }
}

// Spin words and identity index of every determinant; string unions and
// electron-count ranges (one atomic per wavefront and quantity).
template <int W>
__global__ __launch_bounds__(256) void split_kernel(const uint64_t* __restrict__ d, int64_t n,
                                                    uint64_t* __restrict__ ka,
                                                    uint64_t* __restrict__ kb,
                                                    int32_t* __restrict__ idx,
                                                    unsigned long long* __restrict__ ctl,
                                                    int* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  uint64_t a = 0, b = 0;
  int amin = INT_MAX, amax = -1, bmin = INT_MAX, bmax = -1;
  if(i < n) {
    a = alpha_word<W>(d, i);
    b = beta_word<W>(d, i);
    ka[i] = a;
    kb[i] = b;
    idx[i] = (int32_t)i;
    amin = amax = popc64(a);
    bmin = bmax = popc64(b);
  }
  // every lane takes part (inactive ones with neutral values)
  const uint64_t oa = wave_or64(a), ob = wave_or64(b);
#pragma unroll
  for(int off = 32; off > 0; off >>= 1) {
    amin = min(amin, __shfl_xor(amin, off, 64));
    amax = max(amax, __shfl_xor(amax, off, 64));
    bmin = min(bmin, __shfl_xor(bmin, off, 64));
    bmax = max(bmax, __shfl_xor(bmax, off, 64));
  }
  if((threadIdx.x & 63) == 0) {
    atomicOr(ctl, (unsigned long long)oa);
    atomicOr(ctl + 1, (unsigned long long)ob);
    atomicMin(cnt, amin);
    atomicMax(cnt + 1, amax);
    atomicMin(cnt + 2, bmin);
    atomicMax(cnt + 3, bmax);
  }
}

// head[p] = 1 where sorted string p starts a group
__global__ void head_kernel(const uint64_t* __restrict__ sk, int64_t n, int32_t* __restrict__ head) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if(p < n) head[p] = (p == 0 || sk[p] != sk[p - 1]) ? 1 : 0;
}

// gid = inclusive count of heads (group g + 1): group starts and strings,
// each determinant's group, and each sorted member's other-spin string.
__global__ void group_kernel(const uint64_t* __restrict__ sk, const int32_t* __restrict__ sm,
                             const int32_t* __restrict__ gid, int64_t n,
                             const uint64_t* __restrict__ other, int64_t* __restrict__ start,
                             uint64_t* __restrict__ ustr, int32_t* __restrict__ gof,
                             uint64_t* __restrict__ mword) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if(p >= n) return;
  const int32_t g = gid[p] - 1;
  if(p == 0 || gid[p - 1] != gid[p]) {
    start[g] = p;
    ustr[g] = sk[p];
  }
  if(p == n - 1) start[g + 1] = n;
  const int32_t m = sm[p];
  gof[m] = g;
  mword[p] = other[m];
}

// Open-addressing table of the distinct strings of one spin (power-of-two
// size, at most ~40 % full): slot value g + 1 for group g, 0 when empty.
__device__ __forceinline__ uint64_t str_hash(uint64_t x) {
  // 64-bit finaliser (xor-shift / multiply rounds)
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

// thread per group: claims the first free slot from the string's hash (the
// strings are distinct, so no slot holds an equal key)
__global__ void str_insert_kernel(const uint64_t* __restrict__ ustr, int64_t ng, uint64_t mask,
                                  uint64_t* __restrict__ hkey, int32_t* __restrict__ hval) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if(g >= ng) return;
  const uint64_t x = ustr[g];
  for(uint64_t h = str_hash(x) & mask;; h = (h + 1) & mask) {
    if(atomicCAS(hval + h, 0, (int32_t)(g + 1)) == 0) {
      hkey[h] = x;
      return;
    }
  }
}

__device__ __forceinline__ int64_t str_find(const uint64_t* __restrict__ hkey,
                                            const int32_t* __restrict__ hval, uint64_t mask,
                                            uint64_t x) {
  for(uint64_t h = str_hash(x) & mask;; h = (h + 1) & mask) {
    const int32_t v = hval[h];
    if(v == 0) return -1;
    if(hkey[h] == x) return v - 1;
  }
}

// Single-excitation neighbours of every group string among the group
// strings (one wavefront per group): each (hole, particle) pair, particles
// drawn from the union of the strings, is looked up in the strings' hash
// table. Count pass: neighbours and their total membership; fill pass: the
// neighbour list.
template <bool FILL>
__global__ __launch_bounds__(256) void neighbour_kernel(
    const uint64_t* __restrict__ ustr, const int64_t* __restrict__ start, int64_t ng,
    uint64_t uni, const uint64_t* __restrict__ hkey, const int32_t* __restrict__ hval,
    uint64_t hmask, int64_t* __restrict__ ncount, int64_t* __restrict__ nsize,
    const int64_t* __restrict__ noff, int32_t* __restrict__ nbr) {
  __shared__ uint8_t pos[kWaves][2][64];
  const int wv = threadIdx.x >> 6;
  const int64_t g = (int64_t)blockIdx.x * kWaves + wv;
  if(g >= ng) return;  // whole wavefront leaves together
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t s = ustr[g], parts = uni & ~s;
  const uint64_t bit = 1ull << lane;
  if(s & bit) pos[wv][0][popc64(s & lowmask(lane))] = (uint8_t)lane;
  if(parts & bit) pos[wv][1][popc64(parts & lowmask(lane))] = (uint8_t)lane;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  const int32_t nh = popc64(s), np = popc64(parts), tot = nh * np;
  int64_t cnt = 0, msum = 0;
  const int64_t out = FILL ? noff[g] : 0;
  for(int32_t t0 = 0; t0 < tot; t0 += 64) {
    const int32_t t = t0 + (int32_t)lane;
    int64_t found = -1;
    if(t < tot) {
      const uint32_t h = pos[wv][0][t / np], q = pos[wv][1][t % np];
      found = str_find(hkey, hval, hmask, s ^ (1ull << h) ^ (1ull << q));
    }
    const uint64_t m = ballot(found >= 0);
    if(found >= 0) {
      if(FILL) nbr[out + cnt + lanes_below(m)] = (int32_t)found;
      msum += start[found + 1] - start[found];
    }
    cnt += popc64(m);
  }
  if(!FILL) {
    msum = wave_sum(msum);
    if(lane == 0) {
      ncount[g] = cnt;
      nsize[g] = msum;
    }
  }
}

// flat length of group g: own members + neighbour members
__global__ void flat_len_kernel(const int64_t* __restrict__ start,
                                const int64_t* __restrict__ nsize, int64_t ng,
                                int64_t* __restrict__ len) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if(g < ng) len[g] = start[g + 1] - start[g] + nsize[g];
  if(g == ng) len[g] = 0;
}

// Flat list of group g (one wavefront per group): own members, then the
// members of each neighbour group in list order.
__global__ __launch_bounds__(256) void flatten_kernel(
    const int64_t* __restrict__ start, const uint64_t* __restrict__ mword,
    const int32_t* __restrict__ sm, const int64_t* __restrict__ noff,
    const int32_t* __restrict__ nbr, const int64_t* __restrict__ foff, int64_t ng,
    uint64_t* __restrict__ fw, int32_t* __restrict__ fi) {
  const int64_t g = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
  if(g >= ng) return;
  const uint32_t lane = threadIdx.x & 63;
  int64_t out = foff[g];
  // four 64-entry steps per iteration with their loads issued together: a
  // large group (the Hartree-Fock string's) is copied into every neighbour's
  // list, and one load latency per step bounded those copies
  auto copy = [&](int64_t gg) {
    const int64_t p0 = start[gg], p1 = start[gg + 1];
    int64_t p = p0 + lane;
    for(; p + 192 < p1; p += 256) {
      uint64_t w[4];
      int32_t m[4];
#pragma unroll
      for(int u = 0; u < 4; ++u) {
        w[u] = mword[p + 64 * u];
        m[u] = sm[p + 64 * u];
      }
#pragma unroll
      for(int u = 0; u < 4; ++u) {
        fw[out + (p + 64 * u - p0)] = w[u];
        fi[out + (p + 64 * u - p0)] = m[u];
      }
    }
    for(; p < p1; p += 64) {
      fw[out + (p - p0)] = mword[p];
      fi[out + (p - p0)] = sm[p];
    }
    out += p1 - p0;
  };
  copy(g);
  const int64_t j1 = noff[g + 1];
  for(int64_t j = noff[g]; j < j1; ++j) copy((int64_t)__builtin_amdgcn_readfirstlane(nbr[j]));
}

struct SideView {
  const int64_t* start;  // group g's own members: start[g + 1] - start[g]
  const int64_t* nsize;  // members of its neighbour groups
  const int64_t* foff;   // its flat list
  const int32_t* gof;    // group of each determinant
  const uint64_t* fw;    // other-spin string per flat entry
  const int32_t* fi;     // determinant index per flat entry
  const int32_t* perm;   // determinants in group order
};

constexpr int kXcds = 8;
constexpr int kChunkBlocks = 16;  // consecutive blocks (64 rows) kept on one XCD

// One side's list scan for every row (one wavefront per row), rows taken in
// that side's group order so rows sharing a flat list run together; blocks
// come in chunks of 16 dealt round-robin to the 8 XCDs (the hardware deals
// single blocks), so a chunk's rows meet in one L2. SIDE 1 (beta): the beta
// group's list, count into cnt_b / columns from rowptr[row]; SIDE 0 (alpha):
// the alpha group's list, total count cnt_b + own into counts / columns from
// rowptr[row] + cnt_b[row]. The row's class C goes with the side whose
// neighbour part is shorter (alpha on ties), the same choice in both passes.
template <int W, bool FILL, int SIDE>
__global__ __launch_bounds__(256) void group_scan_kernel(const uint64_t* __restrict__ d,
                                                         int64_t n, int64_t nblocks, SideView A,
                                                         SideView B,
                                                         uint32_t* __restrict__ cnt_b,
                                                         uint32_t* __restrict__ counts,
                                                         const int64_t* __restrict__ rowptr,
                                                         int32_t* __restrict__ col) {
  const int64_t hb = blockIdx.x, x = hb % kXcds, j = hb / kXcds;
  const int64_t lb = ((j / kChunkBlocks) * kXcds + x) * kChunkBlocks + j % kChunkBlocks;
  const int64_t k = lb * kWaves + (threadIdx.x >> 6);
  if(lb >= nblocks || k >= n) return;  // whole wavefront leaves together
  const uint32_t lane = threadIdx.x & 63;
  const SideView& S = SIDE ? B : A;
  const int64_t row = __builtin_amdgcn_readfirstlane(S.perm[k]);
  const uint64_t a = alpha_word<W>(d, row), b = beta_word<W>(d, row);
  const int32_t ga = A.gof[row], gb = B.gof[row];
  const int64_t anb = A.nsize[ga], bnb = B.nsize[gb];
  const bool via_alpha = anb <= bnb;
  const int32_t g = SIDE ? gb : ga;
  const int64_t own = S.start[g + 1] - S.start[g], o = S.foff[g];
  const int64_t len = own + (SIDE ? (via_alpha ? 0 : bnb) : (via_alpha ? anb : 0));
  // own part: beta side e <= 4 (class B), alpha side e in {2, 4} (class A);
  // neighbour part: e == 2 (class C)
  const uint64_t xw = SIDE ? a : b;
  const int emin = SIDE ? 0 : 2;
  const uint64_t* __restrict__ fw = S.fw;
  const int32_t* __restrict__ fi = S.fi;
  const int64_t base = FILL ? rowptr[row] + (SIDE ? 0 : (int64_t)cnt_b[row]) : 0;
  int64_t cnt = 0;
  // Groups of eight 64-entry steps, the group's eight loads issued before
  // any is tested (the scan is bound by load latency); in the fill pass the
  // index loads are issued together too, by every lane (lanes without a hit
  // read the list's first entry), so no load sits in a branch and the
  // stores wait once.
  constexpr int U = 8;
  const int64_t ngrp = len / (64 * U);
  for(int64_t gi = 0; gi < ngrp; ++gi) {
    const int64_t q0 = gi * 64 * U;
    uint64_t w[U];
#pragma unroll
    for(int u = 0; u < U; ++u) w[u] = fw[o + q0 + 64 * u + lane];
    bool hit[U];
    uint64_t m[U], any = 0;
#pragma unroll
    for(int u = 0; u < U; ++u) {
      const int64_t q = q0 + 64 * u + lane;
      const int e = popc64(w[u] ^ xw);
      hit[u] = q < own ? (e >= emin && e <= 4) : e == 2;
      m[u] = ballot(hit[u]);
      any |= m[u];
    }
    if(FILL && any) {
      int32_t id[U];
#pragma unroll
      for(int u = 0; u < U; ++u) id[u] = fi[o + (hit[u] ? q0 + 64 * u + lane : 0)];
      int64_t c = cnt;
#pragma unroll
      for(int u = 0; u < U; ++u) {
        if(hit[u]) col[base + c + lanes_below(m[u])] = id[u];
        c += popc64(m[u]);
      }
    }
#pragma unroll
    for(int u = 0; u < U; ++u) cnt += popc64(m[u]);
  }
  int64_t q0 = ngrp * 64 * U;
  for(; q0 < len; q0 += 64) {
    const int64_t q = q0 + lane;
    bool hit = false;
    if(q < len) {
      const int e = popc64(fw[o + q] ^ xw);
      hit = q < own ? (e >= emin && e <= 4) : e == 2;
    }
    const uint64_t m = ballot(hit);
    if(FILL && hit) col[base + cnt + lanes_below(m)] = fi[o + q];
    cnt += popc64(m);
  }
  if(!FILL && lane == 0) {
    if(SIDE)
      cnt_b[row] = (uint32_t)cnt;
    else
      counts[row] = cnt_b[row] + (uint32_t)cnt;
  }
}

// begin offsets of rows [r0, r0 + m] relative to rowptr[r0]
__global__ void rel_offsets_kernel(const int64_t* __restrict__ rowptr, int64_t r0, int64_t m,
                                   uint32_t* __restrict__ rel) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if(r <= m) rel[r] = (uint32_t)(rowptr[r0 + r] - rowptr[r0]);
}

unsigned grid_of(int64_t items, int per_block) {
  return (unsigned)((items + per_block - 1) / per_block);
}

void exclusive_scan_i64(mhip_context* ctx, const int64_t* in, int64_t* out, int64_t n1) {
  size_t tmp = 0;
  MHIP_HIP_CHECK(rocprim::exclusive_scan(nullptr, tmp, in, out, (int64_t)0, (size_t)n1,
                                         rocprim::plus<int64_t>(), ctx->stream));
  ctx->scratch_c.reserve(tmp + 16);
  MHIP_HIP_CHECK(rocprim::exclusive_scan(ctx->scratch_c.p, tmp, in, out, (int64_t)0, (size_t)n1,
                                         rocprim::plus<int64_t>(), ctx->stream));
}

// One spin's groups up to the flat offsets (the flat lists are written
// once both sides' sizes are known). keys: this spin's strings; other: the
// other spin's strings; idx: 0..n-1.
void build_side(mhip_context* ctx, int W, int64_t n, const uint64_t* keys, const uint64_t* other,
                const int32_t* idx, uint64_t uni, GroupSide& S, DevBuf& mword, DevBuf& sm,
                DevBuf& nbr, DevBuf& noff) {
  hipStream_t st = ctx->stream;
  DevBuf sk(8 * n), head(4 * n), gid(4 * n);
  sm.alloc(4 * n);
  size_t tmp = 0;
  const unsigned end_bit = W == 1 ? 32u : 64u;
  MHIP_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tmp, keys, sk.as<uint64_t>(), idx,
                                           sm.as<int32_t>(), (size_t)n, 0u, end_bit, st));
  ctx->scratch_c.reserve(tmp + 16);
  MHIP_HIP_CHECK(rocprim::radix_sort_pairs(ctx->scratch_c.p, tmp, keys, sk.as<uint64_t>(), idx,
                                           sm.as<int32_t>(), (size_t)n, 0u, end_bit, st));
  hipLaunchKernelGGL(head_kernel, dim3(grid_of(n, 256)), dim3(256), 0, st, sk.as<uint64_t>(), n,
                     head.as<int32_t>());
  tmp = 0;
  MHIP_HIP_CHECK(rocprim::inclusive_scan(nullptr, tmp, head.as<int32_t>(), gid.as<int32_t>(),
                                         (size_t)n, rocprim::plus<int32_t>(), st));
  ctx->scratch_c.reserve(tmp + 16);
  MHIP_HIP_CHECK(rocprim::inclusive_scan(ctx->scratch_c.p, tmp, head.as<int32_t>(),
                                         gid.as<int32_t>(), (size_t)n, rocprim::plus<int32_t>(),
                                         st));
  int32_t ng = 0;
  d2h(ctx, &ng, gid.as<int32_t>() + n - 1, 4);
  S.ng = ng;
  S.start.alloc(8 * (ng + 1));
  DevBuf ustr(8 * ng);
  S.gof.alloc(4 * n);
  mword.alloc(8 * n);
  hipLaunchKernelGGL(group_kernel, dim3(grid_of(n, 256)), dim3(256), 0, st, sk.as<uint64_t>(),
                     sm.as<int32_t>(), gid.as<int32_t>(), n, other, S.start.as<int64_t>(),
                     ustr.as<uint64_t>(), S.gof.as<int32_t>(), mword.as<uint64_t>());
  MHIP_HIP_CHECK(hipGetLastError());
  sk.release();
  head.release();
  gid.release();
  // hash table of the distinct strings (size: the power of two >= 2.5 ng)
  int64_t hsize = 1024;
  while(hsize < ng * 5 / 2) hsize <<= 1;
  DevBuf hkey(8 * hsize), hval(4 * hsize);
  MHIP_HIP_CHECK(hipMemsetAsync(hval.p, 0, 4 * hsize, st));
  hipLaunchKernelGGL(str_insert_kernel, dim3(grid_of(ng, 256)), dim3(256), 0, st,
                     ustr.as<uint64_t>(), (int64_t)ng, (uint64_t)(hsize - 1), hkey.as<uint64_t>(),
                     hval.as<int32_t>());
  MHIP_HIP_CHECK(hipGetLastError());
  // neighbour groups: counts, offsets, lists
  DevBuf ncount(8 * (ng + 1));
  S.nsize.alloc(8 * (ng + 1));
  noff.alloc(8 * (ng + 1));
  MHIP_HIP_CHECK(hipMemsetAsync(ncount.as<int64_t>() + ng, 0, 8, st));
  hipLaunchKernelGGL(neighbour_kernel<false>, dim3(grid_of(ng, kWaves)), dim3(256), 0, st,
                     ustr.as<uint64_t>(), S.start.as<int64_t>(), (int64_t)ng, uni,
                     hkey.as<uint64_t>(), hval.as<int32_t>(), (uint64_t)(hsize - 1),
                     ncount.as<int64_t>(), S.nsize.as<int64_t>(), nullptr, nullptr);
  MHIP_HIP_CHECK(hipGetLastError());
  exclusive_scan_i64(ctx, ncount.as<int64_t>(), noff.as<int64_t>(), ng + 1);
  int64_t nn = 0;
  d2h(ctx, &nn, noff.as<int64_t>() + ng, 8);
  nbr.alloc(4 * std::max<int64_t>(nn, 1));
  hipLaunchKernelGGL(neighbour_kernel<true>, dim3(grid_of(ng, kWaves)), dim3(256), 0, st,
                     ustr.as<uint64_t>(), S.start.as<int64_t>(), (int64_t)ng, uni,
                     hkey.as<uint64_t>(), hval.as<int32_t>(), (uint64_t)(hsize - 1), nullptr,
                     nullptr, noff.as<int64_t>(), nbr.as<int32_t>());
  // flat offsets
  DevBuf len(8 * (ng + 1));
  S.foff.alloc(8 * (ng + 1));
  hipLaunchKernelGGL(flat_len_kernel, dim3(grid_of(ng + 1, 256)), dim3(256), 0, st,
                     S.start.as<int64_t>(), S.nsize.as<int64_t>(), (int64_t)ng,
                     len.as<int64_t>());
  MHIP_HIP_CHECK(hipGetLastError());
  exclusive_scan_i64(ctx, len.as<int64_t>(), S.foff.as<int64_t>(), ng + 1);
  d2h(ctx, &S.total, S.foff.as<int64_t>() + ng, 8);
}

SideView view(const GroupSide& S) {
  return SideView{S.start.as<int64_t>(), S.nsize.as<int64_t>(), S.foff.as<int64_t>(),
                  S.gof.as<int32_t>(), S.fw.as<uint64_t>(), S.fi.as<int32_t>(),
                  S.perm.as<int32_t>()};
}

}  // namespace

bool prepare_spin_groups(mhip_context* ctx, int W, const uint64_t* d, int64_t n,
                         double max_bytes, SpinGroups& g) {
  if(n < 2) return false;
  hipStream_t st = ctx->stream;
  DevBuf ka(8 * n), kb(8 * n), idx(4 * n), ctl(16), cnt(16);
  hipLaunchKernelGGL(init_ctl_kernel, dim3(1), dim3(64), 0, st,
                     ctl.as<unsigned long long>(), cnt.as<int>());
  if(W == 1)
    hipLaunchKernelGGL(split_kernel<1>, dim3(grid_of(n, 256)), dim3(256), 0, st, d, n,
                       ka.as<uint64_t>(), kb.as<uint64_t>(), idx.as<int32_t>(),
                       ctl.as<unsigned long long>(), cnt.as<int>());
  else
    hipLaunchKernelGGL(split_kernel<2>, dim3(grid_of(n, 256)), dim3(256), 0, st, d, n,
                       ka.as<uint64_t>(), kb.as<uint64_t>(), idx.as<int32_t>(),
                       ctl.as<unsigned long long>(), cnt.as<int>());
  MHIP_HIP_CHECK(hipGetLastError());
  uint64_t uni[2] = {0, 0};
  int c[4] = {0, 0, 0, 0};
  d2h(ctx, uni, ctl.p, 16);
  d2h(ctx, c, cnt.p, 16);
  // one alpha and one beta electron count: distances are even per spin
  if(c[0] != c[1] || c[2] != c[3]) return false;
  DevBuf mwA, smA, nbrA, noffA, mwB, smB, nbrB, noffB;
  build_side(ctx, W, n, ka.as<uint64_t>(), kb.as<uint64_t>(), idx.as<int32_t>(), uni[0], g.A,
             mwA, smA, nbrA, noffA);
  build_side(ctx, W, n, kb.as<uint64_t>(), ka.as<uint64_t>(), idx.as<int32_t>(), uni[1], g.B,
             mwB, smB, nbrB, noffB);
  ka.release();
  kb.release();
  idx.release();
  // the flat lists: 12 bytes per entry
  if(12.0 * (double)(g.A.total + g.B.total) > max_bytes) return false;
  for(int s = 0; s < 2; ++s) {
    GroupSide& S = s == 0 ? g.A : g.B;
    DevBuf& mw = s == 0 ? mwA : mwB;
    DevBuf& sm = s == 0 ? smA : smB;
    DevBuf& nb = s == 0 ? nbrA : nbrB;
    DevBuf& no = s == 0 ? noffA : noffB;
    S.fw.alloc(8 * std::max<int64_t>(S.total, 1));
    S.fi.alloc(4 * std::max<int64_t>(S.total, 1));
    hipLaunchKernelGGL(flatten_kernel, dim3(grid_of(S.ng, kWaves)), dim3(256), 0, st,
                       S.start.as<int64_t>(), mw.as<uint64_t>(), sm.as<int32_t>(),
                       no.as<int64_t>(), nb.as<int32_t>(), S.foff.as<int64_t>(), S.ng,
                       S.fw.as<uint64_t>(), S.fi.as<int32_t>());
    MHIP_HIP_CHECK(hipGetLastError());
    S.perm = std::move(sm);
  }
  g.cnt_b.alloc(4 * n);
  MHIP_HIP_CHECK(hipStreamSynchronize(st));
  g.n = n;
  return true;
}

namespace {
template <bool FILL>
void launch_scans(mhip_context* ctx, int W, const SpinGroups& g, const uint64_t* d, int64_t n,
                  uint32_t* counts, const int64_t* rowptr, int32_t* col) {
  const int64_t nblocks = (n + kWaves - 1) / kWaves;
  // whole rounds of chunks over the XCDs
  const int64_t per = (int64_t)kXcds * kChunkBlocks;
  const dim3 grid((unsigned)((nblocks + per - 1) / per * per));
  uint32_t* cb = g.cnt_b.as<uint32_t>();
  const SideView A = view(g.A), B = view(g.B);
  hipStream_t st = ctx->stream;
  if(W == 1) {
    hipLaunchKernelGGL((group_scan_kernel<1, FILL, 1>), grid, dim3(256), 0, st, d, n, nblocks, A,
                       B, cb, counts, rowptr, col);
    hipLaunchKernelGGL((group_scan_kernel<1, FILL, 0>), grid, dim3(256), 0, st, d, n, nblocks, A,
                       B, cb, counts, rowptr, col);
  } else {
    hipLaunchKernelGGL((group_scan_kernel<2, FILL, 1>), grid, dim3(256), 0, st, d, n, nblocks, A,
                       B, cb, counts, rowptr, col);
    hipLaunchKernelGGL((group_scan_kernel<2, FILL, 0>), grid, dim3(256), 0, st, d, n, nblocks, A,
                       B, cb, counts, rowptr, col);
  }
  MHIP_HIP_CHECK(hipGetLastError());
}
}  // namespace

void spin_group_count(mhip_context* ctx, int W, const SpinGroups& g, const uint64_t* d,
                      int64_t n, uint32_t* counts) {
  launch_scans<false>(ctx, W, g, d, n, counts, nullptr, nullptr);
}

void spin_group_fill(mhip_context* ctx, int W, const SpinGroups& g, const uint64_t* d, int64_t n,
                     const int64_t* rowptr, int32_t* col) {
  launch_scans<true>(ctx, W, g, d, n, nullptr, rowptr, col);
}

void sort_row_columns(mhip_context* ctx, int64_t n, const int64_t* rowptr, int64_t nnz,
                      int64_t ncols, DevBuf& col) {
  if(nnz == 0 || n == 0) return;
  hipStream_t st = ctx->stream;
  unsigned bits = 1;
  while(bits < 31 && (1ll << bits) < ncols) ++bits;
  // rows in pieces of at most 2^31 - 1 entries (the sort indexes in 32
  // bits); MHIP_ROW_SORT_PIECE lowers the piece size so tests reach the
  // multi-piece path on short lists
  int64_t piece = (1ll << 31) - 1;
  if(const char* env = std::getenv("MHIP_ROW_SORT_PIECE"))
    piece = std::max<int64_t>(1, std::min<int64_t>(piece, std::atoll(env)));
  std::vector<int64_t> cuts{0}, base{0};
  if(nnz > piece) {
    std::vector<int64_t> rp(n + 1);
    d2h(ctx, rp.data(), rowptr, 8 * (n + 1));
    int64_t r0 = 0;
    for(int64_t r = 1; r <= n; ++r)
      if(rp[r] - rp[r0] > piece) {
        MHIP_REQUIRE(r - 1 > r0, MHIP_ERR_INDEX_OVERFLOW, "row longer than the sort piece");
        cuts.push_back(r - 1);
        base.push_back(rp[r - 1]);
        r0 = r - 1;
      }
  }
  cuts.push_back(n);
  base.push_back(nnz);
  DevBuf alt(4 * nnz), rel(4 * (n + 1));
  int32_t* where = nullptr;  // buffer holding the sorted columns
  for(size_t k = 0; k + 1 < cuts.size(); ++k) {
    const int64_t r0 = cuts[k], m = cuts[k + 1] - cuts[k];
    const int64_t b0 = base[k], len = base[k + 1] - base[k];
    if(len == 0) continue;
    hipLaunchKernelGGL(rel_offsets_kernel, dim3(grid_of(m + 1, 256)), dim3(256), 0, st, rowptr,
                       r0, m, rel.as<uint32_t>());
    rocprim::double_buffer<int32_t> kb(col.as<int32_t>() + b0, alt.as<int32_t>() + b0);
    size_t tmp = 0;
    MHIP_HIP_CHECK(rocprim::segmented_radix_sort_keys(nullptr, tmp, kb, (unsigned)len,
                                                      (unsigned)m, rel.as<uint32_t>(),
                                                      rel.as<uint32_t>() + 1, 0u, bits, st));
    ctx->scratch_c.reserve(tmp + 16);
    MHIP_HIP_CHECK(rocprim::segmented_radix_sort_keys(ctx->scratch_c.p, tmp, kb, (unsigned)len,
                                                      (unsigned)m, rel.as<uint32_t>(),
                                                      rel.as<uint32_t>() + 1, 0u, bits, st));
    int32_t* w = kb.current() == col.as<int32_t>() + b0 ? col.as<int32_t>() : alt.as<int32_t>();
    // every piece sorts over the same bits, so all land in one buffer
    MHIP_REQUIRE(where == nullptr || where == w, MHIP_ERR_HIP, "sorted pieces in two buffers");
    where = w;
  }
  MHIP_HIP_CHECK(hipGetLastError());
  if(where == alt.as<int32_t>()) std::swap(col, alt);
  // alt (the unsorted or scratch columns) is freed on return: wait for the sort
  MHIP_HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace mhip
