// K1: Hamiltonian generation on the device.
//
// Replaces DoubleLoopHamiltonianGenerator::make_csr_hamiltonian_block_
// (double_loop.hpp:30-98): for every bra row, every ket j with
// popcount(bra ^ ket) <= 4 is a candidate; the element is evaluated with the
// Slater-Condon rules and kept when |h| > thresh, columns ascending.
//
// Device pipeline (all HBM resident):
//   1. bounds  : AND/OR of every 64-ket tile and every 64-tile superblock.
//                popcount(b & ~OR) + popcount(AND & ~b) is a lower bound of
//                popcount(b ^ k) over the block, so blocks with bound > 4
//                are skipped without touching their kets.
//   2. count   : one wavefront handles R consecutive bra rows (bras held in
//                scalar registers); lanes screen 64 superblocks / 64 tiles at
//                a time, then test 64 kets per step with XOR + popcount and
//                count hits with a wavefront ballot. By default the same walk
//                also stages the columns in pooled chunks (kStage).
//   3. scan    : device exclusive scan of the row counts -> row pointers.
//   4. fill    : a gather of the staged chunks into the CSR, or (pool
//                overflow) the same walk again writing ascending column
//                indices at rowptr + running count + ballot prefix (mbcnt).
//   5. element : one wavefront per row evaluates the Slater-Condon element
//                of each stored column (full lane utilisation) and counts
//                |h| > thresh.
//   6. compact : only when some element fell at or below the threshold,
//                a second scan + a stable ballot compaction drops them.
//
// Complete-space lists (a product of spin strings, see mhip_product.hip)
// replace steps 1, 2, 4 and 5 by a closed-form count and a fill pass that
// walks the spin-string neighbour tables and evaluates the elements in the
// same pass; scan and compaction are shared.
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <memory>

#include "mhip_device.hpp"
#include "mhip_internal.hpp"

namespace mhip {
namespace {

constexpr int kTile = 64;   // kets per tile (one per lane)
constexpr int kSuper = 64;  // tiles per superblock
constexpr int kXcdChunk = 64;  // scan blocks per chunk dealt round-robin to the 8 XCDs
constexpr int kRows = 8;    // bra rows per wavefront in the scan passes
constexpr int kWavesPerBlock = 4;
constexpr int64_t kMinGroupRows = 32768;  // AUTO: spin-group scan from this list length

__device__ __forceinline__ uint64_t uniform_u64(uint64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// AND / OR reduction of groups of 64 items (W words each); one wavefront
// per output group. Used for ket tiles (in_and == in_or == kets) and for
// superblocks (over tile bounds).
template <int W>
__global__ __launch_bounds__(256) void bounds_kernel(
    const uint64_t* __restrict__ in_and, const uint64_t* __restrict__ in_or,
    int64_t nin, int64_t nout, uint64_t* __restrict__ out_and,
    uint64_t* __restrict__ out_or) {
  const int64_t g = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if(g >= nout) return;
  const uint32_t lane = threadIdx.x & 63;
  const int64_t i = g * 64 + lane;
#pragma unroll
  for(int w = 0; w < W; ++w) {
    uint64_t a = ~0ull, o = 0ull;
    if(i < nin) {
      a = in_and[i * W + w];
      o = in_or[i * W + w];
    }
#pragma unroll
    for(int off = 32; off > 0; off >>= 1) {
      a &= __shfl_xor(a, off, 64);
      o |= __shfl_xor(o, off, 64);
    }
    if(lane == 0) {
      out_and[g * W + w] = a;
      out_or[g * W + w] = o;
    }
  }
}

// Value of a per-lane 64-bit register in lane l (wave-uniform result).
__device__ __forceinline__ uint64_t lane_value(uint64_t x, uint32_t l) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)x, l);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

template <int W>
__device__ __forceinline__ int excitation_bound(const uint64_t (&b)[W],
                                                const uint64_t (&a)[W],
                                                const uint64_t (&o)[W]) {
  int lb = 0;
#pragma unroll
  for(int w = 0; w < W; ++w) lb += popc64(b[w] & ~o[w]) + popc64(a[w] & ~b[w]);
  return lb;
}

// Pass modes of the tiled scan. kStage appends every row's columns, in
// scan order (ascending), to chunks of kChunk entries drawn from a device
// pool with one vector atomic per chunk; a row's chunks form a chain
// (head -> next ... -> -1, `used` entries each) that stage_gather_kernel
// copies into the CSR once the row counts are scanned. One scan instead of
// two; when the pool runs out the pass reports it and the build falls back
// to count + fill.
enum ScanMode : int { kCount = 0, kFill = 1, kStage = 2 };
constexpr int kChunk = 128;
constexpr int kChunksPerGrab = 8;  // chunks a wavefront takes per atomic
struct StagePool {
  int32_t* col;                // cap * kChunk entries
  int64_t cap;                 // chunks
  unsigned long long* next;    // chunk allocator
  int64_t* head;               // first chunk of each row, -1 when empty
  int64_t* chunk_next;         // next chunk of the row, -1 at its end
  uint32_t* used;              // entries in each chunk
  int* overflow;               // set when a chunk was refused
  // state carried between the launches of a ket-chunked scan: per row the
  // count, current chunk and its fill; per wavefront the grabbed chunks
  uint32_t* st_cnt;
  int64_t* st_cur;
  uint32_t* st_fil;
  int64_t* st_grab;
};

// Structural count (kCount), fill (kFill) or staged (kStage) pass.
// Without a work-item table, wave w owns rows [w*R, w*R+R) and scans kets
// [kc_lo, kc_hi). With one (batched instances), item w = {row0, nrows,
// ket_lo, ket_hi} names rows of one instance and the ket range of that
// instance; columns stay absolute indices, so a batch yields one
// block-diagonal CSR. A staged scan of a long list runs as several launches
// over consecutive ket chunks sized to one XCD's L2 (chunk_flags: 1 = first,
// 2 = last launch): every row's hits of chunk c are appended after those of
// chunk c - 1, so columns stay ascending, and the row and wavefront state
// (count, chain end, grabbed chunks) is stored between launches.
template <int W, int R, int MODE>
__global__ __launch_bounds__(256, MODE == kStage ? (W == 1 ? 4 : 3) : 1) void scan_pass_kernel(
    const uint64_t* __restrict__ bra, int64_t nbra,
    const uint64_t* __restrict__ ket, int64_t nket,
    const uint64_t* __restrict__ tand, const uint64_t* __restrict__ tor,
    int64_t ntiles, const uint64_t* __restrict__ sand,
    const uint64_t* __restrict__ sor, int64_t nsuper, int64_t nwaves,
    int64_t nblocks, const int64_t* __restrict__ items,
    uint32_t* __restrict__ row_counts, const int64_t* __restrict__ rowptr,
    int32_t* __restrict__ colind, StagePool P, int64_t kc_lo, int64_t kc_hi, int chunk_flags) {
  constexpr bool FILL = MODE == kFill;
  const bool first = chunk_flags & 1, last = chunk_flags & 2;
  // XCD-aware mapping: hardware dispatches block b to XCD b % 8. Logical
  // blocks (consecutive rows, which share spin strings and therefore ket
  // tiles) come in chunks of kXcdChunk blocks, chunk c on XCD c % 8: each
  // XCD's L2 sees contiguous rows, and every XCD samples the whole list (the
  // work per row varies along a selected-CI list: one contiguous eighth per
  // XCD left the first XCD with twice the mean).
  const int64_t xcd = blockIdx.x % 8, k = blockIdx.x / 8;
  const int64_t lblk = (xcd + 8 * (k / kXcdChunk)) * kXcdChunk + k % kXcdChunk;
  if(lblk >= nblocks) return;
  const int64_t wave = lblk * kWavesPerBlock + (threadIdx.x >> 6);
  if(wave >= nwaves) return;
  const uint32_t lane = threadIdx.x & 63;
  int64_t row0 = wave * R, nrows = R, klo = kc_lo, khi = kc_hi;
  if(items) {
    row0 = (int64_t)uniform_u64((uint64_t)items[4 * wave + 0]);
    nrows = (int64_t)uniform_u64((uint64_t)items[4 * wave + 1]);
    klo = (int64_t)uniform_u64((uint64_t)items[4 * wave + 2]);
    khi = (int64_t)uniform_u64((uint64_t)items[4 * wave + 3]);
  }
  const int64_t rend = row0 + nrows < nbra ? row0 + nrows : nbra;

  uint64_t b[R][W];
  uint32_t live = 0;
  int64_t base[R];
  uint32_t cnt[R];
  int64_t cur[R];   // kStage: the row's current chunk (-1 before its first)
  uint32_t fil[R];  // kStage: entries used in it
  bool dead = false;  // kStage: the pool ran out; counting goes on
  // kStage: chunks [grab, grab_end) taken from the pool and not yet used (one
  // atomic per kChunksPerGrab chunks keeps the shared counter uncontended)
  int64_t grab = 0, grab_end = 0;
#pragma unroll
  for(int r = 0; r < R; ++r) {
    const int64_t row = row0 + r;
    uint64_t nz = 0;
#pragma unroll
    for(int w = 0; w < W; ++w) {
      b[r][w] = row < rend ? uniform_u64(bra[row * W + w]) : 0ull;
      nz |= b[r][w];
    }
    if(nz) live |= 1u << r;  // zero bras give empty rows (double_loop.hpp:50)
    cnt[r] = 0;
    base[r] = 0;
    cur[r] = -1;
    fil[r] = 0;
    if constexpr(FILL) {
      if(row < rend) base[r] = (int64_t)uniform_u64((uint64_t)rowptr[row]);
    }
  }
  if(!live || khi <= klo) {
    if constexpr(!FILL) {
      if(lane == 0 && first)
        for(int r = 0; r < R; ++r)
          if(row0 + r < rend) {
            row_counts[row0 + r] = 0;
            if constexpr(MODE == kStage) P.head[row0 + r] = -1;
          }
    }
    return;
  }
  if constexpr(MODE == kStage) {
    if(!first) {  // resume the rows' chains and the wavefront's grabbed chunks
#pragma unroll
      for(int r = 0; r < R; ++r) {
        const int64_t row = row0 + r;
        if(row < rend) {
          cnt[r] = (uint32_t)uniform_u64((uint64_t)P.st_cnt[row]);
          cur[r] = (int64_t)uniform_u64((uint64_t)P.st_cur[row]);
          fil[r] = (uint32_t)uniform_u64((uint64_t)P.st_fil[row]);
        }
      }
      grab = (int64_t)uniform_u64((uint64_t)P.st_grab[2 * wave]);
      grab_end = (int64_t)uniform_u64((uint64_t)P.st_grab[2 * wave + 1]);
      dead = P.overflow[0] != 0;
    }
  }
  // kStage: record one ballot's hits of row r (m = ballot, hit = this lane's
  // bit): the part that fits the current chunk goes there, the rest opens
  // the next chunk of the row's chain
  auto stage = [&](int r, uint64_t m, bool hit, int32_t j) {
    const uint32_t k = popc64(m);
    if(!k) return;
    cnt[r] += k;
    if(dead) return;
    const uint32_t room = cur[r] < 0 ? 0u : (uint32_t)kChunk - fil[r];
    const uint32_t pos = lanes_below(m);
    int64_t nc = -1;
    if(k > room) {
      if(grab >= grab_end) {
        unsigned long long got = 0;
        if(lane == 0) got = atomicAdd(P.next, (unsigned long long)kChunksPerGrab);
        grab = (int64_t)uniform_u64((uint64_t)__shfl(got, 0, 64));
        grab_end = grab + kChunksPerGrab;
        if(grab_end > P.cap) {
          if(lane == 0) P.overflow[0] = 1;
          dead = true;
          return;
        }
      }
      nc = grab++;
      if(lane == 0) {
        if(cur[r] < 0) {
          P.head[row0 + r] = nc;
        } else {
          P.chunk_next[cur[r]] = nc;
          P.used[cur[r]] = (uint32_t)kChunk;
        }
      }
    }
    if(hit) {
      if(pos < room)
        P.col[cur[r] * kChunk + fil[r] + pos] = j;
      else
        P.col[nc * kChunk + (pos - room)] = j;
    }
    if(k > room) {
      cur[r] = nc;
      fil[r] = k - room;
    } else {
      fil[r] += k;
    }
  };
  const int64_t s_lo = klo / (kTile * kSuper), s_hi = (khi + kTile * kSuper - 1) / (kTile * kSuper);
  const int64_t t_lo = klo / kTile, t_hi = (khi + kTile - 1) / kTile;

  for(int64_t g = s_lo; g < s_hi; g += 64) {
    const int64_t s = g + lane;
    const bool sv = s < s_hi;
    uint64_t sa[W], so[W];
#pragma unroll
    for(int w = 0; w < W; ++w) {
      sa[w] = sv ? sand[s * W + w] : ~0ull;
      so[w] = sv ? sor[s * W + w] : 0ull;
    }
    uint64_t smask[R];
    uint64_t sany = 0;
#pragma unroll
    for(int r = 0; r < R; ++r) {
      smask[r] = ((live >> r) & 1u) ? ballot(sv && excitation_bound<W>(b[r], sa, so) <= 4) : 0ull;
      sany |= smask[r];
    }
    // the masks are wavefront-uniform (ballots): kept in scalar registers
    sany = uniform_u64(sany);
    while(sany) {
      const uint32_t sb = __builtin_amdgcn_readfirstlane(ctz64(sany));
      sany &= sany - 1;
      const int64_t S = g + sb;
      const int64_t t = S * kSuper + lane;
      const bool tv = t >= t_lo && t < t_hi;
      uint64_t ta[W], to[W];
#pragma unroll
      for(int w = 0; w < W; ++w) {
        ta[w] = tv ? tand[t * W + w] : ~0ull;
        to[w] = tv ? tor[t * W + w] : 0ull;
      }
      uint64_t tm[R];
      uint64_t tany = 0;
#pragma unroll
      for(int r = 0; r < R; ++r) {
        tm[r] = ((smask[r] >> sb) & 1ull) ? ballot(tv && excitation_bound<W>(b[r], ta, to) <= 4) : 0ull;
        tany |= tm[r];
      }
      // kets of surviving tile tb (one per lane; zero outside [klo, khi))
      auto load_tile = [&](uint32_t tb, uint64_t (&k)[W]) {
        const int64_t j = (S * kSuper + tb) * kTile + lane;
        const bool kv = j >= klo && j < khi;
#pragma unroll
        for(int w = 0; w < W; ++w) k[w] = kv ? ket[j * W + w] : 0ull;
      };
      auto test_tile = [&](uint32_t tb, const uint64_t (&k)[W]) {
        const int64_t j = (S * kSuper + tb) * kTile + lane;
        uint64_t knz = 0;
#pragma unroll
        for(int w = 0; w < W; ++w) knz |= k[w];
        // Tiles whose kets all share one alpha string (AND == OR on the
        // alpha bits): the alpha excitation is a per-row scalar, so each
        // ket costs one XOR + popcount of the beta string.
        const uint64_t tA = lane_value(ta[0], tb), tO = lane_value(to[0], tb);
        const uint64_t amask = W == 1 ? 0xffffffffull : ~0ull;
        const bool one_alpha = ((tA ^ tO) & amask) == 0;
#pragma unroll
        for(int r = 0; r < R; ++r) {
          if((tm[r] >> tb) & 1ull) {
            bool hit;
            if(one_alpha) {
              const int thr = 4 - popc64((b[r][0] ^ tA) & amask);
              const uint64_t bb = W == 1 ? (b[r][0] >> 32) : b[r][W - 1];
              const uint64_t kb = W == 1 ? (k[0] >> 32) : k[W - 1];
              hit = knz && popc64(bb ^ kb) <= thr;
            } else {
              int pc = 0;
#pragma unroll
              for(int w = 0; w < W; ++w) pc += popc64(b[r][w] ^ k[w]);
              hit = knz && pc <= 4;
            }
            const uint64_t m = ballot(hit);
            if constexpr(FILL) {
              if(hit) colind[base[r] + lanes_below(m)] = (int32_t)j;
              base[r] += popc64(m);
            } else if constexpr(MODE == kStage) {
              stage(r, m, hit, (int32_t)j);
            } else {
              cnt[r] += (uint32_t)popc64(m);
            }
          }
        }
      };
      // the next surviving tile's kets are loaded while this one is tested
      tany = uniform_u64(tany);
      if(tany) {
        uint32_t tb = __builtin_amdgcn_readfirstlane(ctz64(tany));
        tany &= tany - 1;
        uint64_t k[W];
        load_tile(tb, k);
        while(true) {
          const bool more = tany != 0;
          const uint32_t tn = more ? __builtin_amdgcn_readfirstlane(ctz64(tany)) : 0u;
          uint64_t kn[W];
          if(more) load_tile(tn, kn);
          test_tile(tb, k);
          if(!more) break;
          tany &= tany - 1;
          tb = tn;
#pragma unroll
          for(int w = 0; w < W; ++w) k[w] = kn[w];
        }
      }
    }
  }
  if constexpr(MODE == kStage) {
    if(!last) {  // keep the state for the next chunk's launch
      if(lane == 0) {
        for(int r = 0; r < R; ++r)
          if(row0 + r < rend) {
            P.st_cnt[row0 + r] = cnt[r];
            P.st_cur[row0 + r] = cur[r];
            P.st_fil[row0 + r] = fil[r];
          }
        P.st_grab[2 * wave] = grab;
        P.st_grab[2 * wave + 1] = grab_end;
      }
      return;
    }
  }
  if constexpr(!FILL) {
    if(lane == 0)
      for(int r = 0; r < R; ++r)
        if(row0 + r < rend) {
          row_counts[row0 + r] = cnt[r];
          if constexpr(MODE == kStage) {
            if(cur[r] < 0) {
              P.head[row0 + r] = -1;
            } else {
              P.used[cur[r]] = fil[r];
              P.chunk_next[cur[r]] = -1;
            }
          }
        }
  }
}

// Staged columns -> CSR: one wavefront per row walks the row's chunk chain.
__global__ __launch_bounds__(256) void stage_gather_kernel(
    const int64_t* __restrict__ head, const int64_t* __restrict__ chunk_next,
    const uint32_t* __restrict__ used, const int32_t* __restrict__ pool,
    const int64_t* __restrict__ rowptr, int64_t nrows, int32_t* __restrict__ colind) {
  const int64_t row = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if(row >= nrows) return;
  const uint32_t lane = threadIdx.x & 63;
  int64_t out = rowptr[row];
  for(int64_t c = head[row]; c >= 0; c = chunk_next[c]) {
    const uint32_t u = used[c];
    for(uint32_t e = lane; e < u; e += 64) colind[out + e] = pool[c * kChunk + e];
    out += u;
  }
}

// OR of all determinant words (range validation of orbital indices).
template <int W>
__global__ __launch_bounds__(256) void or_all_kernel(const uint64_t* __restrict__ d,
                                                      int64_t n,
                                                      unsigned long long* __restrict__ out) {
  uint64_t o[W];
#pragma unroll
  for(int w = 0; w < W; ++w) o[w] = 0;
  for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
      i += (int64_t)gridDim.x * blockDim.x)
#pragma unroll
    for(int w = 0; w < W; ++w) o[w] |= d[i * W + w];
#pragma unroll
  for(int w = 0; w < W; ++w) {
#pragma unroll
    for(int off = 32; off > 0; off >>= 1) o[w] |= __shfl_xor(o[w], off, 64);
    if((threadIdx.x & 63) == 0) atomicOr(out + w, (unsigned long long)o[w]);
  }
}

// Widen uint32 row counts into the int64 scan input (last entry 0).
__global__ void widen_counts_kernel(const uint32_t* __restrict__ c, int64_t n,
                                    int64_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) out[i] = c[i];
  if(i == n) out[n] = 0;
}

// Diagonal element of every determinant (thread per determinant).
template <int W>
__global__ __launch_bounds__(256) void diag_kernel(IntsView I0,
                                                    const uint64_t* __restrict__ dets,
                                                    int64_t n, double* __restrict__ diag,
                                                    const IntsView* __restrict__ views,
                                                    const int32_t* __restrict__ row_inst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const IntsView I = views ? views[row_inst[i]] : I0;
  const Spins<W> s = load_spins<W>(dets, i);
  diag[i] = me_diag(I, s.a, s.b);
}

// Element pass: one wavefront per bra row.
template <int W>
__global__ __launch_bounds__(256) void element_kernel(
    IntsView I0, const uint64_t* __restrict__ bra, int64_t nbra,
    const uint64_t* __restrict__ ket, const double* __restrict__ diag,
    const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
    double* __restrict__ nzval, double thresh, int64_t* __restrict__ kept,
    int* __restrict__ dropped, const IntsView* __restrict__ views,
    const int32_t* __restrict__ row_inst) {
  const int64_t row = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if(row >= nbra) return;
  const IntsView I = views ? views[row_inst[row]] : I0;
  const uint32_t lane = threadIdx.x & 63;
  const int64_t st = rowptr[row], en = rowptr[row + 1];
  if(st == en) {
    if(lane == 0) kept[row] = 0;
    return;
  }
  Spins<W> bs = load_spins<W>(bra, row);
  bs.a = uniform_u64(bs.a);
  bs.b = uniform_u64(bs.b);
  const double d = diag[row];
  int64_t nk = 0;
  // Single excitations (a sum over the bra's occupied orbitals each) are
  // queued per spin in LDS and evaluated 64 at a time, so the sums run once
  // per 64 singles rather than once per 64 columns with most lanes idle;
  // the other elements are evaluated in place. Same values, any order.
  __shared__ int32_t pend[kWavesPerBlock][2][64];
  const int wv = threadIdx.x >> 6;
  int npend[2] = {0, 0};  // wavefront-uniform
  auto flush = [&](int sp) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    if(lane < (uint32_t)npend[sp]) {
      const int64_t j = st + pend[wv][sp][lane];
      const Spins<W> ks = load_spins<W>(ket, colind[j]);
      const double h = sp == 0 ? me_2(I, bs.a, ks.a, bs.a ^ ks.a, bs.a, bs.b)
                               : me_2(I, bs.b, ks.b, bs.b ^ ks.b, bs.b, bs.a);
      nzval[j] = h;
      nk += (fabs(h) > thresh) ? 1 : 0;  // double_loop.hpp:76
    }
    npend[sp] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  };
  for(int64_t j0 = st; j0 < en; j0 += 64) {
    const int64_t j = j0 + lane;
    bool single[2] = {false, false};
    if(j < en) {
      const Spins<W> ks = load_spins<W>(ket, colind[j]);
      const int ca = popc64(bs.a ^ ks.a), cb = popc64(bs.b ^ ks.b);
      single[0] = ca == 2 && cb == 0;
      single[1] = cb == 2 && ca == 0;
      if(!single[0] && !single[1]) {
        const double h = matrix_element(I, bs.a, bs.b, ks.a, ks.b, d);
        nzval[j] = h;
        nk += (fabs(h) > thresh) ? 1 : 0;  // double_loop.hpp:76
      }
    }
#pragma unroll
    for(int sp = 0; sp < 2; ++sp) {
      const uint64_t m = ballot(single[sp]);
      const int c = popc64(m);
      if(npend[sp] + c > 64) flush(sp);
      if(single[sp]) pend[wv][sp][npend[sp] + (int)lanes_below(m)] = (int32_t)(j - st);
      npend[sp] += c;
    }
  }
  flush(0);
  flush(1);
  nk = wave_sum(nk);
  if(lane == 0) {
    kept[row] = nk;
    if(nk != en - st) atomicOr(dropped, 1);
  }
}

// Stable compaction of the kept entries (one wavefront per row).
__global__ __launch_bounds__(256) void compact_kernel(
    int64_t nbra, const int64_t* __restrict__ rp_old,
    const int32_t* __restrict__ col_old, const double* __restrict__ val_old,
    const int64_t* __restrict__ rp_new, int32_t* __restrict__ col_new,
    double* __restrict__ val_new, double thresh) {
  const int64_t row = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if(row >= nbra) return;
  const uint32_t lane = threadIdx.x & 63;
  const int64_t st = rp_old[row], en = rp_old[row + 1];
  int64_t out = rp_new[row];
  for(int64_t j0 = st; j0 < en; j0 += 64) {
    const int64_t j = j0 + lane;
    const double v = j < en ? val_old[j] : 0.;
    const bool keep = j < en && fabs(v) > thresh;
    const uint64_t m = ballot(keep);
    if(keep) {
      const int64_t p = out + lanes_below(m);
      col_new[p] = col_old[j];
      val_new[p] = v;
    }
    out += popc64(m);
  }
}

}  // namespace

IntsView view(const Integrals& I) {
  IntsView v;
  v.n = (int)I.norb;
  v.T = I.T.as<double>();
  v.V = I.V.as<double>();
  v.G = I.G.as<double>();
  v.G_red = I.G_red.as<double>();
  v.V_red = I.V_red.as<double>();
  v.G2_red = I.G2_red.as<double>();
  v.V2_red = I.V2_red.as<double>();
  v.pair_sym = I.pair_sym ? 1 : 0;
  return v;
}

namespace {
// Exclusive scan of n+1 int64 values (in[n] must be 0) into out.
void exclusive_scan(mhip_context* ctx, const int64_t* in, int64_t* out,
                    int64_t n1) {
  size_t tmp = 0;
  MHIP_HIP_CHECK(rocprim::exclusive_scan(nullptr, tmp, in, out, (int64_t)0,
                                         (size_t)n1, rocprim::plus<int64_t>(),
                                         ctx->stream));
  ctx->scratch_c.reserve(tmp + 16);
  MHIP_HIP_CHECK(rocprim::exclusive_scan(ctx->scratch_c.p, tmp, in, out,
                                         (int64_t)0, (size_t)n1,
                                         rocprim::plus<int64_t>(), ctx->stream));
}

}  // namespace

template <int W>
void build_impl(mhip_context* ctx, const uint64_t* d_bra, int64_t nbra,
                const uint64_t* d_ket, int64_t nket, double thresh,
                mhip_csr* csr, const BuildExtra& ex) {
  hipStream_t st = ctx->stream;
  mhip_timings& tm = ctx->timings;
  Timer t_all(st), t(st);
  t_all.start();
  const IntsView I = view(ctx->ints);

  // 1. bounds (tiled path only)
  const bool prod = ex.product || ex.bproduct;
  const bool grp = ex.groups != nullptr;
  t.start();
  const int64_t ntiles = (nket + kTile - 1) / kTile;
  const int64_t nsuper = (ntiles + kSuper - 1) / kSuper;
  DevBuf bnd;
  uint64_t *tand = nullptr, *tor = nullptr, *sand = nullptr, *sor = nullptr;
  if(!prod && !grp) {
    bnd.alloc(sizeof(uint64_t) * W * 2 * (ntiles + nsuper));
    tand = bnd.as<uint64_t>();
    tor = tand + W * ntiles;
    sand = tor + W * ntiles;
    sor = sand + W * nsuper;
    hipLaunchKernelGGL(bounds_kernel<W>, dim3((ntiles + 3) / 4), dim3(256), 0, st,
                       d_ket, d_ket, nket, ntiles, tand, tor);
    hipLaunchKernelGGL(bounds_kernel<W>, dim3((nsuper + 3) / 4), dim3(256), 0, st,
                       tand, tor, ntiles, nsuper, sand, sor);
    MHIP_HIP_CHECK(hipGetLastError());
  }
  t.stop();
  tm.ham_bounds_ms = t.ms() + ex.detect_ms;

  // 2. count
  const int64_t nwaves = ex.items ? ex.nitems : (nbra + kRows - 1) / kRows;
  const int64_t nblocks = (nwaves + kWavesPerBlock - 1) / kWavesPerBlock;
  // whole rounds of 8 chunks (the scan kernel's XCD mapping)
  const int64_t nchunks = (nblocks + kXcdChunk - 1) / kXcdChunk;
  const int64_t grid = (nchunks + 7) / 8 * 8 * kXcdChunk;
  MHIP_REQUIRE(grid < (1ll << 31), MHIP_ERR_INVALID_ARG, "too many bra rows");
  DevBuf counts32(sizeof(uint32_t) * (nbra + 1));
  DevBuf counts64(sizeof(int64_t) * (nbra + 1));
  // tiled path over one list: one staged scan with a pool of 4,096 entries
  // per bra row (at most a third of the free memory; rows longer on average
  // overflow it and take count + fill). The pool is released before the
  // values are allocated, so the peak stays that of count + fill. Batched
  // builds (work items over small per-instance ket ranges) scan cheaply and
  // keep count + fill, which runs at a higher occupancy.
  DevBuf s_pool, s_head, s_next, s_used, s_ctl, s_state;
  StagePool P{};
  bool staged = false;
  // ket chunks of the staged scan: one XCD's L2 (4 MB) holds a chunk of
  // kets (16 bytes each at W = 2) with its tile bounds; a list of more than
  // two chunks is scanned one chunk per launch (MHIP_SCAN_KET_CHUNK, a
  // multiple of 4,096 kets, overrides the chunk so tests reach the path on
  // short lists; 0 disables chunking)
  int64_t ket_chunk = W == 1 ? 262144 : 131072;
  if(const char* env = std::getenv("MHIP_SCAN_KET_CHUNK"))
    ket_chunk = std::max<int64_t>(0, std::atoll(env)) / (kTile * kSuper) * (kTile * kSuper);
  const bool chunked = !prod && !grp && !ex.items && ket_chunk > 0 && nket > 2 * ket_chunk;
  if(!prod && !grp && !ex.items) {
    size_t free_b = 0, total_b = 0;
    device_mem_info(&free_b, &total_b);
    const int64_t per_chunk = (int64_t)kChunk * 4 + 8 + 4;
    int64_t cap = std::min<int64_t>({(int64_t)(free_b / 3) / per_chunk,
                                     nbra * (4096 / kChunk) + 1024, 1ll << 27});
    // MHIP_STAGE_POOL_CHUNKS caps the pool (0: no staging), so tests can
    // force the overflow fallback and the two-pass walk
    if(const char* env = std::getenv("MHIP_STAGE_POOL_CHUNKS"))
      cap = std::min<int64_t>(cap, std::max<int64_t>(0, std::atoll(env)));
    if(cap >= 1024) {
      try {
        s_pool.alloc(sizeof(int32_t) * kChunk * cap);
        s_next.alloc(sizeof(int64_t) * cap);
        s_used.alloc(sizeof(uint32_t) * cap);
        s_head.alloc(sizeof(int64_t) * nbra);
        s_ctl.alloc(16);
        if(chunked) s_state.alloc(16 * (size_t)nbra + 16 * (size_t)nwaves);
        staged = true;
      } catch(const Error& e) {
        if(e.code != MHIP_ERR_OUT_OF_MEMORY) throw;
        (void)hipGetLastError();  // a refused hipMalloc leaves no pending error
        s_pool.release();
        s_next.release();
        s_used.release();
        s_head.release();
        s_state.release();
      }
    }
    if(staged) {
      P.col = s_pool.as<int32_t>();
      P.cap = cap;
      P.next = s_ctl.as<unsigned long long>();
      P.head = s_head.as<int64_t>();
      P.chunk_next = s_next.as<int64_t>();
      P.used = s_used.as<uint32_t>();
      P.overflow = reinterpret_cast<int*>(s_ctl.as<unsigned char>() + 8);
      if(chunked) {
        P.st_cur = s_state.as<int64_t>();
        P.st_grab = P.st_cur + nbra;
        P.st_cnt = reinterpret_cast<uint32_t*>(P.st_grab + 2 * nwaves);
        P.st_fil = P.st_cnt + nbra;
      }
      MHIP_HIP_CHECK(hipMemsetAsync(s_ctl.p, 0, 16, st));
    }
  }
  auto count_pass = [&](bool stage) {
    if(ex.product)
      product_count(ctx, *ex.product, nbra, counts32.as<uint32_t>());
    else if(ex.bproduct)
      batch_product_count(ctx, *ex.bproduct, ex.row_inst, nbra, counts32.as<uint32_t>());
    else if(grp)
      spin_group_count(ctx, W, *ex.groups, d_bra, nbra, counts32.as<uint32_t>());
    else if(stage) {
      const int64_t step = chunked ? ket_chunk : nket;
      for(int64_t c0 = 0; c0 < nket || c0 == 0; c0 += step) {
        const int64_t c1 = std::min(nket, c0 + step);
        const int flags = (c0 == 0 ? 1 : 0) | (c1 >= nket ? 2 : 0);
        hipLaunchKernelGGL((scan_pass_kernel<W, kRows, kStage>), dim3(grid), dim3(256), 0, st,
                           d_bra, nbra, d_ket, nket, tand, tor, ntiles, sand, sor, nsuper,
                           nwaves, nblocks, ex.items, counts32.as<uint32_t>(), nullptr,
                           nullptr, P, c0, c1, flags);
        if(c1 >= nket) break;
      }
    } else
      hipLaunchKernelGGL((scan_pass_kernel<W, kRows, kCount>), dim3(grid), dim3(256), 0, st,
                         d_bra, nbra, d_ket, nket, tand, tor, ntiles, sand, sor, nsuper,
                         nwaves, nblocks, ex.items, counts32.as<uint32_t>(), nullptr, nullptr,
                         StagePool{}, (int64_t)0, nket, 3);
    MHIP_HIP_CHECK(hipGetLastError());
  };
  t.start();
  count_pass(staged);
  t.stop();
  tm.ham_count_ms = t.ms();

  // 3. scan
  DevBuf rp_struct = take_buffer(ctx, sizeof(int64_t) * (nbra + 1));
  int64_t nnz_struct = 0;
  auto scan_counts = [&] {
    hipLaunchKernelGGL(widen_counts_kernel, dim3((nbra + 1 + 255) / 256), dim3(256), 0, st,
                       counts32.as<uint32_t>(), nbra, counts64.as<int64_t>());
    exclusive_scan(ctx, counts64.as<int64_t>(), rp_struct.as<int64_t>(), nbra + 1);
    d2h(ctx, &nnz_struct, rp_struct.as<int64_t>() + nbra, 8);
  };
  t.start();
  scan_counts();
  if(staged) {
    int overflow = 0;
    d2h(ctx, &overflow, P.overflow, sizeof(int));
    if(overflow) {  // the pool was too small: count + fill after all
      staged = false;
      s_pool.release();
      s_next.release();
      s_used.release();
      s_head.release();
      s_state.release();
      count_pass(false);
      scan_counts();
    }
  }
  t.stop();
  tm.ham_scan_ms = t.ms();  // synchronises
  tm.ham_nnz_struct = nnz_struct;
  counts32.release();

  // 4. fill (the product path evaluates the elements in the same pass; a
  // staged scan only gathers its chunks)
  DevBuf col = take_buffer(ctx, sizeof(int32_t) * std::max<int64_t>(nnz_struct, 1));
  if(staged) {
    t.start();
    if(nnz_struct)
      hipLaunchKernelGGL(stage_gather_kernel, dim3((nbra + kWavesPerBlock - 1) / kWavesPerBlock),
                         dim3(256), 0, st, P.head, P.chunk_next, P.used, P.col,
                         rp_struct.as<int64_t>(), nbra, col.as<int32_t>());
    MHIP_HIP_CHECK(hipGetLastError());
    MHIP_HIP_CHECK(hipStreamSynchronize(st));
    s_pool.release();
    s_next.release();
    s_used.release();
    s_head.release();
    s_state.release();
    t.stop();
  }
  DevBuf val = take_buffer(ctx, sizeof(double) * std::max<int64_t>(nnz_struct, 1));
  DevBuf diag(sizeof(double) * nbra);
  DevBuf flag(sizeof(int));
  MHIP_HIP_CHECK(hipMemsetAsync(flag.p, 0, sizeof(int), st));
  if(!staged) t.start();
  if(prod) {
    hipLaunchKernelGGL(diag_kernel<W>, dim3((nbra + 255) / 256), dim3(256), 0, st, I, d_bra,
                       nbra, diag.as<double>(), ex.views, ex.row_inst);
    if(nnz_struct && ex.bproduct)
      batch_product_fill_elements(ctx, *ex.bproduct, ex.row_inst, nbra, diag.as<double>(),
                                  rp_struct.as<int64_t>(), col.as<int32_t>(), val.as<double>(),
                                  thresh, counts64.as<int64_t>(), flag.as<int>());
    else if(nnz_struct)
      product_fill_elements(ctx, W, *ex.product, I, nbra, diag.as<double>(),
                            rp_struct.as<int64_t>(), col.as<int32_t>(), val.as<double>(), thresh,
                            counts64.as<int64_t>(), flag.as<int>());
  } else if(nnz_struct && grp) {
    // scan-order columns, then every row sorted ascending
    spin_group_fill(ctx, W, *ex.groups, d_bra, nbra, rp_struct.as<int64_t>(), col.as<int32_t>());
    sort_row_columns(ctx, nbra, rp_struct.as<int64_t>(), nnz_struct, nket, col);
  } else if(nnz_struct && !staged) {
    hipLaunchKernelGGL((scan_pass_kernel<W, kRows, kFill>), dim3(grid), dim3(256), 0, st,
                       d_bra, nbra, d_ket, nket, tand, tor, ntiles, sand, sor, nsuper,
                       nwaves, nblocks, ex.items, nullptr, rp_struct.as<int64_t>(),
                       col.as<int32_t>(), StagePool{}, (int64_t)0, nket, 3);
  }
  MHIP_HIP_CHECK(hipGetLastError());
  if(!staged) t.stop();
  tm.ham_fill_ms = t.ms();
  bnd.release();

  // 5. elements (tiled path)
  t.start();
  if(!prod) {
    hipLaunchKernelGGL(diag_kernel<W>, dim3((nbra + 255) / 256), dim3(256), 0, st, I, d_bra,
                       nbra, diag.as<double>(), ex.views, ex.row_inst);
    if(nnz_struct)
      hipLaunchKernelGGL(element_kernel<W>, dim3((nbra + 3) / 4), dim3(256), 0, st, I, d_bra,
                         nbra, d_ket, diag.as<double>(), rp_struct.as<int64_t>(),
                         col.as<int32_t>(), val.as<double>(), thresh,
                         counts64.as<int64_t>(), flag.as<int>(), ex.views, ex.row_inst);
  }
  MHIP_HIP_CHECK(hipGetLastError());
  int dropped = 0;
  d2h(ctx, &dropped, flag.p, sizeof(int));
  t.stop();
  tm.ham_elem_ms = t.ms();

  // 6. compaction (only when the threshold removed something)
  tm.ham_compact_ms = 0.;
  if(!dropped) {
    csr->rowptr = std::move(rp_struct);
    csr->colind = std::move(col);
    csr->nzval = std::move(val);
    csr->nnz = nnz_struct;
  } else {
    t.start();
    MHIP_HIP_CHECK(hipMemsetAsync(counts64.as<int64_t>() + nbra, 0, 8, st));
    DevBuf rp_new(sizeof(int64_t) * (nbra + 1));
    exclusive_scan(ctx, counts64.as<int64_t>(), rp_new.as<int64_t>(), nbra + 1);
    int64_t nnz = 0;
    d2h(ctx, &nnz, rp_new.as<int64_t>() + nbra, 8);
    DevBuf col2(sizeof(int32_t) * std::max<int64_t>(nnz, 1));
    DevBuf val2(sizeof(double) * std::max<int64_t>(nnz, 1));
    hipLaunchKernelGGL(compact_kernel, dim3((nbra + 3) / 4), dim3(256), 0, st, nbra,
                       rp_struct.as<int64_t>(), col.as<int32_t>(), val.as<double>(),
                       rp_new.as<int64_t>(), col2.as<int32_t>(), val2.as<double>(),
                       thresh);
    MHIP_HIP_CHECK(hipGetLastError());
    t.stop();
    tm.ham_compact_ms = t.ms();
    csr->rowptr = std::move(rp_new);
    csr->colind = std::move(col2);
    csr->nzval = std::move(val2);
    csr->nnz = nnz;
  }
  t_all.stop();
  tm.ham_total_ms = t_all.ms() + ex.detect_ms;
  tm.ham_nnz = csr->nnz;
  tm.ham_path = prod ? MHIP_BUILD_PRODUCT : grp ? MHIP_BUILD_SPIN_GROUPS : MHIP_BUILD_TILED;
}

template void build_impl<1>(mhip_context*, const uint64_t*, int64_t, const uint64_t*, int64_t,
                            double, mhip_csr*, const BuildExtra&);
template void build_impl<2>(mhip_context*, const uint64_t*, int64_t, const uint64_t*, int64_t,
                            double, mhip_csr*, const BuildExtra&);

namespace {
void check_dets(int W, const uint64_t* p, int64_t n, const char* what) {
  MHIP_REQUIRE(W == 1 || W == 2, MHIP_ERR_UNSUPPORTED,
               "words_per_det must be 1 (wfn_t<64>) or 2 (wfn_t<128>)");
  MHIP_REQUIRE(n >= 0, MHIP_ERR_INVALID_ARG, std::string(what) + ": negative count");
  MHIP_REQUIRE(n == 0 || p, MHIP_ERR_INVALID_ARG, std::string(what) + ": null pointer");
  MHIP_REQUIRE(((uintptr_t)p % (8 * W)) == 0, MHIP_ERR_INVALID_ARG,
               std::string(what) + ": determinant array must be 8*W-byte aligned");
}

}  // namespace

void validate_dets(mhip_context* ctx, int W, const uint64_t* d, int64_t n);

mhip_csr* ham_build(mhip_context* ctx, int W, const uint64_t* d_bra,
                    int64_t nbra, const uint64_t* d_ket, int64_t nket,
                    double thresh) {
  use_device(ctx);
  const auto t_entry = std::chrono::steady_clock::now();
  check_dets(W, d_bra, nbra, "bra");
  check_dets(W, d_ket, nket, "ket");
  MHIP_REQUIRE(nket < (1ll << 31), MHIP_ERR_INDEX_OVERFLOW,
               "ket count exceeds the int32 column index range");
  std::unique_ptr<mhip_csr> csr(new mhip_csr());
  csr->device = ctx->device;
  csr->m = nbra;
  csr->n = nket;
  csr->indexing = 0;
  if(nbra == 0 || nket == 0) {  // csr_hamiltonian.hpp:30-35
    csr->rowptr.alloc(sizeof(int64_t) * (nbra + 1));
    MHIP_HIP_CHECK(hipMemsetAsync(csr->rowptr.p, 0, sizeof(int64_t) * (nbra + 1), ctx->stream));
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    csr->nnz = 0;
    ctx->timings.ham_nnz = ctx->timings.ham_nnz_struct = 0;
    return csr.release();
  }
  MHIP_REQUIRE(ctx->ints.norb > 0, MHIP_ERR_NO_INTEGRALS,
               "mhip_set_integrals must be called before building H");
  MHIP_REQUIRE(ctx->ints.norb <= 32 * W, MHIP_ERR_INVALID_ARG,
               "norb exceeds the spin-string width of the determinant type");
  validate_dets(ctx, W, d_bra, nbra);
  const bool square = d_ket == d_bra && nket == nbra;
  if(!square) validate_dets(ctx, W, d_ket, nket);
  // Product-space structure passes for complete spaces, spin-string group
  // scan for other long lists (same CSR).
  ProductSpace ps;
  SpinGroups sg;
  BuildExtra ex;
  const int path = ctx->build_path;
  if(square && (path == MHIP_BUILD_AUTO || path == MHIP_BUILD_PRODUCT)) {
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    const auto t0 = std::chrono::steady_clock::now();
    if(detect_product_space(ctx, W, d_bra, nbra, ps)) ex.product = &ps;
    ex.detect_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    MHIP_REQUIRE(ex.product || path != MHIP_BUILD_PRODUCT, MHIP_ERR_INVALID_ARG,
                 "determinant list is not a product space");
  } else {
    MHIP_REQUIRE(path != MHIP_BUILD_PRODUCT, MHIP_ERR_INVALID_ARG,
                 "the product-space build needs a square block (bra list == ket list)");
  }
  if(!ex.product && square &&
     (path == MHIP_BUILD_SPIN_GROUPS || (path == MHIP_BUILD_AUTO && nbra >= kMinGroupRows))) {
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    const auto t0 = std::chrono::steady_clock::now();
    // flat lists within a quarter of the free memory (the CSR comes after)
    size_t free_b = 0, total_b = 0;
    device_mem_info(&free_b, &total_b);
    // out of memory while preparing: the tiled loop needs far less
    try {
      if(prepare_spin_groups(ctx, W, d_bra, nbra, 0.25 * (double)free_b, sg)) ex.groups = &sg;
    } catch(const Error& e) {
      if(e.code != MHIP_ERR_OUT_OF_MEMORY) throw;
      (void)hipGetLastError();  // a refused hipMalloc leaves no pending error
      sg = SpinGroups();
    }
    ex.detect_ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    MHIP_REQUIRE(ex.groups || path != MHIP_BUILD_SPIN_GROUPS, MHIP_ERR_INVALID_ARG,
                 "spin-group build: electron counts vary over the list (or its lists exceed "
                 "the memory budget)");
  }
  MHIP_REQUIRE(square || path != MHIP_BUILD_SPIN_GROUPS, MHIP_ERR_INVALID_ARG,
               "the spin-group build needs a square block (bra list == ket list)");
  if(W == 1)
    build_impl<1>(ctx, d_bra, nbra, d_ket, nket, thresh, csr.get(), ex);
  else
    build_impl<2>(ctx, d_bra, nbra, d_ket, nket, thresh, csr.get(), ex);
  // end-to-end host wall time of the call (validation, recognition, all
  // passes; build_impl has synchronised)
  ctx->timings.ham_total_ms = std::chrono::duration<double, std::milli>(
                                  std::chrono::steady_clock::now() - t_entry)
                                  .count();
  return csr.release();
}

void ham_diagonal(mhip_context* ctx, int W, const uint64_t* d_dets,
                  int64_t ndets, double* d_diag) {
  use_device(ctx);
  check_dets(W, d_dets, ndets, "dets");
  MHIP_REQUIRE(ctx->ints.norb > 0, MHIP_ERR_NO_INTEGRALS, "no integrals set");
  if(!ndets) return;
  validate_dets(ctx, W, d_dets, ndets);
  const IntsView I = view(ctx->ints);
  if(W == 1)
    hipLaunchKernelGGL(diag_kernel<1>, dim3((ndets + 255) / 256), dim3(256), 0, ctx->stream,
                       I, d_dets, ndets, d_diag, nullptr, nullptr);
  else
    hipLaunchKernelGGL(diag_kernel<2>, dim3((ndets + 255) / 256), dim3(256), 0, ctx->stream,
                       I, d_dets, ndets, d_diag, nullptr, nullptr);
  MHIP_HIP_CHECK(hipGetLastError());
}

IntsView device_view(const Integrals& I) { return view(I); }

// Rejects determinants with occupied orbitals >= norb (they would index
// the integral tables out of range).
void validate_dets(mhip_context* ctx, int W, const uint64_t* d, int64_t n) {
  if(!n) return;
  DevBuf o(16);
  MHIP_HIP_CHECK(hipMemsetAsync(o.p, 0, 16, ctx->stream));
  const int64_t blocks = std::min<int64_t>(1024, (n + 255) / 256);
  if(W == 1)
    hipLaunchKernelGGL(or_all_kernel<1>, dim3(blocks), dim3(256), 0, ctx->stream, d, n,
                       o.as<unsigned long long>());
  else
    hipLaunchKernelGGL(or_all_kernel<2>, dim3(blocks), dim3(256), 0, ctx->stream, d, n,
                       o.as<unsigned long long>());
  MHIP_HIP_CHECK(hipGetLastError());
  uint64_t h[2] = {0, 0};
  d2h(ctx, h, o.p, 8 * W);
  const int64_t norb = ctx->ints.norb;
  uint64_t a, b;
  if(W == 1) {
    a = h[0] & 0xffffffffull;
    b = h[0] >> 32;
  } else {
    a = h[0];
    b = h[1];
  }
  const uint64_t bad = norb >= 64 ? 0ull : ~((1ull << norb) - 1ull);
  MHIP_REQUIRE(!((a | b) & bad), MHIP_ERR_INVALID_ARG,
               "determinant occupies an orbital index >= norb");
}

}  // namespace mhip

using namespace mhip;

extern "C" {

int mhip_ham_build(mhip_context* ctx, int words_per_det, const uint64_t* d_bra,
                   int64_t nbra, const uint64_t* d_ket, int64_t nket,
                   double thresh, mhip_csr** out) {
  return guarded([&] {
    MHIP_REQUIRE(out, MHIP_ERR_INVALID_ARG, "null output pointer");
    *out = nullptr;
    *out = ham_build(ctx, words_per_det, d_bra, nbra, d_ket, nket, thresh);
  });
}

int mhip_ham_diagonal(mhip_context* ctx, int words_per_det,
                      const uint64_t* d_dets, int64_t ndets, double* d_diag) {
  return guarded([&] {
    MHIP_REQUIRE(ndets == 0 || d_diag, MHIP_ERR_INVALID_ARG, "null output");
    ham_diagonal(ctx, words_per_det, d_dets, ndets, d_diag);
  });
}

}  // extern "C"
