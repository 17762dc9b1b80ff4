// K2: FP64 CSR SpMV (gespmbv, spmbv.hpp:37-73), diagonal extraction
// (extract_diagonal_elements, submatrix.hpp:264-295) and the CSR handle
// entry points.
//
// Two layouts. CSR (the handle's own arrays): four rows per wavefront, 16
// lanes per row streaming 16-byte quads of column indices and values with
// non-temporal loads (read once, kept out of L2 so the gathered x vector
// stays cached); x is gathered through L1/L2 and each row sum is reduced
// across its 16 lanes. Sliced copy (SELL-64, made once per matrix by
// prepare_sell when the device has room, then used by every single-vector
// SpMV of that matrix): one 64-row slice per wavefront, lane = row, slots
// streamed slot-major. Both are HBM bound on the streamed 12 bytes per
// nonzero; the x gathers' L1/L2 requests are the second limiter, far less
// so in the sliced layout.
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <limits>
#include <memory>
#include <vector>

#include "mhip_device.hpp"
#include "mhip_internal.hpp"

namespace mhip {
namespace {

constexpr int kWavesPerBlock = 4;
constexpr int kXcds = 8;
// SpMV: 4 rows per wavefront, 16 lanes per row.
constexpr int kRowsPerWave = 4;
constexpr int kLanesPerRow = 64 / kRowsPerWave;
constexpr int kRowsPerBlock = kWavesPerBlock * kRowsPerWave;

typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// Four rows share a wavefront, 16 lanes each, and every lane streams whole
// 16-byte aligned quads of its row (one column load, two value loads,
// non-temporal); the up-to-3 elements before the first aligned quad and
// after the last one are taken by lanes 0-5 of the row's group. Neighbouring
// rows of a determinant-ordered Hamiltonian gather nearly the same x lines
// at the same nonzero position, so the four rows' gathers coalesce in one
// instruction. Blocks are mapped XCD-aware: hardware block b runs on XCD
// b % 8, and XCD x owns logical blocks [x*per, (x+1)*per), a contiguous
// range of rows whose x working set then stays in that XCD's L2.
__global__ __launch_bounds__(256) void spmv_kernel(
    int64_t m, int64_t per_xcd, const int64_t* __restrict__ rowptr,
    const int32_t* __restrict__ colind, const double* __restrict__ nzval,
    int64_t indexing, double alpha, const double* __restrict__ x, double beta,
    double* __restrict__ y) {
  const int64_t lblk = (int64_t)(blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  const int64_t row = lblk * kRowsPerBlock + threadIdx.x / kLanesPerRow;
  const int l = threadIdx.x % kLanesPerRow;
  const double* xb = x - indexing;
  double acc0 = 0., acc1 = 0., acc2 = 0., acc3 = 0.;
  if(row < m) {
    const int64_t st = rowptr[row] - indexing, en = rowptr[row + 1] - indexing;
    const int64_t qa = (st + 3) >> 2, qb = en >> 2;  // whole quads [qa, qb)
    const i32x4* cq = reinterpret_cast<const i32x4*>(colind);
    const f64x2* vq = reinterpret_cast<const f64x2*>(nzval);
    int64_t q = qa + l;
    for(; q + kLanesPerRow < qb; q += 2 * kLanesPerRow) {
      const i32x4 c0 = __builtin_nontemporal_load(cq + q);
      const i32x4 c1 = __builtin_nontemporal_load(cq + q + kLanesPerRow);
      const f64x2 v0 = __builtin_nontemporal_load(vq + 2 * q);
      const f64x2 v1 = __builtin_nontemporal_load(vq + 2 * q + 1);
      const f64x2 v2 = __builtin_nontemporal_load(vq + 2 * (q + kLanesPerRow));
      const f64x2 v3 = __builtin_nontemporal_load(vq + 2 * (q + kLanesPerRow) + 1);
      acc0 = fma(v0.x, xb[c0.x], acc0);
      acc1 = fma(v0.y, xb[c0.y], acc1);
      acc2 = fma(v1.x, xb[c0.z], acc2);
      acc3 = fma(v1.y, xb[c0.w], acc3);
      acc0 = fma(v2.x, xb[c1.x], acc0);
      acc1 = fma(v2.y, xb[c1.y], acc1);
      acc2 = fma(v3.x, xb[c1.z], acc2);
      acc3 = fma(v3.y, xb[c1.w], acc3);
    }
    if(q < qb) {
      const i32x4 c0 = __builtin_nontemporal_load(cq + q);
      const f64x2 v0 = __builtin_nontemporal_load(vq + 2 * q);
      const f64x2 v1 = __builtin_nontemporal_load(vq + 2 * q + 1);
      acc0 = fma(v0.x, xb[c0.x], acc0);
      acc1 = fma(v0.y, xb[c0.y], acc1);
      acc2 = fma(v1.x, xb[c0.z], acc2);
      acc3 = fma(v1.y, xb[c0.w], acc3);
    }
    // head [st, min(en, 4 qa)) and tail [max(4 qb, 4 qa), en): at most 3 each,
    // after the quads so the streamed loads are issued first
    const int64_t he = en < 4 * qa ? en : 4 * qa;
    const int64_t ts = 4 * qb > 4 * qa ? 4 * qb : 4 * qa;
    int64_t j = -1;
    if(l < 3) j = st + l < he ? st + l : -1;
    else if(l < 6) j = ts + (l - 3) < en ? ts + (l - 3) : -1;
    if(j >= 0) acc0 = fma(nzval[j], xb[colind[j]], acc0);
  }
  double acc = (acc0 + acc1) + (acc2 + acc3);
#pragma unroll
  for(int off = kLanesPerRow / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if(row < m && l == 0) y[row] = alpha * acc + beta * y[row];  // spmbv.hpp:71
}

// ---- column panels ----------------------------------------------------
// The CSR kernel above gathers x from everywhere in every row; once x (8
// bytes per column) is larger than an XCD's 4 MB L2, most gathers miss and
// the same x lines are fetched again and again (the 2,000,000-determinant
// selected-CI matrix: 52.8 GB fetched per product for 33.6 GB of matrix).
// Here the columns are cut into nnz-balanced panels, a multiple of 8, and
// XCD x multiplies every row's entries of panels x, x + 8, ...: that XCD's
// L2 then holds one panel of x (a few MB at most) at a time while
// the matrix still streams through once in total. Each panel writes its own
// partial sums, added in a fixed order by panel_sum_kernel.
constexpr int kHistBins = 4096;

// Column histogram of every `stride`-th row into kHistBins equal bins.
__global__ __launch_bounds__(256) void col_hist_kernel(int64_t m, int64_t n, int64_t stride,
                                                       const int64_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ colind,
                                                       int64_t indexing,
                                                       unsigned long long* __restrict__ hist) {
  __shared__ unsigned int h[kHistBins];
  for(int i = threadIdx.x; i < kHistBins; i += 256) h[i] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t nrows = (m + stride - 1) / stride;
  for(int64_t s = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); s < nrows;
      s += (int64_t)gridDim.x * 4) {
    const int64_t row = s * stride;
    const int64_t st = rowptr[row] - indexing, en = rowptr[row + 1] - indexing;
    for(int64_t j = st + lane; j < en; j += 64) {
      const int64_t c = colind[j] - indexing;
      atomicAdd(&h[(int)((c * kHistBins) / n)], 1u);
    }
  }
  __syncthreads();
  for(int i = threadIdx.x; i < kHistBins; i += 256)
    if(h[i]) atomicAdd(&hist[i], (unsigned long long)h[i]);
}

// off[row * np + p] = first entry of row `row` with column >= bounds[p]
// (relative to the row start), p < np; bounds[0] = 0.
struct PanelBounds {
  int32_t b[kMaxPanels];
};
__global__ __launch_bounds__(256) void panel_offsets_kernel(int64_t m,
                                                            const int64_t* __restrict__ rowptr,
                                                            const int32_t* __restrict__ colind,
                                                            int64_t indexing, int np,
                                                            PanelBounds pb,
                                                            int32_t* __restrict__ off) {
  const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if(row >= m) return;
  const int64_t st = rowptr[row] - indexing, en = rowptr[row + 1] - indexing;
  off[row * np] = 0;
  int64_t lo = st;
  for(int p = 1; p < np; ++p) {
    const int32_t c = pb.b[p] + (int32_t)indexing;
    int64_t a = lo, z = en;  // first j in [lo, en) with colind[j] >= c
    while(a < z) {
      const int64_t mid = (a + z) >> 1;
      if(colind[mid] < c) a = mid + 1;
      else z = mid;
    }
    off[row * np + p] = (int32_t)(a - st);
    lo = a;
  }
}

// part[p * m + row] = sum over row's entries in panel p of A(row, c) x(c).
// Block b runs on XCD b % 8 and takes panel (b % 8) + 8 * round, round =
// (b / 8) / nblk, for row block (b / 8) % nblk: XCD x multiplies panels x,
// x + 8, ... one after the other. The row loop is spmv_kernel's (16 lanes
// per row, 16-byte quads, non-temporal matrix loads) over the panel's run.
__global__ __launch_bounds__(256) void spmv_panel_kernel(
    int64_t m, int64_t nblk, int np, const int64_t* __restrict__ rowptr,
    const int32_t* __restrict__ colind, const double* __restrict__ nzval, int64_t indexing,
    const int32_t* __restrict__ off, const double* __restrict__ x, double* __restrict__ part) {
  const int64_t i = blockIdx.x / kXcds;
  const int p = (int)(blockIdx.x % kXcds) + kXcds * (int)(i / nblk);
  const int64_t row = (i % nblk) * kRowsPerBlock + threadIdx.x / kLanesPerRow;
  const int l = threadIdx.x % kLanesPerRow;
  const double* xb = x - indexing;
  double acc0 = 0., acc1 = 0., acc2 = 0., acc3 = 0.;
  if(row < m) {
    const int64_t rs = rowptr[row] - indexing;
    const int64_t st = rs + off[row * np + p];
    const int64_t en = p + 1 < np ? rs + off[row * np + p + 1] : rowptr[row + 1] - indexing;
    const int64_t qa = (st + 3) >> 2, qb = en >> 2;  // whole quads [qa, qb)
    const i32x4* cq = reinterpret_cast<const i32x4*>(colind);
    const f64x2* vq = reinterpret_cast<const f64x2*>(nzval);
    int64_t q = qa + l;
    for(; q + kLanesPerRow < qb; q += 2 * kLanesPerRow) {
      const i32x4 c0 = __builtin_nontemporal_load(cq + q);
      const i32x4 c1 = __builtin_nontemporal_load(cq + q + kLanesPerRow);
      const f64x2 v0 = __builtin_nontemporal_load(vq + 2 * q);
      const f64x2 v1 = __builtin_nontemporal_load(vq + 2 * q + 1);
      const f64x2 v2 = __builtin_nontemporal_load(vq + 2 * (q + kLanesPerRow));
      const f64x2 v3 = __builtin_nontemporal_load(vq + 2 * (q + kLanesPerRow) + 1);
      acc0 = fma(v0.x, xb[c0.x], acc0);
      acc1 = fma(v0.y, xb[c0.y], acc1);
      acc2 = fma(v1.x, xb[c0.z], acc2);
      acc3 = fma(v1.y, xb[c0.w], acc3);
      acc0 = fma(v2.x, xb[c1.x], acc0);
      acc1 = fma(v2.y, xb[c1.y], acc1);
      acc2 = fma(v3.x, xb[c1.z], acc2);
      acc3 = fma(v3.y, xb[c1.w], acc3);
    }
    if(q < qb) {
      const i32x4 c0 = __builtin_nontemporal_load(cq + q);
      const f64x2 v0 = __builtin_nontemporal_load(vq + 2 * q);
      const f64x2 v1 = __builtin_nontemporal_load(vq + 2 * q + 1);
      acc0 = fma(v0.x, xb[c0.x], acc0);
      acc1 = fma(v0.y, xb[c0.y], acc1);
      acc2 = fma(v1.x, xb[c0.z], acc2);
      acc3 = fma(v1.y, xb[c0.w], acc3);
    }
    const int64_t he = en < 4 * qa ? en : 4 * qa;
    const int64_t ts = 4 * qb > 4 * qa ? 4 * qb : 4 * qa;
    int64_t j = -1;
    if(l < 3) j = st + l < he ? st + l : -1;
    else if(l < 6) j = ts + (l - 3) < en ? ts + (l - 3) : -1;
    if(j >= 0) acc0 = fma(nzval[j], xb[colind[j]], acc0);
  }
  double acc = (acc0 + acc1) + (acc2 + acc3);
#pragma unroll
  for(int o = kLanesPerRow / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if(row < m && l == 0) part[(int64_t)p * m + row] = acc;
}

// y = alpha * (sum of the panels' partials in panel order) + beta * y
__global__ __launch_bounds__(256) void panel_sum_kernel(int64_t m, int np,
                                                        const double* __restrict__ part,
                                                        double alpha, double beta,
                                                        double* __restrict__ y) {
  for(int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < m;
      r += (int64_t)gridDim.x * 256) {
    double s = part[r];
    for(int p = 1; p < np; ++p) s += part[(int64_t)p * m + r];
    y[r] = alpha * s + beta * y[r];  // spmbv.hpp:71
  }
}

// ---- x tiles --------------------------------------------------------------
// The panel kernel above still gathers x through L1 and L2 one line per
// nonzero, and on the 2,000,000-determinant selected-CI matrix those L1
// misses, not the streamed matrix, set its time. Here a panel is narrow
// enough (kTileCols columns, 96 KB) that its share of x sits in LDS: the
// matrix is copied once panel-major (16-bit local columns), and a workgroup
// owns kTileRows consecutive rows, walks the panels in order, stages each
// panel's x in LDS and streams the panel's entries of its rows contiguously,
// kTilePiece at a time: every thread forms kTilePiece / 1024 products
// (gathers from LDS) into an LDS buffer, then a thread per row adds the
// products of its row's run in that piece, in order, into the row's sum
// kept in LDS. Row sums are written once, after the last panel; the
// summation order (panels, pieces, entries within a row's run) is fixed.
// Measured on the 2,000,000-determinant selected-CI matrix: 24.9 ms per
// product against 14.5 ms for the panel kernel (a first version with one
// thread per row's run and row offsets read from global memory: 75 ms).
// One 96 KB tile fills most of a CU's LDS, so the CU's waves synchronise
// on every 4,096-entry piece and on every panel's tile load; the layout is
// kept as an option (MHIP_SPMV_X_TILES), not the Davidson default.
constexpr int kTileRows = 2048;
constexpr int kTilePiece = 4096;
constexpr int kTileThreads = 1024;
constexpr int kTilePer = kTilePiece / kTileThreads;

// counts[p * m + r] += entries of row r in panel p (counts zeroed by the
// caller); one wavefront per row, one add per run of equal panels in each
// 64-entry window (columns ascend, so a panel's entries form one run).
__global__ __launch_bounds__(256) void xtile_count_kernel(int64_t m,
                                                          const int64_t* __restrict__ rowptr,
                                                          const int32_t* __restrict__ colind,
                                                          int64_t indexing,
                                                          uint32_t* __restrict__ counts) {
  const int64_t row = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if(row >= m) return;
  const int lane = threadIdx.x & 63;
  const int64_t st = rowptr[row] - indexing, en = rowptr[row + 1] - indexing;
  for(int64_t j0 = st; j0 < en; j0 += 64) {
    const int64_t j = j0 + lane;
    const bool v = j < en;
    const int p = v ? (int)((colind[j] - indexing) / kTileCols) : -1;
    // neighbours' panels, read with every lane active (a cross-lane read
    // from a lane masked off by a branch returns 0, not its value)
    const int pn = __shfl_down(p, 1, 64);
    const int pp = __shfl_up(p, 1, 64);
    // last lane of its run in this window
    const bool last = v && (lane == 63 || j + 1 >= en || pn != p);
    const bool first = v && (lane == 0 || pp != p);
    const uint64_t firsts = ballot(first);
    if(last) {
      const int f = 63 - __clzll((long long)(firsts & lowmask((uint32_t)lane + 1u)));
      atomicAdd(&counts[(int64_t)p * m + row], (uint32_t)(lane - f + 1));
    }
  }
}

// rowoff[p * (m + 1) + r] = scan[p * m + r] - base[p] (r < m), the panel's
// entry count at r = m; base[p] = scan[p * m] (base[np] = total).
__global__ __launch_bounds__(256) void xtile_rowoff_kernel(int64_t m, int np,
                                                           const int64_t* __restrict__ scan,
                                                           int64_t total,
                                                           uint32_t* __restrict__ rowoff,
                                                           int64_t* __restrict__ base) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t nn = (int64_t)np * (m + 1);
  if(i >= nn) return;
  const int p = (int)(i / (m + 1));
  const int64_t r = i % (m + 1);
  const int64_t b = scan[(int64_t)p * m];
  const int64_t e = p + 1 < np ? scan[(int64_t)(p + 1) * m] : total;
  rowoff[i] = (uint32_t)((r < m ? scan[(int64_t)p * m + r] : e) - b);
  if(r == 0) base[p] = b;
  if(i == 0) base[np] = total;
}

// Entries into the panel-major copy: entry j of row r, in panel p with
// rank k within the row's run of p, goes to base[p] + rowoff[p][r] + k.
__global__ __launch_bounds__(256) void xtile_scatter_kernel(
    int64_t m, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
    const double* __restrict__ nzval, int64_t indexing, const int64_t* __restrict__ base,
    const uint32_t* __restrict__ rowoff, uint16_t* __restrict__ cols,
    double* __restrict__ vals) {
  const int64_t row = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if(row >= m) return;
  const int lane = threadIdx.x & 63;
  const int64_t st = rowptr[row] - indexing, en = rowptr[row + 1] - indexing;
  int prev_p = -1;       // panel of the previous window's last entry
  int64_t prev_s = st;   // first entry of that panel's run
  for(int64_t j0 = st; j0 < en; j0 += 64) {
    const int64_t j = j0 + lane;
    const bool v = j < en;
    const int64_t c = v ? colind[j] - indexing : 0;
    const int p = v ? (int)(c / kTileCols) : -1;
    const int pp = __shfl_up(p, 1, 64);  // every lane active (see xtile_count_kernel)
    const int pu = lane == 0 ? prev_p : pp;
    const bool first = v && pu != p;
    const uint64_t firsts = ballot(first);
    const uint64_t below = firsts & lowmask((uint32_t)lane + 1u);
    // the run's first entry: the last run start at or below this lane, or
    // (no start in this window yet) the carried start of the previous window
    const int64_t s = below ? j0 + (63 - __clzll((long long)below)) : prev_s;
    if(v) {
      const int64_t dst = base[p] + rowoff[(int64_t)p * (m + 1) + row] + (j - s);
      cols[dst] = (uint16_t)(c - (int64_t)p * kTileCols);
      vals[dst] = nzval[j];
    }
    // carry the last valid lane's panel and run start to the next window
    const uint64_t vm = ballot(v);
    const int lastlane = 63 - __clzll((long long)vm);
    prev_p = __shfl(p, lastlane, 64);
    prev_s = __shfl(s, lastlane, 64);
  }
}

// First row r in [lo_r, hi_r) with key[r] > v (key ascending; hi_r if none).
__device__ __forceinline__ int first_above(const uint32_t* key, int lo_r, int hi_r, uint32_t v) {
  while(lo_r < hi_r) {
    const int mid = (lo_r + hi_r) >> 1;
    if(key[mid] > v)
      hi_r = mid;
    else
      lo_r = mid + 1;
  }
  return lo_r;
}

// y = alpha A x + beta y on the panel-major copy (block = kTileRows rows).
// Per panel: x's share and the rows' offsets into LDS; per piece: every
// thread forms its kTilePer products (16-bit columns index the LDS x tile)
// into an LDS buffer, then loads the next piece's entries into registers
// while 16 lanes per row add the row's run in the piece (lanes strided,
// then a 16-lane tree) into the row's sum in LDS.
__global__ __launch_bounds__(kTileThreads) void spmv_xtile_kernel(
    int64_t m, int64_t n, int np, const int64_t* __restrict__ base,
    const uint32_t* __restrict__ rowoff, const uint16_t* __restrict__ cols,
    const double* __restrict__ vals, double alpha, const double* __restrict__ x, double beta,
    double* __restrict__ y) {
  __shared__ double xs[kTileCols];
  __shared__ double prod[kTilePiece];
  __shared__ double acc[kTileRows];
  __shared__ uint32_t rof[kTileRows + 1];
  const int t = threadIdx.x;
  const int sub = t & 15;  // lane within the row's 16
  const int64_t r0 = (int64_t)blockIdx.x * kTileRows;
  const int nr = (int)(m - r0 < kTileRows ? m - r0 : kTileRows);
  for(int i = t; i < kTileRows; i += kTileThreads) acc[i] = 0.;
  for(int p = 0; p < np; ++p) {
    const uint32_t* __restrict__ ro = rowoff + (int64_t)p * (m + 1) + r0;
    const uint32_t e0 = ro[0], e1 = ro[nr];
    if(e0 == e1) continue;  // uniform: no entry of these rows in panel p
    const int64_t pb = base[p];
    const int64_t c0 = (int64_t)p * kTileCols;
    const int ncol = (int)(n - c0 < kTileCols ? n - c0 : kTileCols);
    __syncthreads();  // the previous panel's reads of xs, rof and prod are done
    for(int i = t; i < ncol; i += kTileThreads) xs[i] = x[c0 + i];
    for(int i = t; i <= nr; i += kTileThreads) rof[i] = ro[i];
    // the first piece's entries
    uint16_t cv[kTilePer];
    double vv[kTilePer];
    auto load_piece = [&](uint32_t lo, uint32_t hi) {
#pragma unroll
      for(int k = 0; k < kTilePer; ++k) {
        const uint32_t q = (uint32_t)(k * kTileThreads + t);
        const bool in = lo + q < hi;
        const int64_t g = pb + lo + (in ? q : 0u);
        cv[k] = in ? __builtin_nontemporal_load(cols + g) : (uint16_t)0;
        vv[k] = in ? __builtin_nontemporal_load(vals + g) : 0.;
      }
    };
    uint32_t lo = e0;
    uint32_t hi = e1 - lo < (uint32_t)kTilePiece ? e1 : lo + kTilePiece;
    load_piece(lo, hi);
    __syncthreads();
    int ra = 0;  // first row whose run ends after lo (uniform)
    while(lo < e1) {
#pragma unroll
      for(int k = 0; k < kTilePer; ++k) {
        const uint32_t q = (uint32_t)(k * kTileThreads + t);
        if(lo + q < hi) prod[q] = vv[k] * xs[cv[k]];
      }
      __syncthreads();
      const uint32_t nlo = hi;
      const uint32_t nhi = e1 - nlo < (uint32_t)kTilePiece ? e1 : nlo + kTilePiece;
      if(nlo < e1) load_piece(nlo, nhi);  // in flight during the row sums
      ra = first_above(rof + 1, ra, nr, lo);      // rof[r + 1] > lo
      const int rb = first_above(rof, ra, nr, hi - 1);  // rof[r] >= hi
      for(int r = ra + (t >> 4); r < rb; r += kTileThreads / 16) {
        const uint32_t a = rof[r] > lo ? rof[r] : lo;
        const uint32_t b = rof[r + 1] < hi ? rof[r + 1] : hi;
        double part = 0.;
        for(uint32_t i = a + sub; i < b; i += 16) part += prod[i - lo];
#pragma unroll
        for(int o = 8; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
        if(sub == 0) acc[r] += part;
      }
      __syncthreads();
      lo = nlo;
      hi = nhi;
    }
  }
  __syncthreads();
  for(int i = t; i < nr; i += kTileThreads) y[r0 + i] = alpha * acc[i] + beta * y[r0 + i];
}

__global__ __launch_bounds__(256) void extract_diag_kernel(
    int64_t m, const int64_t* __restrict__ rowptr,
    const int32_t* __restrict__ colind, const double* __restrict__ nzval,
    int64_t indexing, double* __restrict__ D) {
  const int64_t row = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if(row >= m) return;
  const int lane = threadIdx.x & 63;
  const int64_t st = rowptr[row] - indexing, en = rowptr[row + 1] - indexing;
  double d = 0.;
  for(int64_t j0 = st; j0 < en; j0 += 64) {
    const int64_t j = j0 + lane;
    const bool hit = j < en && (int64_t)colind[j] == row + indexing;
    const uint64_t mask = ballot(hit);
    if(mask) {  // first occurrence in row order (std::find)
      const int first = (int)ctz64(mask);
      const double v = __shfl(hit ? nzval[j] : 0., first, 64);
      d = v;
      break;
    }
  }
  if(lane == 0) D[row] = d;
}


// ---- sliced (SELL-64) copy ----------------------------------------------
// One wavefront owns one 64-row slice and lane l one row: every load of a
// slot is one contiguous 256-byte (columns) / 512-byte (values) access. For
// product-space matrices all rows have one length (no padding) and the
// lanes' gathers at one slot fall in one x segment (rows of one outer string
// share their outer neighbours slot by slot), so x lines are reused across
// the wavefront. Selected-CI matrices have ragged rows and scattered
// columns: there the CSR kernel's rows-along-lanes gathers are faster even
// without padding, and the padding rule of prepare_sell keeps them on it.
constexpr int kSliceRows = 64;
constexpr int kSellChunk = 32;  // slots staged through LDS per round of the copy
constexpr int kSellUnroll = 8;  // slots in flight per lane

// slen[s] = longest row of slice s (one wavefront per slice).
__global__ __launch_bounds__(256) void sell_lengths_kernel(int64_t m, int64_t nsl,
                                                           const int64_t* __restrict__ rowptr,
                                                           int32_t* __restrict__ slen) {
  const int64_t s = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / 64;
  if(s >= nsl) return;
  const int64_t row = s * kSliceRows + (threadIdx.x & 63);
  int len = row < m ? (int)(rowptr[row + 1] - rowptr[row]) : 0;
#pragma unroll
  for(int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(len, off, 64);
    len = o > len ? o : len;
  }
  if((threadIdx.x & 63) == 0) slen[s] = len;
}

// soff = exclusive scan of 64 * slen, total entries in *total (one block).
__global__ __launch_bounds__(1024) void sell_offsets_kernel(int64_t nsl,
                                                            const int32_t* __restrict__ slen,
                                                            int64_t* __restrict__ soff,
                                                            int64_t* __restrict__ total) {
  __shared__ int64_t part[1024];
  const int t = threadIdx.x;
  const int64_t chunk = (nsl + 1023) / 1024;
  const int64_t lo = std::min<int64_t>(nsl, t * chunk), hi = std::min<int64_t>(nsl, lo + chunk);
  int64_t sum = 0;
  for(int64_t i = lo; i < hi; ++i) sum += (int64_t)slen[i] * kSliceRows;
  part[t] = sum;
  __syncthreads();
  for(int off = 1; off < 1024; off <<= 1) {
    const int64_t v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int64_t run = part[t] - sum;
  for(int64_t i = lo; i < hi; ++i) {
    soff[i] = run;
    run += (int64_t)slen[i] * kSliceRows;
  }
  if(t == 1023) *total = part[1023];
}

// *total = entries of the narrow slices (one block).
__global__ __launch_bounds__(1024) void sell_narrow_count_kernel(int64_t nsl,
                                                                 const int32_t* __restrict__ slen,
                                                                 const int32_t* __restrict__ narrow,
                                                                 int64_t* __restrict__ total) {
  __shared__ int64_t part[1024];
  const int t = threadIdx.x;
  int64_t sum = 0;
  for(int64_t i = t; i < nsl; i += 1024) sum += narrow[i] ? (int64_t)slen[i] * kSliceRows : 0;
  part[t] = sum;
  __syncthreads();
  for(int off = 512; off > 0; off >>= 1) {
    if(t < off) part[t] += part[t + off];
    __syncthreads();
  }
  if(t == 0) *total = part[0];
}

// Copy of slice blockIdx.x, kSellChunk slots at a time: rows are read
// contiguously into an LDS tile and written back slot-major. A first pass
// over the slice's columns finds every slot's smallest and largest column
// (padding excluded), stores the smallest as the slot base and decides the
// slice's width: narrow when every slot spans less than 65,536 columns.
// Padding entries get value 0 and, narrow, offset 0 (the base column),
// wide, the row itself as column (past the last row, column 0).
__global__ __launch_bounds__(256) void sell_fill_kernel(
    int64_t m, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
    const double* __restrict__ nzval, int64_t indexing, const int64_t* __restrict__ soff,
    const int32_t* __restrict__ slen, int32_t* __restrict__ cols, double* __restrict__ vals,
    int32_t* __restrict__ narrow, int32_t* __restrict__ sbase) {
  __shared__ int64_t rst[kSliceRows];
  __shared__ int32_t rlen[kSliceRows], rid[kSliceRows];
  __shared__ int32_t tc[kSellChunk][kSliceRows + 1];
  __shared__ double tv[kSellChunk][kSliceRows + 1];
  __shared__ int32_t tb[kSellChunk];
  const int64_t s = blockIdx.x;
  const int t = threadIdx.x;
  if(t < kSliceRows) {
    const int64_t row = s * kSliceRows + t;
    rid[t] = row < m ? (int32_t)row : 0;
    rst[t] = row < m ? rowptr[row] - indexing : 0;
    rlen[t] = row < m ? (int32_t)(rowptr[row + 1] - rowptr[row]) : 0;
  }
  __syncthreads();
  const int L = slen[s];
  const int64_t base = soff[s];
  int32_t* __restrict__ bs = sbase + base / kSliceRows;
  // pass 1: slot bases and spans; thread t takes slot t / 8 of the chunk and
  // rows 8 (t % 8) .. 8 (t % 8) + 7
  int wide = 0;
  for(int j0 = 0; j0 < L; j0 += kSellChunk) {
    const int jj = t >> 3, j = j0 + jj, r0 = (t & 7) * 8;
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    if(j < L) {
      for(int r = r0; r < r0 + 8; ++r) {
        if(j < rlen[r]) {
          const int32_t c = (int32_t)(colind[rst[r] + j] - indexing);
          lo = c < lo ? c : lo;
          hi = c > hi ? c : hi;
        }
      }
    }
#pragma unroll
    for(int off = 1; off < 8; off <<= 1) {
      const int32_t ol = __shfl_xor(lo, off, 64), oh = __shfl_xor(hi, off, 64);
      lo = ol < lo ? ol : lo;
      hi = oh > hi ? oh : hi;
    }
    if(j < L && (t & 7) == 0) {
      bs[j] = lo;  // every slot below L holds at least one entry
      if((int64_t)hi - (int64_t)lo >= 65536) wide = 1;
    }
  }
  const bool nar = !__syncthreads_or(wide);
  if(t == 0) narrow[s] = nar ? 1 : 0;
  uint32_t* __restrict__ cw = reinterpret_cast<uint32_t*>(cols) + base;
  for(int j0 = 0; j0 < L; j0 += kSellChunk) {
    if(nar && t < kSellChunk) tb[t] = j0 + t < L ? bs[j0 + t] : 0;
    for(int e = t; e < kSliceRows * kSellChunk; e += 256) {
      const int r = e / kSellChunk, jj = e % kSellChunk, j = j0 + jj;
      int32_t c = nar ? -1 : rid[r];  // narrow padding: -1 marks offset 0
      double v = 0.;
      if(j < rlen[r]) {
        c = (int32_t)(colind[rst[r] + j] - indexing);
        v = nzval[rst[r] + j];
      }
      tc[jj][r] = c;
      tv[jj][r] = v;
    }
    __syncthreads();
    if(nar) {
      // one 32-bit word per row and slot pair (kSellChunk is even)
      for(int e = t; e < kSliceRows * kSellChunk / 2; e += 256) {
        const int pj = e / kSliceRows, r = e % kSliceRows, jj = 2 * pj, j = j0 + jj;
        if(j < L) {
          const int32_t c0 = tc[jj][r], c1 = j + 1 < L ? tc[jj + 1][r] : -1;
          const uint32_t o0 = c0 < 0 ? 0u : (uint32_t)(c0 - tb[jj]);
          const uint32_t o1 = c1 < 0 ? 0u : (uint32_t)(c1 - tb[jj + 1]);
          cw[(int64_t)(j >> 1) * kSliceRows + r] = o0 | (o1 << 16);
        }
      }
    } else {
      for(int e = t; e < kSliceRows * kSellChunk; e += 256) {
        const int jj = e / kSliceRows, r = e % kSliceRows, j = j0 + jj;
        if(j < L) cols[base + (int64_t)j * kSliceRows + r] = tc[jj][r];
      }
    }
    for(int e = t; e < kSliceRows * kSellChunk; e += 256) {
      const int jj = e / kSliceRows, r = e % kSliceRows, j = j0 + jj;
      if(j < L) vals[base + (int64_t)j * kSliceRows + r] = tv[jj][r];
    }
    __syncthreads();
  }
}

// y = alpha A x + beta y on the sliced copy: one slice per wavefront, lane =
// row, kSellUnroll slots per step with non-temporal column/value loads;
// blocks mapped XCD-aware as in spmv_kernel.
// With row_inst / conv (a block-diagonal batch), rows of converged
// instances are neither computed nor written, and slices holding only such
// rows are skipped.
__global__ __launch_bounds__(256) void spmv_sell_kernel(
    int64_t m, int64_t nsl, int64_t per_xcd, const int64_t* __restrict__ soff,
    const int32_t* __restrict__ slen, const int32_t* __restrict__ cols,
    const double* __restrict__ vals, const int32_t* __restrict__ narrow,
    const int32_t* __restrict__ sbase, double alpha, const double* __restrict__ x, double beta,
    double* __restrict__ y, const int32_t* __restrict__ row_inst,
    const int32_t* __restrict__ conv) {
  const int64_t lblk = (int64_t)(blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  // the slice index is uniform over the wavefront: slice data and slot
  // bases come through scalar loads
  const int64_t s = lblk * kWavesPerBlock + __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  if(s >= nsl) return;
  const int lane = threadIdx.x & 63;
  bool active = s * kSliceRows + lane < m;
  if(conv) {
    active = active && !conv[row_inst[s * kSliceRows + lane]];
    if(!ballot(active)) return;  // whole wavefront: nothing to do in this slice
  }
  const int len = slen[s];
  const int64_t so = soff[s];
  const double* v = vals + so + lane;
  double a[kSellUnroll];
#pragma unroll
  for(int u = 0; u < kSellUnroll; ++u) a[u] = 0.;
  int j = 0;
  if(narrow[s]) {
    // 16-bit offsets from the slot bases, two slots per 32-bit word
    const uint32_t* c = reinterpret_cast<const uint32_t*>(cols) + so + lane;
    const int32_t* b = sbase + so / kSliceRows;
    for(; j + kSellUnroll <= len; j += kSellUnroll) {
      uint32_t w[kSellUnroll / 2];
      double vv[kSellUnroll];
#pragma unroll
      for(int u = 0; u < kSellUnroll / 2; ++u)
        w[u] = __builtin_nontemporal_load(c + (int64_t)((j >> 1) + u) * kSliceRows);
#pragma unroll
      for(int u = 0; u < kSellUnroll; ++u)
        vv[u] = __builtin_nontemporal_load(v + (int64_t)(j + u) * kSliceRows);
#pragma unroll
      for(int u = 0; u < kSellUnroll; ++u) {
        const int32_t cc = b[j + u] + (int32_t)((u & 1) ? w[u >> 1] >> 16 : w[u >> 1] & 0xffffu);
        a[u] = fma(vv[u], x[cc], a[u]);
      }
    }
    for(; j < len; ++j) {
      const uint32_t w1 = c[(int64_t)(j >> 1) * kSliceRows];
      const int32_t cc = b[j] + (int32_t)((j & 1) ? w1 >> 16 : w1 & 0xffffu);
      a[0] = fma(v[(int64_t)j * kSliceRows], x[cc], a[0]);
    }
  } else {
    const int32_t* c = cols + so + lane;
    for(; j + kSellUnroll <= len; j += kSellUnroll) {
      int32_t cc[kSellUnroll];
      double vv[kSellUnroll];
#pragma unroll
      for(int u = 0; u < kSellUnroll; ++u) {
        cc[u] = __builtin_nontemporal_load(c + (int64_t)(j + u) * kSliceRows);
        vv[u] = __builtin_nontemporal_load(v + (int64_t)(j + u) * kSliceRows);
      }
#pragma unroll
      for(int u = 0; u < kSellUnroll; ++u) a[u] = fma(vv[u], x[cc[u]], a[u]);
    }
    for(; j < len; ++j)
      a[0] = fma(v[(int64_t)j * kSliceRows], x[c[(int64_t)j * kSliceRows]], a[0]);
  }
  const double acc = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  const int64_t row = s * kSliceRows + lane;
  if(active) y[row] = alpha * acc + beta * y[row];
}

}  // namespace

bool prepare_sell(mhip_context* ctx, const mhip_csr* A, bool any_padding) {
  if(A->sell) return true;
  if(A->m == 0 || A->m != A->n || A->m > INT32_MAX || A->xtiles) return false;
  hipStream_t st = ctx->stream;
  auto S = std::make_unique<SellCopy>();
  S->nsl = (A->m + kSliceRows - 1) / kSliceRows;
  S->meta.alloc(12 * S->nsl + 8);  // soff[nsl] | total | slen[nsl]
  int64_t* soff = S->meta.as<int64_t>();
  int64_t* d_total = soff + S->nsl;
  int32_t* slen = reinterpret_cast<int32_t*>(soff + S->nsl + 1);
  hipLaunchKernelGGL(sell_lengths_kernel, dim3((S->nsl + kWavesPerBlock - 1) / kWavesPerBlock),
                     dim3(256), 0, st, A->m, S->nsl, A->rowptr.as<int64_t>(), slen);
  hipLaunchKernelGGL(sell_offsets_kernel, dim3(1), dim3(1024), 0, st, S->nsl, slen, soff,
                     d_total);
  MHIP_HIP_CHECK(hipGetLastError());
  d2h(ctx, &S->entries, d_total, 8);
  // padding beyond a quarter of the nonzeros costs more streamed bytes than
  // the layout saves
  if(!any_padding && S->entries > A->nnz + A->nnz / 4) return false;
  // room for the copy (12 bytes per stored entry) plus 1 GiB for the caller
  const size_t need = 12 * (size_t)S->entries + (size_t(1) << 30);
  size_t free_b = 0, total_b = 0;
  device_mem_info(&free_b, &total_b);
  size_t spare_b = 0;
  for(auto& b : ctx->spare) spare_b += b.bytes;
  if(free_b + spare_b < need) return false;
  try {
    S->cols = take_buffer(ctx, 4 * (size_t)S->entries);
    S->vals = take_buffer(ctx, 8 * (size_t)S->entries);
  } catch(const Error& e) {  // another allocation won the race: keep the CSR kernel
    if(e.code != MHIP_ERR_OUT_OF_MEMORY) throw;
    (void)hipGetLastError();
    return false;
  }
  S->narrow = take_buffer(ctx, 4 * (size_t)S->nsl);
  S->base = take_buffer(ctx, 4 * (size_t)std::max<int64_t>(1, S->entries / kSliceRows));
  hipLaunchKernelGGL(sell_fill_kernel, dim3(S->nsl), dim3(256), 0, st, A->m,
                     A->rowptr.as<int64_t>(), A->colind.as<int32_t>(), A->nzval.as<double>(),
                     A->indexing, soff, slen, S->cols.as<int32_t>(), S->vals.as<double>(),
                     S->narrow.as<int32_t>(), S->base.as<int32_t>());
  hipLaunchKernelGGL(sell_narrow_count_kernel, dim3(1), dim3(1024), 0, st, S->nsl, slen,
                     S->narrow.as<int32_t>(), d_total);
  MHIP_HIP_CHECK(hipGetLastError());
  d2h(ctx, &S->narrow_entries, d_total, 8);
  A->sell = std::move(S);
  return true;
}

bool prepare_xtiles(mhip_context* ctx, const mhip_csr* A, bool any_size) {
  if(A->xtiles) return true;
  if(A->sell || A->m != A->n || A->m == 0 || A->nnz == 0) return false;
  if(!any_size && A->n < kPanelMinCols) return false;
  if(A->n > std::numeric_limits<int32_t>::max()) return false;
  const int64_t m = A->m;
  const int np = (int)((A->n + kTileCols - 1) / kTileCols);
  // the panel-major copy (10 bytes per nonzero, 4 per row and panel) and,
  // while it is made, the counts and their scan (12 bytes per row and
  // panel), with 1 GiB left for the caller
  const size_t cells = (size_t)np * (size_t)m;
  const size_t need = 10 * (size_t)A->nnz + 4 * (cells + np) + 12 * cells + (size_t(1) << 30);
  size_t free_b = 0, total_b = 0;
  device_mem_info(&free_b, &total_b);
  size_t spare_b = 0;
  for(auto& b : ctx->spare) spare_b += b.bytes;
  if(free_b + spare_b < need) return false;
  // one panel's entries of one block's rows are indexed by 32-bit offsets
  hipStream_t st = ctx->stream;
  std::unique_ptr<XTiles> X(new XTiles());
  X->np = np;
  try {
    DevBuf counts(4 * cells);
    MHIP_HIP_CHECK(hipMemsetAsync(counts.p, 0, 4 * cells, st));
    hipLaunchKernelGGL(xtile_count_kernel, dim3((unsigned)((m + kWavesPerBlock - 1) / kWavesPerBlock)),
                       dim3(256), 0, st, m, A->rowptr.as<int64_t>(), A->colind.as<int32_t>(),
                       A->indexing, counts.as<uint32_t>());
    MHIP_HIP_CHECK(hipGetLastError());
    DevBuf scan(8 * cells);
    size_t tmp = 0;
    MHIP_HIP_CHECK(rocprim::exclusive_scan(nullptr, tmp, counts.as<uint32_t>(), scan.as<int64_t>(),
                                           (int64_t)0, cells, rocprim::plus<int64_t>(), st));
    DevBuf tmpb(tmp + 16);
    MHIP_HIP_CHECK(rocprim::exclusive_scan(tmpb.p, tmp, counts.as<uint32_t>(), scan.as<int64_t>(),
                                           (int64_t)0, cells, rocprim::plus<int64_t>(), st));
    X->entries = A->nnz;
    // the largest panel's entry count must fit the 32-bit row offsets
    X->base.alloc(8 * (size_t)(np + 1));
    X->rowoff.alloc(4 * (size_t)np * (size_t)(m + 1));
    const int64_t nn = (int64_t)np * (m + 1);
    hipLaunchKernelGGL(xtile_rowoff_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, m,
                       np, scan.as<int64_t>(), A->nnz, X->rowoff.as<uint32_t>(),
                       X->base.as<int64_t>());
    MHIP_HIP_CHECK(hipGetLastError());
    X->base_h.resize(np + 1);
    d2h(ctx, X->base_h.data(), X->base.p, 8 * (size_t)(np + 1));
    for(int p = 0; p < np; ++p)
      if(X->base_h[p + 1] - X->base_h[p] > (int64_t)std::numeric_limits<uint32_t>::max() - 4096)
        return false;
    X->cols = take_buffer(ctx, 2 * (size_t)A->nnz);
    X->vals = take_buffer(ctx, 8 * (size_t)A->nnz);
    hipLaunchKernelGGL(xtile_scatter_kernel, dim3((unsigned)((m + kWavesPerBlock - 1) / kWavesPerBlock)),
                       dim3(256), 0, st, m, A->rowptr.as<int64_t>(), A->colind.as<int32_t>(),
                       A->nzval.as<double>(), A->indexing, X->base.as<int64_t>(),
                       X->rowoff.as<uint32_t>(), X->cols.as<uint16_t>(), X->vals.as<double>());
    MHIP_HIP_CHECK(hipGetLastError());
    MHIP_HIP_CHECK(hipStreamSynchronize(st));
  } catch(const Error& e) {
    if(e.code != MHIP_ERR_OUT_OF_MEMORY) throw;
    (void)hipGetLastError();
    return false;
  }
  A->xtiles = std::move(X);
  return true;
}

bool prepare_panels(mhip_context* ctx, const mhip_csr* A, bool any_size) {
  if(A->panels) return true;
  if(A->sell || A->xtiles || A->m != A->n || A->m == 0 || A->nnz == 0) return false;
  if(!any_size && A->n < kPanelMinCols) return false;
  if(A->n > std::numeric_limits<int32_t>::max()) return false;
  hipStream_t st = ctx->stream;
  // nnz-balanced column bounds from a histogram of about 65,536 rows
  DevBuf hist(8 * kHistBins);
  MHIP_HIP_CHECK(hipMemsetAsync(hist.p, 0, 8 * kHistBins, st));
  const int64_t stride = std::max<int64_t>(1, A->m / 65536);
  hipLaunchKernelGGL(col_hist_kernel, dim3(1024), dim3(256), 0, st, A->m, A->n, stride,
                     A->rowptr.as<int64_t>(), A->colind.as<int32_t>(), A->indexing,
                     hist.as<unsigned long long>());
  MHIP_HIP_CHECK(hipGetLastError());
  std::vector<unsigned long long> h(kHistBins);
  d2h(ctx, h.data(), hist.p, 8 * kHistBins);
  double total = 0.;
  for(auto v : h) total += (double)v;
  if(total <= 0.) return false;
  // 8 panels, one per XCD: 16 and 32 (two or four rounds per XCD, smaller x
  // panels) were measured slower on the 2,000,000-determinant matrix, 15.3
  // and 17.9 ms against 14.7 ms, as the rows' runs per panel get shorter
  std::unique_ptr<PanelSplit> P(new PanelSplit());
  const int np = P->np;
  P->bounds.assign(np + 1, 0);
  P->bounds[np] = A->n;
  double run = 0.;
  int p = 1;
  for(int b = 0; b < kHistBins && p < np; ++b) {
    run += (double)h[b];
    while(p < np && run >= total * p / np) {
      // first column of bin b + 1
      P->bounds[p] = std::min<int64_t>(A->n, ((int64_t)(b + 1) * A->n + kHistBins - 1) / kHistBins);
      ++p;
    }
  }
  for(; p < np; ++p) P->bounds[p] = A->n;
  for(int q = 1; q <= np; ++q) P->bounds[q] = std::max(P->bounds[q], P->bounds[q - 1]);
  PanelBounds pb{};
  for(int q = 0; q < np; ++q) pb.b[q] = (int32_t)P->bounds[q];
  try {
    P->off = take_buffer(ctx, 4 * (size_t)A->m * np);
    P->part = take_buffer(ctx, 8 * (size_t)A->m * np);
  } catch(const Error& e) {
    if(e.code != MHIP_ERR_OUT_OF_MEMORY) throw;
    (void)hipGetLastError();
    return false;
  }
  hipLaunchKernelGGL(panel_offsets_kernel, dim3((unsigned)((A->m + 255) / 256)), dim3(256), 0, st,
                     A->m, A->rowptr.as<int64_t>(), A->colind.as<int32_t>(), A->indexing, np, pb,
                     P->off.as<int32_t>());
  MHIP_HIP_CHECK(hipGetLastError());
  A->panels = std::move(P);
  return true;
}

void spmv(mhip_context* ctx, const mhip_csr* A, double alpha, const double* x,
          double beta, double* y) {
  if(A->m == 0) return;
  if(A->xtiles && !A->sell) {
    const XTiles& X = *A->xtiles;
    hipLaunchKernelGGL(spmv_xtile_kernel, dim3((unsigned)((A->m + kTileRows - 1) / kTileRows)),
                       dim3(kTileThreads), 0, ctx->stream, A->m, A->n, X.np, X.base.as<int64_t>(),
                       X.rowoff.as<uint32_t>(), X.cols.as<uint16_t>(), X.vals.as<double>(),
                       alpha, x, beta, y);
    MHIP_HIP_CHECK(hipGetLastError());
    return;
  }
  if(A->panels && !A->sell) {
    const PanelSplit& P = *A->panels;
    const int64_t nblk = (A->m + kRowsPerBlock - 1) / kRowsPerBlock;
    hipLaunchKernelGGL(spmv_panel_kernel, dim3((unsigned)(nblk * P.np)), dim3(256), 0,
                       ctx->stream, A->m, nblk, P.np, A->rowptr.as<int64_t>(),
                       A->colind.as<int32_t>(), A->nzval.as<double>(), A->indexing,
                       P.off.as<int32_t>(), x, P.part.as<double>());
    const unsigned g = (unsigned)std::min<int64_t>(4096, (A->m + 255) / 256);
    hipLaunchKernelGGL(panel_sum_kernel, dim3(g), dim3(256), 0, ctx->stream, A->m, P.np,
                       P.part.as<double>(), alpha, beta, y);
    MHIP_HIP_CHECK(hipGetLastError());
    return;
  }
  if(A->sell) {
    const SellCopy& S = *A->sell;
    const int64_t nblk = (S.nsl + kWavesPerBlock - 1) / kWavesPerBlock;
    const int64_t per = (nblk + kXcds - 1) / kXcds;
    hipLaunchKernelGGL(spmv_sell_kernel, dim3(per * kXcds), dim3(256), 0, ctx->stream, A->m,
                       S.nsl, per, S.soff(), S.slen(), S.cols.as<int32_t>(),
                       S.vals.as<double>(), S.narrow.as<int32_t>(), S.base.as<int32_t>(), alpha,
                       x, beta, y, nullptr, nullptr);
    MHIP_HIP_CHECK(hipGetLastError());
    return;
  }
  const int64_t nblk = (A->m + kRowsPerBlock - 1) / kRowsPerBlock;
  const int64_t per = (nblk + kXcds - 1) / kXcds;
  hipLaunchKernelGGL(spmv_kernel, dim3(per * kXcds), dim3(256), 0, ctx->stream, A->m, per,
                     A->rowptr.as<int64_t>(), A->colind.as<int32_t>(), A->nzval.as<double>(),
                     A->indexing, alpha, x, beta, y);
  MHIP_HIP_CHECK(hipGetLastError());
}

void spmv_sliced_masked(mhip_context* ctx, const mhip_csr* A, const double* x, double* y,
                        const int32_t* row_inst, const int32_t* conv) {
  const SellCopy& S = *A->sell;
  const int64_t nblk = (S.nsl + kWavesPerBlock - 1) / kWavesPerBlock;
  const int64_t per = (nblk + kXcds - 1) / kXcds;
  hipLaunchKernelGGL(spmv_sell_kernel, dim3(per * kXcds), dim3(256), 0, ctx->stream, A->m, S.nsl,
                     per, S.soff(), S.slen(), S.cols.as<int32_t>(), S.vals.as<double>(),
                     S.narrow.as<int32_t>(), S.base.as<int32_t>(), 1., x, 0., y, row_inst, conv);
  MHIP_HIP_CHECK(hipGetLastError());
}

void extract_diagonal(mhip_context* ctx, const mhip_csr* A, double* d_D) {
  if(A->m == 0) return;
  hipLaunchKernelGGL(extract_diag_kernel, dim3((A->m + kWavesPerBlock - 1) / kWavesPerBlock),
                     dim3(256), 0, ctx->stream, A->m, A->rowptr.as<int64_t>(),
                     A->colind.as<int32_t>(), A->nzval.as<double>(), A->indexing, d_D);
  MHIP_HIP_CHECK(hipGetLastError());
}

}  // namespace mhip

using namespace mhip;

extern "C" {

int mhip_csr_info(const mhip_csr* csr, int64_t* m, int64_t* n, int64_t* nnz) {
  return guarded([&] {
    MHIP_REQUIRE(csr, MHIP_ERR_INVALID_ARG, "null csr");
    if(m) *m = csr->m;
    if(n) *n = csr->n;
    if(nnz) *nnz = csr->nnz;
  });
}

int mhip_csr_device_arrays(const mhip_csr* csr, const int64_t** d_rowptr,
                           const int32_t** d_colind, const double** d_nzval) {
  return guarded([&] {
    MHIP_REQUIRE(csr, MHIP_ERR_INVALID_ARG, "null csr");
    if(d_rowptr) *d_rowptr = csr->rowptr.as<int64_t>();
    if(d_colind) *d_colind = csr->colind.as<int32_t>();
    if(d_nzval) *d_nzval = csr->nzval.as<double>();
  });
}

int mhip_csr_copy_to_host_i32(mhip_context* ctx, const mhip_csr* csr,
                              int32_t* rowptr, int32_t* colind,
                              double* nzval) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(csr, MHIP_ERR_INVALID_ARG, "null csr");
    MHIP_REQUIRE(csr->nnz + csr->indexing <= std::numeric_limits<int32_t>::max() &&
                     csr->n <= std::numeric_limits<int32_t>::max(),
                 MHIP_ERR_INDEX_OVERFLOW, "nnz does not fit int32 row pointers");
    std::vector<int64_t> rp(csr->m + 1);
    MHIP_HIP_CHECK(hipMemcpyAsync(rp.data(), csr->rowptr.p, 8 * (csr->m + 1),
                                  hipMemcpyDeviceToHost, ctx->stream));
    if(csr->nnz) {
      if(colind)
        MHIP_HIP_CHECK(hipMemcpyAsync(colind, csr->colind.p, 4 * csr->nnz,
                                      hipMemcpyDeviceToHost, ctx->stream));
      if(nzval)
        MHIP_HIP_CHECK(hipMemcpyAsync(nzval, csr->nzval.p, 8 * csr->nnz,
                                      hipMemcpyDeviceToHost, ctx->stream));
    }
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if(rowptr)
      for(int64_t i = 0; i <= csr->m; ++i) rowptr[i] = (int32_t)rp[i];
  });
}

int mhip_csr_copy_to_host_i64(mhip_context* ctx, const mhip_csr* csr,
                              int64_t* rowptr, int64_t* colind,
                              double* nzval) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(csr, MHIP_ERR_INVALID_ARG, "null csr");
    if(rowptr)
      MHIP_HIP_CHECK(hipMemcpyAsync(rowptr, csr->rowptr.p, 8 * (csr->m + 1),
                                    hipMemcpyDeviceToHost, ctx->stream));
    std::vector<int32_t> c;
    if(csr->nnz) {
      if(colind) {
        c.resize(csr->nnz);
        MHIP_HIP_CHECK(hipMemcpyAsync(c.data(), csr->colind.p, 4 * csr->nnz,
                                      hipMemcpyDeviceToHost, ctx->stream));
      }
      if(nzval)
        MHIP_HIP_CHECK(hipMemcpyAsync(nzval, csr->nzval.p, 8 * csr->nnz,
                                      hipMemcpyDeviceToHost, ctx->stream));
    }
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if(colind)
      for(size_t i = 0; i < c.size(); ++i) colind[i] = c[i];
  });
}

int mhip_csr_create_from_host(mhip_context* ctx, int64_t m, int64_t n,
                              int64_t nnz, int64_t indexing,
                              const int64_t* rowptr, const int32_t* colind,
                              const double* nzval, mhip_csr** out) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(out, MHIP_ERR_INVALID_ARG, "null output pointer");
    *out = nullptr;
    MHIP_REQUIRE(m >= 0 && n >= 0 && nnz >= 0, MHIP_ERR_INVALID_ARG, "negative size");
    MHIP_REQUIRE(indexing == 0 || indexing == 1, MHIP_ERR_INVALID_ARG,
                 "indexing must be 0 or 1");
    MHIP_REQUIRE(rowptr && (nnz == 0 || (colind && nzval)), MHIP_ERR_INVALID_ARG,
                 "null array");
    MHIP_REQUIRE(rowptr[m] - indexing == nnz && rowptr[0] == indexing,
                 MHIP_ERR_INVALID_ARG, "rowptr inconsistent with nnz / indexing");
    for(int64_t i = 0; i < m; ++i)
      MHIP_REQUIRE(rowptr[i + 1] >= rowptr[i], MHIP_ERR_INVALID_ARG,
                   "rowptr not monotone");
    for(int64_t j = 0; j < nnz; ++j)
      MHIP_REQUIRE(colind[j] - indexing >= 0 && colind[j] - indexing < n,
                   MHIP_ERR_INVALID_ARG, "column index out of range");
    std::unique_ptr<mhip_csr> csr(new mhip_csr());
    csr->device = ctx->device;
    csr->m = m;
    csr->n = n;
    csr->nnz = nnz;
    csr->indexing = indexing;
    csr->rowptr.alloc(8 * (m + 1));
    csr->colind.alloc(4 * std::max<int64_t>(nnz, 1));
    csr->nzval.alloc(8 * std::max<int64_t>(nnz, 1));
    MHIP_HIP_CHECK(hipMemcpyAsync(csr->rowptr.p, rowptr, 8 * (m + 1), hipMemcpyHostToDevice,
                                  ctx->stream));
    if(nnz) {
      MHIP_HIP_CHECK(hipMemcpyAsync(csr->colind.p, colind, 4 * nnz, hipMemcpyHostToDevice,
                                    ctx->stream));
      MHIP_HIP_CHECK(hipMemcpyAsync(csr->nzval.p, nzval, 8 * nnz, hipMemcpyHostToDevice,
                                    ctx->stream));
    }
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    *out = csr.release();
  });
}

int mhip_csr_destroy(mhip_csr* csr) {
  return guarded([&] {
    if(!csr) return;
    (void)hipSetDevice(csr->device);
    (void)hipDeviceSynchronize();
    delete csr;
  });
}

int mhip_spmv(mhip_context* ctx, const mhip_csr* A, int64_t K, double alpha,
              const double* d_V, int64_t LDV, double beta, double* d_AV,
              int64_t LDAV) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(A, MHIP_ERR_INVALID_ARG, "null matrix");
    MHIP_REQUIRE(K >= 0 && LDV >= A->n && LDAV >= A->m, MHIP_ERR_INVALID_ARG,
                 "bad block dimensions");
    MHIP_REQUIRE(K == 0 || A->m == 0 || (d_V && d_AV), MHIP_ERR_INVALID_ARG,
                 "null vector");
    for(int64_t k = 0; k < K; ++k) spmv(ctx, A, alpha, d_V + k * LDV, beta, d_AV + k * LDAV);
  });
}

int mhip_csr_prepare_spmv(mhip_context* ctx, const mhip_csr* A, int policy, int* sliced) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(A, MHIP_ERR_INVALID_ARG, "null matrix");
    MHIP_REQUIRE(policy == MHIP_SPMV_SLICED_AUTO || policy == MHIP_SPMV_SLICED_ALWAYS ||
                     policy == MHIP_SPMV_COLUMN_PANELS || policy == MHIP_SPMV_X_TILES,
                 MHIP_ERR_INVALID_ARG, "unknown SpMV layout policy");
    if(policy == MHIP_SPMV_X_TILES) {
      MHIP_REQUIRE(!A->sell && !A->panels, MHIP_ERR_INVALID_ARG,
                   "the matrix already has another SpMV layout");
      MHIP_REQUIRE(A->m == A->n, MHIP_ERR_INVALID_ARG, "x tiles need a square matrix");
      const bool ok = prepare_xtiles(ctx, A, true);
      if(sliced) *sliced = ok ? 3 : 0;
      return;
    }
    if(policy == MHIP_SPMV_COLUMN_PANELS) {
      MHIP_REQUIRE(!A->sell && !A->xtiles, MHIP_ERR_INVALID_ARG,
                   "the matrix already has another SpMV layout");
      MHIP_REQUIRE(A->m == A->n, MHIP_ERR_INVALID_ARG, "column panels need a square matrix");
      const bool ok = prepare_panels(ctx, A, true);
      if(sliced) *sliced = ok ? 2 : 0;
      return;
    }
    const bool ok = prepare_sell(ctx, A, policy == MHIP_SPMV_SLICED_ALWAYS);
    if(sliced) *sliced = ok ? 1 : 0;
  });
}

int mhip_extract_diagonal(mhip_context* ctx, const mhip_csr* A, double* d_D) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(A && (A->m == 0 || d_D), MHIP_ERR_INVALID_ARG, "null argument");
    extract_diagonal(ctx, A, d_D);
  });
}

}  // extern "C"
