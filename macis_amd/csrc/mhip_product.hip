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
#include <cstring>
#include <vector>

#include "mhip_device.hpp"
#include "mhip_internal.hpp"

namespace mhip {
namespace {

constexpr int kWavesPerBlock = 4;
constexpr int64_t kMaxHead = 1 << 16;      // determinants read to find L
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
__global__ void product_count_kernel(int64_t n, int64_t L, const int32_t* __restrict__ ocnt,
                                     const int32_t* __restrict__ ptr1,
                                     const int32_t* __restrict__ ptr2,
                                     uint32_t* __restrict__ counts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const int64_t r = i / L, k = i % L;
  const uint32_t m1 = (uint32_t)(ptr1[k + 1] - ptr1[k]);
  const uint32_t m2 = (uint32_t)(ptr2[k + 1] - ptr2[k]);
  counts[i] = (uint32_t)ocnt[3 * r] * m2 + (uint32_t)ocnt[3 * r + 1] * m1 +
              (uint32_t)ocnt[3 * r + 2];
}

// Fill and elements in one pass: one wavefront per row. Lanes take 64 outer
// neighbours at a time; a wavefront prefix sum of their segment lengths
// gives every segment's offset. Level-2 neighbours (one column each:
// the same inner string) are evaluated by their own lane; level-0/1
// segments are then written by the whole wavefront, one segment at a time.
// Elements are the Slater-Condon values of the tiled path's element pass
// (same function, same diagonal), so the CSR is bit-identical.
__global__ __launch_bounds__(256) void product_fill_elements_kernel(
    IntsView I, int64_t n, int64_t L, int outer_alpha, const uint64_t* __restrict__ ostr,
    const uint64_t* __restrict__ istr, const int32_t* __restrict__ onb_ptr,
    const int32_t* __restrict__ onb, const int32_t* __restrict__ ptr1,
    const int32_t* __restrict__ idx1, const int32_t* __restrict__ ptr2,
    const int32_t* __restrict__ idx2, const double* __restrict__ diag,
    const int64_t* __restrict__ rowptr, int32_t* __restrict__ colind, double* __restrict__ nzval,
    double thresh, int64_t* __restrict__ kept, int* __restrict__ dropped) {
  const int64_t row = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if(row >= n) return;
  const int lane = threadIdx.x & 63;
  const int64_t r = row / L, k = row % L;
  const uint64_t bo = ostr[r], bi = istr[k];
  const uint64_t ba = outer_alpha ? bo : bi, bb = outer_alpha ? bi : bo;
  const double d = diag[row];
  const int64_t st = rowptr[row];
  const int32_t s1 = ptr1[k], len1 = ptr1[k + 1] - s1;
  const int32_t s2 = ptr2[k], len2 = ptr2[k + 1] - s2;
  int64_t pos = st, nk = 0;
  const int32_t j0 = onb_ptr[r], j1 = onb_ptr[r + 1];
  for(int32_t jb = j0; jb < j1; jb += 64) {
    const int32_t j = jb + lane;
    const bool valid = j < j1;
    const int32_t packed = valid ? onb[j] : 0;
    const int64_t r2 = packed >> 2;
    const int e = packed & 3;
    const int32_t len = !valid ? 0 : (e == 2 ? 1 : (e == 1 ? len1 : len2));
    // inclusive wavefront prefix sum of the segment lengths
    int64_t incl = len;
#pragma unroll
    for(int off = 1; off < 64; off <<= 1) {
      const int64_t v = __shfl_up(incl, off, 64);
      if(lane >= off) incl += v;
    }
    const int64_t my = pos + incl - len;
    const int64_t total = __shfl(incl, 63, 64);
    if(valid && e == 2) {  // same inner string, outer double excitation
      const uint64_t ko = ostr[r2];
      const uint64_t ka = outer_alpha ? ko : bi, kb = outer_alpha ? bi : ko;
      const double h = matrix_element(I, ba, bb, ka, kb, d);
      colind[my] = (int32_t)(r2 * L + k);
      nzval[my] = h;
      nk += fabs(h) > thresh ? 1 : 0;
    }
    uint64_t seg = ballot(valid && e < 2);
    while(seg) {
      const int src = (int)ctz64(seg);
      seg &= seg - 1;
      const int64_t sp = __shfl(my, src, 64);
      const int64_t sr = __shfl(r2, src, 64);
      const int se = __shfl(e, src, 64);
      const uint64_t ko = ostr[sr];
      const int32_t sl = se == 1 ? len1 : len2;
      const int32_t* lst = se == 1 ? idx1 + s1 : idx2 + s2;
      for(int32_t t = lane; t < sl; t += 64) {
        const int32_t k2 = lst[t];
        const uint64_t ki = istr[k2];
        const uint64_t ka = outer_alpha ? ko : ki, kb = outer_alpha ? ki : ko;
        const double h = matrix_element(I, ba, bb, ka, kb, d);
        colind[sp + t] = (int32_t)(sr * L + k2);
        nzval[sp + t] = h;
        nk += fabs(h) > thresh ? 1 : 0;
      }
    }
    pos += total;
  }
  nk = wave_sum(nk);
  if(lane == 0) {
    kept[row] = nk;
    if(nk != pos - st) atomicOr(dropped, 1);
  }
}

int popc(uint64_t x) { return __builtin_popcountll(x); }

// Neighbour lists of `s` within `t` at levels <= maxlvl (ascending index).
void neighbour_lists(const std::vector<uint64_t>& s, int maxlvl, std::vector<int32_t>& ptr,
                     std::vector<int32_t>& idx, std::vector<int32_t>* lvl) {
  const size_t n = s.size();
  ptr.assign(n + 1, 0);
  idx.clear();
  if(lvl) lvl->clear();
  for(size_t i = 0; i < n; ++i) {
    for(size_t j = 0; j < n; ++j) {
      const int e = popc(s[i] ^ s[j]);  // 2 * excitation level (equal electron counts)
      if(e <= 2 * maxlvl) {
        idx.push_back((int32_t)j);
        if(lvl) lvl->push_back(e / 2);
      }
    }
    ptr[i + 1] = (int32_t)idx.size();
  }
}

template <typename T>
void upload(mhip_context* ctx, DevBuf& b, const std::vector<T>& h) {
  b.alloc(sizeof(T) * std::max<size_t>(h.size(), 1));
  if(!h.empty())
    MHIP_HIP_CHECK(hipMemcpyAsync(b.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice,
                                  ctx->stream));
}

template <int W>
bool detect(mhip_context* ctx, const uint64_t* d, int64_t n, ProductSpace& ps) {
  if(n < 2) return false;
  hipStream_t st = ctx->stream;
  const int64_t nh = std::min(n, kMaxHead);
  std::vector<uint64_t> head(nh * W);
  MHIP_HIP_CHECK(hipMemcpyAsync(head.data(), d, 8 * W * nh, hipMemcpyDeviceToHost, st));
  MHIP_HIP_CHECK(hipStreamSynchronize(st));
  auto alpha = [&](int64_t i) { return W == 1 ? head[i] & 0xffffffffull : head[2 * i]; };
  auto beta = [&](int64_t i) { return W == 1 ? head[i] >> 32 : head[2 * i + 1]; };
  bool outer_alpha;
  if(alpha(0) == alpha(1)) outer_alpha = true;
  else if(beta(0) == beta(1)) outer_alpha = false;
  else return false;
  auto outer = [&](int64_t i) { return outer_alpha ? alpha(i) : beta(i); };
  auto inner = [&](int64_t i) { return outer_alpha ? beta(i) : alpha(i); };
  int64_t L = 1;
  while(L < nh && outer(L) == outer(0)) ++L;
  if(L == nh && nh < n) return false;  // first run longer than the head read
  if(n % L) return false;
  const int64_t nruns = n / L;
  if((double)nruns * nruns + (double)L * L > kMaxPairTests) return false;
  // inner strings: distinct, equal electron count
  std::vector<uint64_t> I(L);
  for(int64_t k = 0; k < L; ++k) I[k] = inner(k);
  {
    std::vector<uint64_t> s(I);
    std::sort(s.begin(), s.end());
    if(std::adjacent_find(s.begin(), s.end()) != s.end()) return false;
    for(uint64_t v : I)
      if(popc(v) != popc(I[0])) return false;
  }
  // device check of the whole list, outer strings of every run
  DevBuf flag(sizeof(int)), heads(8 * nruns);
  MHIP_HIP_CHECK(hipMemsetAsync(flag.p, 0, sizeof(int), st));
  hipLaunchKernelGGL(product_check_kernel<W>, dim3((n + 255) / 256), dim3(256), 0, st, d, n, L,
                     outer_alpha, flag.as<int>());
  hipLaunchKernelGGL(run_heads_kernel<W>, dim3((nruns + 255) / 256), dim3(256), 0, st, d, nruns,
                     L, outer_alpha, heads.as<uint64_t>());
  MHIP_HIP_CHECK(hipGetLastError());
  int bad = 0;
  std::vector<uint64_t> O(nruns);
  MHIP_HIP_CHECK(hipMemcpyAsync(&bad, flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
  MHIP_HIP_CHECK(hipMemcpyAsync(O.data(), heads.p, 8 * nruns, hipMemcpyDeviceToHost, st));
  MHIP_HIP_CHECK(hipStreamSynchronize(st));
  if(bad) return false;
  {
    std::vector<uint64_t> s(O);
    std::sort(s.begin(), s.end());
    if(std::adjacent_find(s.begin(), s.end()) != s.end()) return false;
    for(uint64_t v : O)
      if(popc(v) != popc(O[0])) return false;
  }
  // neighbour tables
  std::vector<int32_t> optr, oidx, olvl, p1, i1, p2, i2;
  neighbour_lists(O, 2, optr, oidx, &olvl);
  neighbour_lists(I, 1, p1, i1, nullptr);
  neighbour_lists(I, 2, p2, i2, nullptr);
  std::vector<int32_t> onb(oidx.size()), ocnt(3 * nruns, 0);
  for(size_t j = 0; j < oidx.size(); ++j) onb[j] = (oidx[j] << 2) | olvl[j];
  for(int64_t r = 0; r < nruns; ++r)
    for(int32_t j = optr[r]; j < optr[r + 1]; ++j) ocnt[3 * r + olvl[j]]++;
  ps.L = L;
  ps.nruns = nruns;
  ps.outer_alpha = outer_alpha;
  upload(ctx, ps.onb_ptr, optr);
  upload(ctx, ps.onb, onb);
  upload(ctx, ps.ocnt, ocnt);
  upload(ctx, ps.ptr1, p1);
  upload(ctx, ps.idx1, i1);
  upload(ctx, ps.ptr2, p2);
  upload(ctx, ps.idx2, i2);
  upload(ctx, ps.ostr, O);
  upload(ctx, ps.istr, I);
  MHIP_HIP_CHECK(hipStreamSynchronize(st));
  return true;
}

}  // namespace

bool detect_product_space(mhip_context* ctx, int W, const uint64_t* d, int64_t n,
                          ProductSpace& ps) {
  return W == 1 ? detect<1>(ctx, d, n, ps) : detect<2>(ctx, d, n, ps);
}

void product_count(mhip_context* ctx, const ProductSpace& ps, int64_t n, uint32_t* counts) {
  hipLaunchKernelGGL(product_count_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n,
                     ps.L, ps.ocnt.as<int32_t>(), ps.ptr1.as<int32_t>(), ps.ptr2.as<int32_t>(),
                     counts);
  MHIP_HIP_CHECK(hipGetLastError());
}

void product_fill_elements(mhip_context* ctx, int W, const ProductSpace& ps, const IntsView& I,
                           int64_t n, const double* diag, const int64_t* rowptr, int32_t* colind,
                           double* nzval, double thresh, int64_t* kept, int* dropped) {
  (void)W;  // the strings are stored split; the kernel does not read determinants
  hipLaunchKernelGGL(product_fill_elements_kernel,
                     dim3((n + kWavesPerBlock - 1) / kWavesPerBlock), dim3(256), 0, ctx->stream, I,
                     n, ps.L, ps.outer_alpha ? 1 : 0, ps.ostr.as<uint64_t>(),
                     ps.istr.as<uint64_t>(), ps.onb_ptr.as<int32_t>(), ps.onb.as<int32_t>(),
                     ps.ptr1.as<int32_t>(), ps.idx1.as<int32_t>(), ps.ptr2.as<int32_t>(),
                     ps.idx2.as<int32_t>(), diag, rowptr, colind, nzval, thresh, kept, dropped);
  MHIP_HIP_CHECK(hipGetLastError());
}

}  // namespace mhip
