// Batched CAS-CI: many independent instances of uneven size solved together
// (no reference counterpart; the reference runs one compute_casci_rdms per
// instance, util/cas.hpp:32-64).
//
// All instances share one concatenated determinant array; their
// Hamiltonians form one block-diagonal CSR (absolute column indices) built
// by a single launch per stage of the Hamiltonian kernels, driven by a
// work-item table {rows of one instance, that instance's ket range} and a
// per-instance integral table. Davidson runs all instances in lockstep
// (same subspace index, same restart points), which is exactly the
// per-instance algorithm: every reduction is segmented per instance and the
// k x k eigenproblems are solved on the device (one wavefront per instance,
// cyclic Jacobi), so an iteration has no host round trip and is captured
// once per subspace size as a hipGraph and replayed. Converged instances
// are frozen (their rows are skipped by the SpMV and the vector kernels).
// An autotuner picks, per instance-size bucket, the lanes per row of the
// SpMV and the rows per work item of the Hamiltonian scan.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <unordered_map>
#include <vector>

#include "mhip_device.hpp"
#include "mhip_internal.hpp"

namespace mhip {
namespace {

typedef double d4_t __attribute__((ext_vector_type(4)));
constexpr int kMaxK = 32;

// ---- SpMV over a row range with a lanes-per-row schedule, skipping rows
// of converged instances.
template <int LPR>
__global__ __launch_bounds__(256) void bspmv_kernel(
    int64_t r0, int64_t r1, const int64_t* __restrict__ rowptr,
    const int32_t* __restrict__ colind, const double* __restrict__ nzval,
    const double* __restrict__ x, double* __restrict__ y, const int32_t* __restrict__ row_inst,
    const int32_t* __restrict__ conv) {
  constexpr int RPW = 64 / LPR;  // rows per wavefront
  const int lane = threadIdx.x & 63;
  const int sub = lane / LPR, sl = lane % LPR;
  const int64_t row = r0 + ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * RPW + sub;
  const bool active = row < r1 && !(conv && conv[row_inst[row]]);
  double acc = 0.;
  if(active) {
    const int64_t st = rowptr[row], en = rowptr[row + 1];
    for(int64_t j = st + sl; j < en; j += LPR)
      acc = fma(__builtin_nontemporal_load(nzval + j), x[__builtin_nontemporal_load(colind + j)], acc);
  }
#pragma unroll
  for(int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if(active && sl == 0) y[row] = acc;
}

// ---- per-instance state: row offsets and convergence flags ---------------
struct Seg {
  const int64_t* __restrict__ off;  // [ninst + 1] row offsets
  const int32_t* __restrict__ conv;  // [ninst]
};

// Lowest eigenpair of each C_b (lower triangle) by cyclic Jacobi in LDS,
// one wavefront per instance: lam[b], coef[b][:k], the second-lowest
// eigenvector in coef2[b][:k]; the previous coef[b] moves to coef_prev[b]
// (restart coefficients of the restarted variant).
__global__ __launch_bounds__(64) void bjacobi_kernel(Seg sg, int k, const double* __restrict__ C,
                                                     double* __restrict__ lam, double* coef,
                                                     double* __restrict__ coef_prev,
                                                     double* __restrict__ coef2) {
  const int b = blockIdx.x;
  if(sg.conv[b]) return;
  __shared__ double a[kMaxK][kMaxK + 1], v[kMaxK][kMaxK + 1];
  const int t = threadIdx.x;
  if(t < kMaxK) coef_prev[(int64_t)b * kMaxK + t] = coef[(int64_t)b * kMaxK + t];
  for(int e = t; e < k * k; e += 64) {
    const int i = e % k, j = e / k;
    const double x = i >= j ? C[(int64_t)b * kMaxK * kMaxK + i + j * kMaxK]
                            : C[(int64_t)b * kMaxK * kMaxK + j + i * kMaxK];
    a[i][j] = x;
    v[i][j] = i == j ? 1. : 0.;
  }
  __syncthreads();
  for(int sweep = 0; sweep < 60; ++sweep) {
    double off = 0., tot = 0.;
    for(int e = t; e < k * k; e += 64) {
      const int i = e % k, j = e / k;
      tot += a[i][j] * a[i][j];
      if(i != j) off += a[i][j] * a[i][j];
    }
    off = wave_sum(off);
    tot = wave_sum(tot);
    if(off <= 1e-32 * tot || off == 0.) break;
    for(int p = 0; p < k - 1; ++p)
      for(int q = p + 1; q < k; ++q) {
        const double apq = a[p][q];
        if(apq == 0.) continue;
        // elements below the diagonals' rounding are zeroed after four
        // sweeps instead of rotated (Numerical Recipes, jacobi); every
        // thread reads the same values, so the branch is uniform
        const double g = 100. * fabs(apq), app = fabs(a[p][p]), aqq = fabs(a[q][q]);
        if(sweep > 3 && app + g == app && aqq + g == aqq) {
          __syncthreads();
          if(t == 0) {
            a[p][q] = 0.;
            a[q][p] = 0.;
          }
          __syncthreads();
          continue;
        }
        const double tau = (a[q][q] - a[p][p]) / (2. * apq);
        const double tt = (tau >= 0. ? 1. : -1.) / (fabs(tau) + sqrt(1. + tau * tau));
        const double c = 1. / sqrt(1. + tt * tt), s = tt * c;
        __syncthreads();
        if(t < k) {
          const double akp = a[t][p], akq = a[t][q];
          a[t][p] = c * akp - s * akq;
          a[t][q] = s * akp + c * akq;
          const double vkp = v[t][p], vkq = v[t][q];
          v[t][p] = c * vkp - s * vkq;
          v[t][q] = s * vkp + c * vkq;
        }
        __syncthreads();
        if(t < k) {
          const double apk = a[p][t], aqk = a[q][t];
          a[p][t] = c * apk - s * aqk;
          a[q][t] = s * apk + c * aqk;
        }
        __syncthreads();
      }
  }
  // lowest eigenvalue (first on ties)
  int best = 0;
  for(int i = 1; i < k; ++i)
    if(a[i][i] < a[best][best]) best = i;
  int second = best == 0 ? 1 : 0;
  for(int i = 0; i < k; ++i)
    if(i != best && a[i][i] < a[second][second]) second = i;
  if(t == 0) lam[b] = a[best][best];
  if(t < kMaxK) {
    coef[(int64_t)b * kMaxK + t] = t < k ? v[t][best] : 0.;
    coef2[(int64_t)b * kMaxK + t] = t < k && k > 1 ? v[t][second] : 0.;
  }
}

// Restart coefficients of every unconverged instance (one thread each),
// Q[b] = [q0 | q1] with kMaxK entries per column: q0 = c/|c|; for nb == 2,
// q1 = (c_prev, 0) orthogonalised against q0 by two classical Gram-Schmidt
// passes and normalised, or the second Ritz vector's coefficients treated
// the same way when prev_valid is 0 or that norm is <= 1e-10.
__global__ void brestart_coef_kernel(int ninst, const int32_t* __restrict__ conv, int k, int nb,
                                     int prev_valid, const double* __restrict__ coef,
                                     const double* __restrict__ coef_prev,
                                     const double* __restrict__ coef2, double* __restrict__ Q) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if(b >= ninst || conv[b]) return;
  const double* c = coef + (int64_t)b * kMaxK;
  double* q0 = Q + (int64_t)b * 2 * kMaxK;
  double* q1 = q0 + kMaxK;
  double cn = 0.;
  for(int j = 0; j < k; ++j) cn += c[j] * c[j];
  cn = 1. / sqrt(cn);
  for(int j = 0; j < kMaxK; ++j) {
    q0[j] = j < k ? c[j] * cn : 0.;
    q1[j] = 0.;
  }
  if(nb < 2) return;
  for(int attempt = prev_valid ? 0 : 1; attempt < 2; ++attempt) {
    for(int j = 0; j < k; ++j)
      q1[j] = attempt ? coef2[(int64_t)b * kMaxK + j]
                      : (j < k - 1 ? coef_prev[(int64_t)b * kMaxK + j] : 0.);
    for(int pass = 0; pass < 2; ++pass) {
      double s = 0.;
      for(int j = 0; j < k; ++j) s += q0[j] * q1[j];
      for(int j = 0; j < k; ++j) q1[j] -= s * q0[j];
    }
    double pn = 0.;
    for(int j = 0; j < k; ++j) pn += q1[j] * q1[j];
    pn = sqrt(pn);
    if(pn > 1e-10) {
      for(int j = 0; j < k; ++j) q1[j] /= pn;
      return;
    }
  }
}

// In-place restart of the rows of unconverged instances:
// V(:,0..nb-1) = V(:, :k) Q_b, AV(:,0..nb-1) = AV(:, :k) Q_b,
// V(:,nb) = V(:,tcol). Every thread reads its row before writing it.
__global__ __launch_bounds__(256) void brestart_apply_kernel(
    int64_t N, int k, int nb, const int32_t* __restrict__ row_inst,
    const int32_t* __restrict__ conv, const double* __restrict__ Q, double* V, double* AV,
    int64_t ld, int64_t tcol) {
  for(int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N; n += (int64_t)gridDim.x * 256) {
    const int b = row_inst[n];
    if(conv[b]) continue;
    const double* q0 = Q + (int64_t)b * 2 * kMaxK;
    const double* q1 = q0 + kMaxK;
    double v0 = 0., v1 = 0., a0 = 0., a1 = 0.;
    for(int j = 0; j < k; ++j) {
      const double x = V[n + j * ld], y = AV[n + j * ld];
      v0 = fma(x, q0[j], v0);
      a0 = fma(y, q0[j], a0);
      v1 = fma(x, q1[j], v1);
      a1 = fma(y, q1[j], a1);
    }
    const double t = V[n + tcol * ld];
    V[n] = v0;
    AV[n] = a0;
    if(nb > 1) {
      V[n + ld] = v1;
      AV[n + ld] = a1;
    }
    V[n + nb * ld] = t;
  }
}

// X = V c_b; R = AV c_b - lam_b X (rows of unconverged instances).
__global__ __launch_bounds__(256) void bresidual_kernel(
    int64_t N, int k, const double* __restrict__ V, const double* __restrict__ AV, int64_t ld,
    const int32_t* __restrict__ row_inst, const int32_t* __restrict__ conv,
    const double* __restrict__ coef, const double* __restrict__ lam, double* __restrict__ X,
    double* __restrict__ R) {
  for(int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N; n += (int64_t)gridDim.x * 256) {
    const int b = row_inst[n];
    if(conv[b]) continue;
    const double* c = coef + (int64_t)b * kMaxK;
    double x = 0., r = 0.;
    for(int j = 0; j < k; ++j) {
      x = fma(V[n + j * ld], c[j], x);
      r = fma(AV[n + j * ld], c[j], r);
    }
    X[n] = x;
    R[n] = r + (-lam[b]) * x;
  }
}

// ---- tiled per-instance reductions ----------------------------------------
// Instances are cut into tiles of at most kTileRows rows (tile t: rows
// [tiles[2t], tiles[2t+1]) of instance tinst[t]); a block per tile writes
// partial results, a second kernel sums an instance's tiles in a fixed
// order, so the large instances spread over many CUs and results stay
// run-to-run deterministic.
constexpr int64_t kTileRows = 4096;

// part[t*kMaxK + j] = V_j . w over tile t, all j < k in one pass.
__global__ __launch_bounds__(256) void btdots_kernel(const int64_t* __restrict__ tiles,
                                                     const int32_t* __restrict__ tinst,
                                                     const int32_t* __restrict__ conv, int k,
                                                     const double* __restrict__ V, int64_t ld,
                                                     const double* __restrict__ w,
                                                     double* __restrict__ part) {
  const int t = blockIdx.x;
  if(conv[tinst[t]]) return;
  const int64_t t0 = tiles[2 * t], t1 = tiles[2 * t + 1];
  double acc[kMaxK];
#pragma unroll
  for(int j = 0; j < kMaxK; ++j) acc[j] = 0.;
  for(int64_t n = t0 + threadIdx.x; n < t1; n += 256) {
    const double wn = w[n];
#pragma unroll
    for(int j = 0; j < kMaxK; ++j)
      if(j < k) acc[j] = fma(V[n + j * ld], wn, acc[j]);
  }
  __shared__ double lds[4][kMaxK];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for(int j = 0; j < kMaxK; ++j)
    if(j < k) {
      const double v = wave_sum(acc[j]);
      if(lane == 0) lds[wv][j] = v;
    }
  __syncthreads();
  if(threadIdx.x < k)
    part[(int64_t)t * kMaxK + threadIdx.x] =
        (lds[0][threadIdx.x] + lds[1][threadIdx.x]) + (lds[2][threadIdx.x] + lds[3][threadIdx.x]);
}

// out[b*kMaxK + j] = sum over instance b's tiles of part[t*kMaxK + j].
__global__ __launch_bounds__(64) void btreduce_kernel(const int32_t* __restrict__ toff,
                                                      const int32_t* __restrict__ conv, int k,
                                                      const double* __restrict__ part,
                                                      double* __restrict__ out) {
  const int b = blockIdx.x, lane = threadIdx.x;
  if(conv[b]) return;
  for(int j = 0; j < k; ++j) {
    double s = 0.;
    for(int t = toff[b] + lane; t < toff[b + 1]; t += 64) s += part[(int64_t)t * kMaxK + j];
    s = wave_sum(s);
    if(lane == 0) out[(int64_t)b * kMaxK + j] = s;
  }
}

// Partial Gram matrix V^T AV of tile t on MFMA (k <= 32): the four waves'
// accumulators are combined in LDS and stored in the accumulator layout.
template <int TI>  // 16-column tiles per side: 1 for k <= 16, else 2
__global__ __launch_bounds__(256) void btgram_kernel(const int64_t* __restrict__ tiles,
                                                     const int32_t* __restrict__ tinst,
                                                     const int32_t* __restrict__ conv, int k,
                                                     const double* __restrict__ V,
                                                     const double* __restrict__ AV, int64_t ld,
                                                     double* __restrict__ part) {
  const int t = blockIdx.x;
  if(conv[tinst[t]]) return;
  const int64_t lo = tiles[2 * t], hi = tiles[2 * t + 1];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = lane & 15, grp = lane >> 4;
  d4_t acc[TI][TI];
#pragma unroll
  for(int i = 0; i < TI; ++i)
#pragma unroll
    for(int j = 0; j < TI; ++j) acc[i][j] = d4_t{0., 0., 0., 0.};
  for(int64_t n0 = lo + wv * 32; n0 < hi; n0 += 4 * 32) {
    const int64_t nb = n0 + grp * 8;
    double a[TI][8], bb[TI][8];
#pragma unroll
    for(int i = 0; i < TI; ++i) {
      const int c = i * 16 + col;
#pragma unroll
      for(int s = 0; s < 8; ++s) {
        const bool ok = c < k && nb + s < hi;
        a[i][s] = ok ? V[(int64_t)c * ld + nb + s] : 0.;
        bb[i][s] = ok ? AV[(int64_t)c * ld + nb + s] : 0.;
      }
    }
#pragma unroll
    for(int s = 0; s < 8; ++s)
#pragma unroll
      for(int i = 0; i < TI; ++i)
#pragma unroll
        for(int j = 0; j < TI; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][s], bb[j][s], acc[i][j], 0, 0, 0);
  }
  // partial layout of a tile: 2 x 2 blocks of 256 (only TI x TI written;
  // the reduction reads rows and columns below k)
  __shared__ double red[4][1024];
#pragma unroll
  for(int i = 0; i < TI; ++i)
#pragma unroll
    for(int j = 0; j < TI; ++j)
#pragma unroll
      for(int r = 0; r < 4; ++r) red[wv][(i * 2 + j) * 256 + r * 64 + lane] = acc[i][j][r];
  __syncthreads();
  for(int e = threadIdx.x; e < 1024; e += 256)
    if((e >> 8) / 2 < TI && (e >> 8) % 2 < TI)
      part[(int64_t)t * 1024 + e] = (red[0][e] + red[1][e]) + (red[2][e] + red[3][e]);
}

// C_b (k x k, ld kMaxK) = sum over instance b's tiles of the partial Gram
// matrices, in tile order.
__global__ __launch_bounds__(256) void btgram_reduce_kernel(const int32_t* __restrict__ toff,
                                                            const int32_t* __restrict__ conv, int k,
                                                            const double* __restrict__ part,
                                                            double* __restrict__ C) {
  const int b = blockIdx.x;
  if(conv[b]) return;
  for(int e = threadIdx.x; e < 1024; e += 256) {
    const int tile = e / 256, rem = e % 256, r = rem / 64, ln = rem % 64;
    // v_mfma_f64_16x16x4f64: lane ln, register r holds D(ln/16 + 4*r, ln%16)
    const int i = (tile / 2) * 16 + ln / 16 + 4 * r, j = (tile % 2) * 16 + ln % 16;
    if(i >= k || j >= k) continue;
    double s = 0.;
    for(int t = toff[b]; t < toff[b + 1]; ++t) s += part[(int64_t)t * 1024 + e];
    C[(int64_t)b * kMaxK * kMaxK + i + j * kMaxK] = s;
  }
}

// w -= V c_b (CGS update) over unconverged rows.
__global__ __launch_bounds__(256) void bupdate_kernel(int64_t N, int k, const double* __restrict__ V,
                                                      int64_t ld, const int32_t* __restrict__ row_inst,
                                                      const int32_t* __restrict__ conv,
                                                      const double* __restrict__ coef,
                                                      double* __restrict__ w) {
  for(int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N; n += (int64_t)gridDim.x * 256) {
    const int b = row_inst[n];
    if(conv[b]) continue;
    const double* c = coef + (int64_t)b * kMaxK;
    double s = 0.;
    for(int j = 0; j < k; ++j) s = fma(V[n + j * ld], c[j], s);
    w[n] = w[n] - s;
  }
}

// R = -R / (D - lam_b)
__global__ void bprecond_kernel(int64_t N, double* __restrict__ R, const double* __restrict__ D,
                                const int32_t* __restrict__ row_inst,
                                const int32_t* __restrict__ conv, const double* __restrict__ lam) {
  for(int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N; n += (int64_t)gridDim.x * 256) {
    const int b = row_inst[n];
    if(!conv[b]) R[n] = -R[n] / (D[n] - lam[b]);
  }
}

// w *= 1/sqrt(nrm2_b)
__global__ void bscale_kernel(int64_t N, double* __restrict__ w, const int32_t* __restrict__ row_inst,
                              const int32_t* __restrict__ conv, const double* __restrict__ nrm2) {
  for(int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N; n += (int64_t)gridDim.x * 256) {
    const int b = row_inst[n];
    if(!conv[b]) w[n] *= 1. / sqrt(nrm2[(int64_t)b * kMaxK]);
  }
}

// Convergence check after the residual norms: converged instances record E
// and are frozen; counts the unconverged ones.
__global__ void bcheck_kernel(int ninst, int32_t* __restrict__ conv, const double* __restrict__ rn2,
                              const double* __restrict__ lam, double tol, double* __restrict__ E,
                              double* __restrict__ res, int32_t* __restrict__ nleft) {
  __shared__ int cnt;
  if(threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int mine = 0;
  for(int b = threadIdx.x; b < ninst; b += blockDim.x) {
    if(conv[b]) continue;
    const double r = sqrt(rn2[(int64_t)b * kMaxK]);
    res[b] = r;
    E[b] = lam[b];
    if(r < tol)
      conv[b] = 1;
    else
      ++mine;
  }
  atomicAdd(&cnt, mine);
  __syncthreads();
  if(threadIdx.x == 0) nleft[0] = cnt;
}

// Unit guess at the first minimum of D per instance; zeroes X elsewhere.
__global__ __launch_bounds__(256) void bguess_kernel(Seg sg, const double* __restrict__ D,
                                                     double* __restrict__ X) {
  const int b = blockIdx.x;
  const int64_t lo = sg.off[b], hi = sg.off[b + 1];
  __shared__ double sv[256];
  __shared__ int64_t si[256];
  double mn = INFINITY;
  int64_t mi = INT64_MAX;
  for(int64_t n = lo + threadIdx.x; n < hi; n += 256) {
    X[n] = 0.;
    if(D[n] < mn) {
      mn = D[n];
      mi = n;
    }
  }
  sv[threadIdx.x] = mn;
  si[threadIdx.x] = mi;
  __syncthreads();
  for(int o = 128; o > 0; o >>= 1) {
    if(threadIdx.x < o) {
      const int q = threadIdx.x + o;
      if(sv[q] < sv[threadIdx.x] || (sv[q] == sv[threadIdx.x] && si[q] < si[threadIdx.x])) {
        sv[threadIdx.x] = sv[q];
        si[threadIdx.x] = si[q];
      }
    }
    __syncthreads();
  }
  if(threadIdx.x == 0 && si[0] < hi) X[si[0]] = 1.;
}

__global__ void fill_i32_kernel(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) p[i] = v;
}

}  // namespace
}  // namespace mhip

struct mhip_batch {
  int device = 0;
  int W = 1;  // words per determinant (1: wfn_t<64>, 2: wfn_t<128>)
  int ninst = 0;
  int64_t N = 0;
  std::vector<int> order;          // internal position -> caller's instance index
  std::vector<int64_t> off;        // internal row offsets [ninst + 1]
  std::vector<int32_t> norb;       // internal order
  mhip::DevBuf dets, row_inst, d_off, views, ints;
  std::unique_ptr<mhip_csr> H;
  // autotuned schedule per size bucket: rows [r0, r1), lanes per row
  struct Bucket {
    int64_t r0, r1, ndets;
    int lpr;
    int rows_per_item;
  };
  std::vector<Bucket> buckets;
  bool tuned = false;
  // SpMV layout: 1 the sliced copy of the block-diagonal matrix, 0 the CSR
  // kernels with the bucket schedule, -1 undecided (the sliced copy when its
  // padding rule allows it). The autotuner times both and sets it.
  int layout = -1;
  // Davidson workspace
  int64_t max_m = 0;
  mhip::DevBuf V, AV, X, D, C, coef, coef_prev, coef2, Q, lam, tmp, conv, E, res, nleft;
  // row tiles of the per-instance reductions and their partial results
  int ntiles = 0;
  mhip::DevBuf tiles, tinst, toff, part;
  std::map<int, hipGraphExec_t> graphs;
  // list_id[p]: first internal instance with the same determinant list as p
  // (p itself when none); product-space structures of the instances, made
  // at the first build (product_tried: made or found impossible)
  std::vector<int> list_id;
  std::unique_ptr<mhip::ProductBatch> product;
  bool product_tried = false;
  // caller-supplied Davidson guess (internal order); used instead of the
  // diagonal guess while has_guess is set
  mhip::DevBuf guess;
  bool has_guess = false;
  ~mhip_batch() {
    for(auto& g : graphs) (void)hipGraphExecDestroy(g.second);
  }
};

namespace mhip {
namespace {

void launch_spmv(mhip_context* ctx, mhip_batch* B, const double* x, double* y, bool masked) {
  const int32_t* conv = masked ? B->conv.as<int32_t>() : nullptr;
  if(B->H->sell && B->layout != 0) {  // sliced copy of the block-diagonal matrix
    spmv_sliced_masked(ctx, B->H.get(), x, y, B->row_inst.as<int32_t>(), conv);
    return;
  }
  for(auto& bk : B->buckets) {
    const int64_t rows = bk.r1 - bk.r0;
    if(rows <= 0) continue;
    const int rpw = 64 / bk.lpr;
    const int64_t waves = (rows + rpw - 1) / rpw;
    const dim3 grid((waves + 3) / 4);
#define MHIP_BSPMV(L)                                                                           \
  hipLaunchKernelGGL(bspmv_kernel<L>, grid, dim3(256), 0, ctx->stream, bk.r0, bk.r1,            \
                     B->H->rowptr.as<int64_t>(), B->H->colind.as<int32_t>(),                    \
                     B->H->nzval.as<double>(), x, y, B->row_inst.as<int32_t>(), conv)
    if(bk.lpr == 8)
      MHIP_BSPMV(8);
    else if(bk.lpr == 16)
      MHIP_BSPMV(16);
    else if(bk.lpr == 32)
      MHIP_BSPMV(32);
    else
      MHIP_BSPMV(64);
#undef MHIP_BSPMV
  }
  MHIP_HIP_CHECK(hipGetLastError());
}

int red_grid(int64_t N) { return (int)std::max<int64_t>(1, std::min<int64_t>(2048, (N + 255) / 256)); }

// B->tmp[b*kMaxK + j] = V_j . w over instance b (j < k), tiled.
void tdots(mhip_context* ctx, mhip_batch* B, int k, const double* V, const double* w) {
  hipStream_t st = ctx->stream;
  hipLaunchKernelGGL(btdots_kernel, dim3(B->ntiles), dim3(256), 0, st, B->tiles.as<int64_t>(),
                     B->tinst.as<int32_t>(), B->conv.as<int32_t>(), k, V, B->N, w,
                     B->part.as<double>());
  hipLaunchKernelGGL(btreduce_kernel, dim3(B->ninst), dim3(64), 0, st, B->toff.as<int32_t>(),
                     B->conv.as<int32_t>(), k, B->part.as<double>(), B->tmp.as<double>());
}

// CGS2 of w against V(:, :k) per instance, then per-instance normalisation.
void bgram_schmidt(mhip_context* ctx, mhip_batch* B, int k, double* w) {
  hipStream_t st = ctx->stream;
  const int64_t N = B->N, ld = B->N;
  for(int pass = 0; pass < 2; ++pass) {
    tdots(ctx, B, k, B->V.as<double>(), w);
    hipLaunchKernelGGL(bupdate_kernel, dim3(red_grid(N)), dim3(256), 0, st, N, k,
                       B->V.as<double>(), ld, B->row_inst.as<int32_t>(), B->conv.as<int32_t>(),
                       B->tmp.as<double>(), w);
  }
  tdots(ctx, B, 1, w, w);
  hipLaunchKernelGGL(bscale_kernel, dim3(red_grid(N)), dim3(256), 0, st, N, w,
                     B->row_inst.as<int32_t>(), B->conv.as<int32_t>(), B->tmp.as<double>());
}

// One Davidson iteration with subspace index i (k = i + 1), all on device.
void biteration(mhip_context* ctx, mhip_batch* B, int i, double tol) {
  hipStream_t st = ctx->stream;
  const int k = i + 1;
  const int64_t N = B->N, ld = B->N;
  Seg sg{B->d_off.as<int64_t>(), B->conv.as<int32_t>()};
  double* V = B->V.as<double>();
  double* AV = B->AV.as<double>();
  launch_spmv(ctx, B, V + i * ld, AV + i * ld, true);
  if(k <= 16)
    hipLaunchKernelGGL(btgram_kernel<1>, dim3(B->ntiles), dim3(256), 0, st, B->tiles.as<int64_t>(),
                     B->tinst.as<int32_t>(), B->conv.as<int32_t>(), k, V, AV, ld,
                     B->part.as<double>());
  else
    hipLaunchKernelGGL(btgram_kernel<2>, dim3(B->ntiles), dim3(256), 0, st, B->tiles.as<int64_t>(),
                     B->tinst.as<int32_t>(), B->conv.as<int32_t>(), k, V, AV, ld,
                     B->part.as<double>());
  hipLaunchKernelGGL(btgram_reduce_kernel, dim3(B->ninst), dim3(256), 0, st,
                     B->toff.as<int32_t>(), B->conv.as<int32_t>(), k, B->part.as<double>(),
                     B->C.as<double>());
  hipLaunchKernelGGL(bjacobi_kernel, dim3(B->ninst), dim3(64), 0, st, sg, k, B->C.as<double>(),
                     B->lam.as<double>(), B->coef.as<double>(), B->coef_prev.as<double>(),
                     B->coef2.as<double>());
  double* R = V + (i + 1) * ld;
  hipLaunchKernelGGL(bresidual_kernel, dim3(red_grid(N)), dim3(256), 0, st, N, k, V, AV, ld,
                     B->row_inst.as<int32_t>(), B->conv.as<int32_t>(), B->coef.as<double>(),
                     B->lam.as<double>(), B->X.as<double>(), R);
  tdots(ctx, B, 1, R, R);
  hipLaunchKernelGGL(bcheck_kernel, dim3(1), dim3(256), 0, st, B->ninst, B->conv.as<int32_t>(),
                     B->tmp.as<double>(), B->lam.as<double>(), tol, B->E.as<double>(),
                     B->res.as<double>(), B->nleft.as<int32_t>());
  hipLaunchKernelGGL(bprecond_kernel, dim3(red_grid(N)), dim3(256), 0, st, N, R, B->D.as<double>(),
                     B->row_inst.as<int32_t>(), B->conv.as<int32_t>(), B->lam.as<double>());
  bgram_schmidt(ctx, B, k, R);
  MHIP_HIP_CHECK(hipGetLastError());
}

}  // namespace
}  // namespace mhip

using namespace mhip;

extern "C" {

int mhip_batch_create_w(mhip_context* ctx, int words_per_det, int ninst, const int32_t* norb,
                        const int64_t* det_offsets, const uint64_t* dets,
                        const double* const* T, const double* const* V, mhip_batch** out) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(out && ninst > 0 && norb && det_offsets && dets && T && V,
                 MHIP_ERR_INVALID_ARG, "null argument");
    MHIP_REQUIRE(words_per_det == 1 || words_per_det == 2, MHIP_ERR_UNSUPPORTED,
                 "words_per_det must be 1 (wfn_t<64>) or 2 (wfn_t<128>)");
    const int W = words_per_det;
    *out = nullptr;
    std::unique_ptr<mhip_batch> B(new mhip_batch());
    B->device = ctx->device;
    B->W = W;
    B->ninst = ninst;
    // internal order: by size, largest first (contiguous size buckets)
    B->order.resize(ninst);
    std::iota(B->order.begin(), B->order.end(), 0);
    std::stable_sort(B->order.begin(), B->order.end(), [&](int a, int b) {
      return det_offsets[a + 1] - det_offsets[a] > det_offsets[b + 1] - det_offsets[b];
    });
    B->off.assign(ninst + 1, 0);
    B->norb.resize(ninst);
    for(int p = 0; p < ninst; ++p) {
      const int b = B->order[p];
      const int64_t n = det_offsets[b + 1] - det_offsets[b];
      MHIP_REQUIRE(n > 0, MHIP_ERR_INVALID_ARG, "empty instance");
      MHIP_REQUIRE(norb[b] >= 1 && norb[b] <= 32 * W, MHIP_ERR_INVALID_ARG,
                   W == 1 ? "wfn_t<64> instances need norb <= 32"
                          : "wfn_t<128> instances need norb <= 64");
      B->off[p + 1] = B->off[p] + n;
      B->norb[p] = norb[b];
    }
    const int64_t N = B->N = B->off[ninst];
    MHIP_REQUIRE(N < (1ll << 31), MHIP_ERR_INDEX_OVERFLOW, "too many determinants");
    std::vector<uint64_t> hd(W * N);
    std::vector<int32_t> ri(N);
    for(int p = 0; p < ninst; ++p) {
      const int b = B->order[p];
      const uint64_t bad = B->norb[p] >= 64 ? 0ull : ~((1ull << B->norb[p]) - 1ull);
      for(int64_t i = 0; i < B->off[p + 1] - B->off[p]; ++i) {
        const uint64_t* d = dets + W * (det_offsets[b] + i);
        const uint64_t spins = W == 1 ? ((d[0] & 0xffffffffull) | (d[0] >> 32)) : (d[0] | d[1]);
        MHIP_REQUIRE(!(spins & bad), MHIP_ERR_INVALID_ARG,
                     "determinant occupies an orbital >= norb");
        for(int w = 0; w < W; ++w) hd[W * (B->off[p] + i) + w] = d[w];
        ri[B->off[p] + i] = p;
      }
    }
    // instances with identical determinant lists (the usual case for
    // complete spaces of one size) share one product-space structure
    {
      B->list_id.assign(ninst, 0);
      std::unordered_map<uint64_t, std::vector<int>> seen;
      for(int p = 0; p < ninst; ++p) {
        const uint64_t* a = hd.data() + W * B->off[p];
        const int64_t len = W * (B->off[p + 1] - B->off[p]);
        uint64_t h = 1469598103934665603ull ^ (uint64_t)len;
        for(int64_t i = 0; i < len; ++i) h = (h ^ a[i]) * 1099511628211ull;
        B->list_id[p] = p;
        for(int q : seen[h])
          if(B->off[q + 1] - B->off[q] == B->off[p + 1] - B->off[p] &&
             std::equal(a, a + len, hd.data() + W * B->off[q])) {
            B->list_id[p] = q;
            break;
          }
        if(B->list_id[p] == p) seen[h].push_back(p);
      }
    }
    B->dets.alloc(8 * W * N);
    B->row_inst.alloc(4 * N);
    B->d_off.alloc(8 * (ninst + 1));
    MHIP_HIP_CHECK(hipMemcpyAsync(B->dets.p, hd.data(), 8 * W * N, hipMemcpyHostToDevice,
                                  ctx->stream));
    MHIP_HIP_CHECK(hipMemcpyAsync(B->row_inst.p, ri.data(), 4 * N, hipMemcpyHostToDevice, ctx->stream));
    MHIP_HIP_CHECK(hipMemcpyAsync(B->d_off.p, B->off.data(), 8 * (ninst + 1), hipMemcpyHostToDevice,
                                  ctx->stream));
    // integrals + intermediates per instance, packed (impl.hpp:25-70)
    size_t total = 0;
    std::vector<size_t> ioff(ninst);
    for(int p = 0; p < ninst; ++p) {
      const size_t n = B->norb[p];
      ioff[p] = total;
      total += 2 * n * n + 2 * n * n * n * n + 2 * n * n * n + 2 * n * n;
    }
    std::vector<double> packed(total);
    std::vector<IntsView> hv(ninst);
    B->ints.alloc(8 * total);
    double* dbase = B->ints.as<double>();
    for(int p = 0; p < ninst; ++p) {
      const int b = B->order[p];
      const size_t n = B->norb[p], n2 = n * n, n3 = n2 * n, n4 = n3 * n;
      double* h = packed.data() + ioff[p];
      double *hT = h, *hV = hT + n2, *hG = hV + n4, *hGr = hG + n4, *hVr = hGr + n3,
             *hG2 = hVr + n3, *hV2 = hG2 + n2;
      std::memcpy(hT, T[b], 8 * n2);
      std::memcpy(hV, V[b], 8 * n4);
      auto idx = [&](size_t i, size_t j, size_t kk, size_t l) { return i + n * j + n2 * kk + n3 * l; };
      for(size_t l = 0; l < n; ++l)
        for(size_t kk = 0; kk < n; ++kk)
          for(size_t j = 0; j < n; ++j)
            for(size_t i = 0; i < n; ++i) hG[idx(i, j, kk, l)] = hV[idx(i, j, kk, l)] - hV[idx(i, l, kk, j)];
      for(size_t j = 0; j < n; ++j)
        for(size_t i = 0; i < n; ++i)
          for(size_t kk = 0; kk < n; ++kk) {
            hGr[kk + n * i + n2 * j] = hG[idx(kk, kk, i, j)];
            hVr[kk + n * i + n2 * j] = hV[idx(kk, kk, i, j)];
          }
      for(size_t j = 0; j < n; ++j)
        for(size_t i = 0; i < n; ++i) {
          hG2[i + n * j] = 0.5 * hG[idx(i, i, j, j)];
          hV2[i + n * j] = hV[idx(i, i, j, j)];
        }
      double* d = dbase + ioff[p];
      hv[p].n = (int)n;
      hv[p].T = d;
      hv[p].V = d + n2;
      hv[p].G = d + n2 + n4;
      hv[p].G_red = d + n2 + 2 * n4;
      hv[p].V_red = d + n2 + 2 * n4 + n3;
      hv[p].G2_red = d + n2 + 2 * n4 + 2 * n3;
      hv[p].V2_red = d + 2 * n2 + 2 * n4 + 2 * n3;
      // V(pq|rs) stored equal to V(rs|pq) bit for bit: the product-space
      // fill may read the pair-swapped (contiguous) entry
      bool sym = true;
      for(size_t q = 0; q < n4 && sym; ++q) {
        const size_t i = q % n, j = (q / n) % n, k = (q / n2) % n, l = q / n3;
        sym = std::memcmp(&hV[q], &hV[k + n * (l + n * (i + n * j))], sizeof(double)) == 0;
      }
      hv[p].pair_sym = sym ? 1 : 0;
    }
    MHIP_HIP_CHECK(hipMemcpyAsync(B->ints.p, packed.data(), 8 * total, hipMemcpyHostToDevice, ctx->stream));
    B->views.alloc(sizeof(IntsView) * ninst);
    MHIP_HIP_CHECK(hipMemcpyAsync(B->views.p, hv.data(), sizeof(IntsView) * ninst,
                                  hipMemcpyHostToDevice, ctx->stream));
    // size buckets (instances with equal determinant counts are contiguous)
    for(int p = 0; p < ninst;) {
      int q = p;
      while(q < ninst && B->off[q + 1] - B->off[q] == B->off[p + 1] - B->off[p]) ++q;
      B->buckets.push_back({B->off[p], B->off[q], B->off[p + 1] - B->off[p], 64, 8});
      p = q;
    }
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    *out = B.release();
  });
}

int mhip_batch_create(mhip_context* ctx, int ninst, const int32_t* norb,
                      const int64_t* det_offsets, const uint64_t* dets, const double* const* T,
                      const double* const* V, mhip_batch** out) {
  return mhip_batch_create_w(ctx, 1, ninst, norb, det_offsets, dets, T, V, out);
}

int mhip_batch_destroy(mhip_batch* B) {
  return guarded([&] {
    if(!B) return;
    (void)hipSetDevice(B->device);
    (void)hipDeviceSynchronize();
    delete B;
  });
}

int mhip_batch_info(const mhip_batch* B, int64_t* ndets, int64_t* nnz) {
  return guarded([&] {
    MHIP_REQUIRE(B, MHIP_ERR_INVALID_ARG, "null batch");
    if(ndets) *ndets = B->N;
    if(nnz) *nnz = B->H ? B->H->nnz : 0;
  });
}

// Builds the block-diagonal Hamiltonian of all instances.
int mhip_batch_build(mhip_context* ctx, mhip_batch* B, double thresh) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(B, MHIP_ERR_INVALID_ARG, "null batch");
    B->H.reset(new mhip_csr());
    B->H->device = ctx->device;
    B->H->m = B->H->n = B->N;
    // product-space instances (complete active spaces): the neighbour-table
    // walk of every instance in one launch per stage, unless the tiled path
    // is forced
    if(!B->product_tried && ctx->build_path != MHIP_BUILD_TILED) {
      B->product_tried = true;
      std::unique_ptr<ProductBatch> pb(new ProductBatch());
      if(prepare_batch_product(ctx, B->W, B->dets.as<uint64_t>(), B->off, B->list_id,
                               B->views.as<IntsView>(), *pb))
        B->product = std::move(pb);
    }
    if(B->product && ctx->build_path != MHIP_BUILD_TILED) {
      BuildExtra ex;
      ex.views = B->views.as<IntsView>();
      ex.row_inst = B->row_inst.as<int32_t>();
      ex.bproduct = B->product.get();
      if(B->W == 1)
        build_impl<1>(ctx, B->dets.as<uint64_t>(), B->N, B->dets.as<uint64_t>(), B->N, thresh,
                      B->H.get(), ex);
      else
        build_impl<2>(ctx, B->dets.as<uint64_t>(), B->N, B->dets.as<uint64_t>(), B->N, thresh,
                      B->H.get(), ex);
      return;
    }
    MHIP_REQUIRE(ctx->build_path != MHIP_BUILD_PRODUCT, MHIP_ERR_INVALID_ARG,
                 "product-space build forced but an instance is not a product space");
    // work items: rows of one instance, in groups of the bucket's size
    std::vector<int64_t> items;
    for(auto& bk : B->buckets)
      for(int64_t r = bk.r0; r < bk.r1; r += bk.ndets)
        for(int64_t q = r; q < r + bk.ndets; q += bk.rows_per_item) {
          items.push_back(q);
          items.push_back(std::min<int64_t>(bk.rows_per_item, r + bk.ndets - q));
          items.push_back(r);
          items.push_back(r + bk.ndets);
        }
    DevBuf ditems(8 * items.size());
    MHIP_HIP_CHECK(hipMemcpyAsync(ditems.p, items.data(), 8 * items.size(), hipMemcpyHostToDevice,
                                  ctx->stream));
    BuildExtra ex;
    ex.items = ditems.as<int64_t>();
    ex.nitems = (int64_t)items.size() / 4;
    ex.views = B->views.as<IntsView>();
    ex.row_inst = B->row_inst.as<int32_t>();
    if(B->W == 1)
      build_impl<1>(ctx, B->dets.as<uint64_t>(), B->N, B->dets.as<uint64_t>(), B->N, thresh,
                    B->H.get(), ex);
    else
      build_impl<2>(ctx, B->dets.as<uint64_t>(), B->N, B->dets.as<uint64_t>(), B->N, thresh,
                    B->H.get(), ex);
  });
}

// Autotuner: per size bucket, time the SpMV with 8/16/32/64 lanes per row
// on the built matrix and keep the fastest. Rows per scan work item are set
// from the same choice (short rows -> more rows per item). Then the whole
// tuned CSR schedule is timed against the sliced copy (when its padding rule
// allows one) and the faster layout is kept for the Davidson.
int mhip_batch_tune(mhip_context* ctx, mhip_batch* B, int32_t* lanes_out) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(B && B->H, MHIP_ERR_INVALID_ARG, "build the batch first");
    DevBuf x(8 * B->N), y(8 * B->N);
    MHIP_HIP_CHECK(hipMemsetAsync(x.p, 0, 8 * B->N, ctx->stream));
    Timer t(ctx->stream);
    B->layout = 0;  // the buckets are timed on the CSR kernels
    int bi = 0;
    for(auto& bk : B->buckets) {
      double best = 1e30;
      int bestl = 64;
      for(int l : {8, 16, 32, 64}) {
        bk.lpr = l;
        std::vector<mhip_batch::Bucket> saved = B->buckets;
        B->buckets = {bk};
        launch_spmv(ctx, B, x.as<double>(), y.as<double>(), false);  // warm
        t.start();
        for(int r = 0; r < 3; ++r) launch_spmv(ctx, B, x.as<double>(), y.as<double>(), false);
        t.stop();
        const double ms = t.ms();
        B->buckets = saved;
        if(ms < best) {
          best = ms;
          bestl = l;
        }
      }
      bk.lpr = bestl;
      bk.rows_per_item = bestl >= 32 ? 8 : 4;
      if(lanes_out) lanes_out[bi] = bestl;
      ++bi;
    }
    auto time3 = [&] {
      launch_spmv(ctx, B, x.as<double>(), y.as<double>(), false);  // warm
      t.start();
      for(int r = 0; r < 3; ++r) launch_spmv(ctx, B, x.as<double>(), y.as<double>(), false);
      t.stop();
      return t.ms();
    };
    const double csr_ms = time3();
    B->layout = 0;
    if(prepare_sell(ctx, B->H.get())) {
      B->layout = 1;
      const double sell_ms = time3();
      B->layout = sell_ms < csr_ms ? 1 : 0;
    }
    B->tuned = true;
  });
}

int mhip_batch_spmv_layout(mhip_context* ctx, mhip_batch* B, int layout, int* current) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(B && B->H, MHIP_ERR_INVALID_ARG, "build the batch first");
    MHIP_REQUIRE(layout >= -2 && layout <= 1, MHIP_ERR_INVALID_ARG, "bad layout");
    if(layout == 1)
      MHIP_REQUIRE(prepare_sell(ctx, B->H.get()), MHIP_ERR_UNSUPPORTED,
                   "no sliced copy for this batch (padding rule or memory)");
    if(layout != -2) B->layout = layout;
    if(current) *current = B->layout == 0 ? 0 : (B->H->sell ? 1 : -1);
  });
}

// Lockstep Davidson over all instances (per-instance algorithm of
// davidson.hpp:128-225 with the restart variant when max_restarts > 0).
// E / iters / residual are per CALLER instance index. Returns
// MHIP_ERR_NOT_CONVERGED if some instance did not converge.
int mhip_batch_davidson(mhip_context* ctx, mhip_batch* B, int64_t max_m, int64_t max_restarts,
                        double tol, int use_graphs, double* E, double* residual,
                        double* ms_out) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(B && B->H && E, MHIP_ERR_INVALID_ARG, "build the batch first");
    int64_t min_n = B->N;
    for(int p = 0; p < B->ninst; ++p) min_n = std::min(min_n, B->off[p + 1] - B->off[p]);
    MHIP_REQUIRE(max_m >= 2 && max_m < kMaxK && max_m <= min_n, MHIP_ERR_INVALID_ARG,
                 "max_m must be in [2, 31] and <= the smallest instance");
    hipStream_t st = ctx->stream;
    const int64_t N = B->N, ld = B->N;
    if(B->max_m != max_m) {
      for(auto& g : B->graphs) (void)hipGraphExecDestroy(g.second);
      B->graphs.clear();
      B->max_m = max_m;
      B->V.alloc(8 * N * (max_m + 1));
      B->AV.alloc(8 * N * (max_m + 1));
      B->X.alloc(8 * N);
      B->D.alloc(8 * N);
      B->C.alloc(8 * (size_t)B->ninst * kMaxK * kMaxK);
      B->coef.alloc(8 * (size_t)B->ninst * kMaxK);
      B->coef_prev.alloc(8 * (size_t)B->ninst * kMaxK);
      B->coef2.alloc(8 * (size_t)B->ninst * kMaxK);
      B->Q.alloc(8 * (size_t)B->ninst * 2 * kMaxK);
      B->lam.alloc(8 * B->ninst);
      B->tmp.alloc(8 * (size_t)B->ninst * kMaxK);
      B->conv.alloc(4 * B->ninst);
      B->E.alloc(8 * B->ninst);
      B->res.alloc(8 * B->ninst);
      B->nleft.alloc(4);
      // the SpMV forms AV = 1 * H v + 0 * AV (gespmbv, spmbv.hpp:71), which
      // reads AV: a reused block may hold NaN or Inf patterns, so the
      // subspace storage starts at zero (outside any graph capture)
      MHIP_HIP_CHECK(hipMemsetAsync(B->V.p, 0, 8 * N * (max_m + 1), st));
      MHIP_HIP_CHECK(hipMemsetAsync(B->AV.p, 0, 8 * N * (max_m + 1), st));
    }
    if(!B->ntiles) {
      std::vector<int64_t> tl;
      std::vector<int32_t> ti, to(B->ninst + 1, 0);
      for(int b = 0; b < B->ninst; ++b) {
        for(int64_t r = B->off[b]; r < B->off[b + 1]; r += kTileRows) {
          tl.push_back(r);
          tl.push_back(std::min<int64_t>(r + kTileRows, B->off[b + 1]));
          ti.push_back(b);
        }
        to[b + 1] = (int32_t)ti.size();
      }
      B->ntiles = (int)ti.size();
      B->tiles.alloc(8 * tl.size());
      B->tinst.alloc(4 * ti.size());
      B->toff.alloc(4 * to.size());
      B->part.alloc(8 * (size_t)B->ntiles * 1024);
      MHIP_HIP_CHECK(hipMemcpyAsync(B->tiles.p, tl.data(), 8 * tl.size(), hipMemcpyHostToDevice, st));
      MHIP_HIP_CHECK(hipMemcpyAsync(B->tinst.p, ti.data(), 4 * ti.size(), hipMemcpyHostToDevice, st));
      MHIP_HIP_CHECK(hipMemcpyAsync(B->toff.p, to.data(), 4 * to.size(), hipMemcpyHostToDevice, st));
      MHIP_HIP_CHECK(hipStreamSynchronize(st));
    }
    // the sliced SpMV copy when padding stays small (rows of one instance
    // have one length) and memory allows, unless the autotuner chose the
    // CSR schedule; made once per built batch
    if(B->layout != 0) prepare_sell(ctx, B->H.get());
    Timer tall(st);
    tall.start();
    Seg sg{B->d_off.as<int64_t>(), B->conv.as<int32_t>()};
    double* V = B->V.as<double>();
    double* AV = B->AV.as<double>();
    hipLaunchKernelGGL(fill_i32_kernel, dim3((B->ninst + 255) / 256), dim3(256), 0, st,
                       B->conv.as<int32_t>(), (int64_t)B->ninst, 0);
    extract_diagonal(ctx, B->H.get(), B->D.as<double>());
    if(B->has_guess)
      MHIP_HIP_CHECK(hipMemcpyAsync(B->X.p, B->guess.p, 8 * N, hipMemcpyDeviceToDevice, st));
    else
      hipLaunchKernelGGL(bguess_kernel, dim3(B->ninst), dim3(256), 0, st, sg, B->D.as<double>(),
                         B->X.as<double>());
    MHIP_HIP_CHECK(hipMemcpyAsync(V, B->X.p, 8 * N, hipMemcpyDeviceToDevice, st));
    launch_spmv(ctx, B, V, AV, false);
    MHIP_HIP_CHECK(hipMemcpyAsync(V + ld, AV, 8 * N, hipMemcpyDeviceToDevice, st));
    bgram_schmidt(ctx, B, 1, V + ld);
    int32_t nleft = B->ninst;
    int i0 = 1;
    const int nb = max_m >= 3 ? 2 : 1;
    // restart unconverged instances from [V c, V q1, t] (AV by the same
    // combinations, no SpMV); the last two iterations shared the basis when
    // the cycle ran at least two of them. Converged instances are frozen.
    auto restart = [&](int first) {
      const int k = (int)max_m;
      hipLaunchKernelGGL(brestart_coef_kernel, dim3((B->ninst + 63) / 64), dim3(64), 0, st,
                         B->ninst, B->conv.as<int32_t>(), k, nb, (int)(max_m - 1 > first),
                         B->coef.as<double>(), B->coef_prev.as<double>(), B->coef2.as<double>(),
                         B->Q.as<double>());
      hipLaunchKernelGGL(brestart_apply_kernel, dim3(red_grid(N)), dim3(256), 0, st, N, k, nb,
                         B->row_inst.as<int32_t>(), B->conv.as<int32_t>(), B->Q.as<double>(), V,
                         AV, ld, max_m);
    };
    if(use_graphs) {
      // one graph per restart cycle: its iterations and the restart after
      // them, replayed with one host read of the unconverged count per
      // cycle. Every kernel skips converged instances, so iterations after
      // the last convergence, and a restart after the last cycle, change no
      // instance's result.
      for(int64_t rst = 0; nleft; ++rst) {
        auto it = B->graphs.find(i0);
        if(it == B->graphs.end()) {
          hipGraph_t g;
          MHIP_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
          for(int i = i0; i < max_m; ++i) biteration(ctx, B, i, tol);
          restart(i0);
          MHIP_HIP_CHECK(hipStreamEndCapture(st, &g));
          hipGraphExec_t ge;
          MHIP_HIP_CHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
          (void)hipGraphDestroy(g);
          it = B->graphs.emplace(i0, ge).first;
        }
        MHIP_HIP_CHECK(hipGraphLaunch(it->second, st));
        d2h(ctx, &nleft, B->nleft.p, 4);
        if(rst >= max_restarts) break;
        i0 = nb;
      }
    } else {
      for(int64_t rst = 0;; ++rst) {
        for(int i = i0; i < max_m && nleft; ++i) {
          biteration(ctx, B, i, tol);
          d2h(ctx, &nleft, B->nleft.p, 4);
        }
        if(!nleft || rst >= max_restarts) break;
        restart(i0);
        MHIP_HIP_CHECK(hipGetLastError());
        i0 = nb;
      }
    }
    tall.stop();
    const double ms = tall.ms();
    if(ms_out) *ms_out = ms;
    std::vector<double> hE(B->ninst), hr(B->ninst);
    d2h(ctx, hE.data(), B->E.p, 8 * B->ninst);
    d2h(ctx, hr.data(), B->res.p, 8 * B->ninst);
    for(int p = 0; p < B->ninst; ++p) {
      E[B->order[p]] = hE[p];
      if(residual) residual[B->order[p]] = hr[p];
    }
    if(nleft) throw Error(MHIP_ERR_NOT_CONVERGED, "Davidson Did Not Converge!");
  });
}

// Position of caller instance `inst` in the internal (size-sorted) order.
static int internal_pos(const mhip_batch* B, int inst) {
  for(int p = 0; p < B->ninst; ++p)
    if(B->order[p] == inst) return p;
  return -1;
}

// Caller-order host copy <-> internal-order device array of N doubles.
static void batch_to_host(mhip_context* ctx, const mhip_batch* B, const double* d, double* h) {
  std::vector<double> tmp(B->N);
  d2h(ctx, tmp.data(), d, 8 * B->N);
  int64_t o = 0;
  for(int b = 0; b < B->ninst; ++b) {  // caller offsets follow caller order
    const int p = internal_pos(B, b);
    const int64_t n = B->off[p + 1] - B->off[p];
    std::memcpy(h + o, tmp.data() + B->off[p], 8 * n);
    o += n;
  }
}

int mhip_batch_diagonal(mhip_context* ctx, mhip_batch* B, double* D) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(B && B->H && D, MHIP_ERR_INVALID_ARG, "build the batch first");
    DevBuf d(8 * B->N);
    extract_diagonal(ctx, B->H.get(), d.as<double>());
    batch_to_host(ctx, B, d.as<double>(), D);
  });
}

int mhip_batch_set_guess(mhip_context* ctx, mhip_batch* B, const double* X) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(B, MHIP_ERR_INVALID_ARG, "null batch");
    if(!X) {
      B->has_guess = false;
      return;
    }
    std::vector<double> tmp(B->N);
    int64_t o = 0;
    for(int b = 0; b < B->ninst; ++b) {
      const int p = internal_pos(B, b);
      const int64_t n = B->off[p + 1] - B->off[p];
      std::memcpy(tmp.data() + B->off[p], X + o, 8 * n);
      o += n;
    }
    B->guess.alloc(8 * B->N);
    MHIP_HIP_CHECK(hipMemcpyAsync(B->guess.p, tmp.data(), 8 * B->N, hipMemcpyHostToDevice,
                                  ctx->stream));
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    B->has_guess = true;
  });
}

int mhip_batch_vectors(mhip_context* ctx, mhip_batch* B, double* X) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(B && X, MHIP_ERR_INVALID_ARG, "null argument");
    MHIP_REQUIRE(B->max_m > 0, MHIP_ERR_INVALID_ARG, "run mhip_batch_davidson first");
    batch_to_host(ctx, B, B->X.as<double>(), X);
  });
}

int mhip_batch_instance_csr(mhip_context* ctx, mhip_batch* B, int inst, int64_t* rowptr,
                            int32_t* colind, double* nzval, int64_t* nnz) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(B && B->H && nnz, MHIP_ERR_INVALID_ARG, "build the batch first");
    MHIP_REQUIRE(inst >= 0 && inst < B->ninst, MHIP_ERR_INVALID_ARG, "instance out of range");
    const int p = internal_pos(B, inst);
    const int64_t r0 = B->off[p], m = B->off[p + 1] - r0;
    std::vector<int64_t> rp(m + 1);
    d2h(ctx, rp.data(), B->H->rowptr.as<int64_t>() + r0, 8 * (m + 1));
    const int64_t base = rp[0], cnt = rp[m] - base;
    *nnz = cnt;
    if(!rowptr) return;
    MHIP_REQUIRE(colind && nzval, MHIP_ERR_INVALID_ARG, "null output array");
    for(int64_t i = 0; i <= m; ++i) rowptr[i] = rp[i] - base;
    if(!cnt) return;
    d2h(ctx, colind, B->H->colind.as<int32_t>() + base, 4 * cnt);
    d2h(ctx, nzval, B->H->nzval.as<double>() + base, 8 * cnt);
    for(int64_t j = 0; j < cnt; ++j) colind[j] -= (int32_t)r0;
  });
}

}  // extern "C"
