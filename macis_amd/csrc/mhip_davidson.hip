// K3-K5: device-resident Davidson (davidson.hpp:61-225) and the serial
// selected_ci_diag routine (selected_ci_diag.hpp:75-207).
//
// Per iteration everything N-long stays in HBM:
//   SpMV AV(:,i) = A V(:,i)                                      (K2)
//   Rayleigh-Ritz C = V^T AV (k x k) as a tall-skinny FP64 GEMM on MFMA
//     (v_mfma_f64_16x16x4f64), recomputed in full as rayleigh_ritz.hpp:73
//     does; the k x k (k <= 21) eigenproblem runs on the host (Jacobi).
//   Residual X = V c, R = AV c - lam X, ||R||^2 in one fused pass.
//   Preconditioner R = -R / (D - lam), CGS2 projections (two dot/update
//     passes) and normalisation (davidson.hpp:111-126, 212-217).
// Reductions are written as per-block partials and summed in a fixed order,
// so results are run-to-run deterministic.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "mhip_device.hpp"
#include "mhip_internal.hpp"

namespace mhip {
namespace {

typedef double d4_t __attribute__((ext_vector_type(4)));

constexpr int kMaxK = 32;  // columns handled by one launch of the projection kernels
constexpr int64_t kMaxSubspace = 1024;  // widest Davidson subspace (max_m)
constexpr int kRedBlocks = 512;
constexpr int kGramBlocks = 256;  // partial Gram blocks (4 wavefronts each)

int gram_blocks(int64_t N) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(kGramBlocks, (N + 127) / 128));
}

int red_blocks(int64_t N) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(kRedBlocks, (N + 255) / 256));
}

// ---- tall-skinny GEMM on MFMA: partial C = A^T B per block --------------
// A: N x k (lda), B: N x m (ldb), k, m <= 16*TI, 16*TJ. Lane l owns output
// column l%16 and reduction slot l/16; each lane feeds 8 consecutive rows
// per step, one MFMA (K = 4) per row offset s: slot g = l/16 carries row
// n0 + 8g + s in both operands, so the pairing of A and B rows is exact.
template <int TI, int TJ>
__global__ __launch_bounds__(256) void gram_mfma_kernel(
    int64_t N, int k, const double* __restrict__ A, int64_t lda, int m,
    const double* __restrict__ B, int64_t ldb, double* __restrict__ partials) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t gw = (int64_t)blockIdx.x * 4 + wv, nw = (int64_t)gridDim.x * 4;
  const int col = lane & 15, grp = lane >> 4;
  d4_t acc[TI][TJ];
#pragma unroll
  for(int i = 0; i < TI; ++i)
#pragma unroll
    for(int j = 0; j < TJ; ++j) acc[i][j] = d4_t{0., 0., 0., 0.};
  for(int64_t n0 = gw * 32; n0 < N; n0 += nw * 32) {
    const int64_t nb = n0 + grp * 8;
    double a[TI][8], b[TJ][8];
#pragma unroll
    for(int i = 0; i < TI; ++i) {
      const int c = i * 16 + col;
      const double* p = A + (int64_t)c * lda;
#pragma unroll
      for(int s = 0; s < 8; ++s) a[i][s] = (c < k && nb + s < N) ? p[nb + s] : 0.;
    }
#pragma unroll
    for(int j = 0; j < TJ; ++j) {
      const int c = j * 16 + col;
      const double* p = B + (int64_t)c * ldb;
#pragma unroll
      for(int s = 0; s < 8; ++s) b[j][s] = (c < m && nb + s < N) ? p[nb + s] : 0.;
    }
#pragma unroll
    for(int s = 0; s < 8; ++s)
#pragma unroll
      for(int i = 0; i < TI; ++i)
#pragma unroll
        for(int j = 0; j < TJ; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][s], b[j][s], acc[i][j], 0, 0, 0);
  }
  __shared__ double red[4][TI * TJ * 256];
#pragma unroll
  for(int i = 0; i < TI; ++i)
#pragma unroll
    for(int j = 0; j < TJ; ++j)
#pragma unroll
      for(int r = 0; r < 4; ++r) red[wv][(i * TJ + j) * 256 + r * 64 + lane] = acc[i][j][r];
  __syncthreads();
  for(int e = threadIdx.x; e < TI * TJ * 256; e += 256)
    partials[(int64_t)blockIdx.x * (TI * TJ * 256) + e] =
        (red[0][e] + red[1][e]) + (red[2][e] + red[3][e]);
}

// Sum the gram partials over blocks and scatter to column-major C (ldc).
// Block: 64 outputs x 4 block slices (fixed combination order).
// Output layout of v_mfma_f64_16x16x4f64 (measured on gfx950): lane l,
// register r holds D(l/16 + 4*r, l%16).
template <int TI, int TJ>
__global__ __launch_bounds__(256) void gram_finalize_kernel(const double* __restrict__ partials,
                                                            int nblocks, int k, int m,
                                                            double* __restrict__ C, int ldc) {
  const int e = blockIdx.x * 64 + (threadIdx.x & 63), slice = threadIdx.x >> 6;
  double s = 0.;
  for(int b = slice; b < nblocks; b += 4) s += partials[(int64_t)b * (TI * TJ * 256) + e];
  __shared__ double red[4][64];
  red[slice][threadIdx.x & 63] = s;
  __syncthreads();
  if(slice) return;
  const int tile = e / 256, rem = e % 256, r = rem / 64, lane = rem % 64;
  const int ti = tile / TJ, tj = tile % TJ;
  const int i = ti * 16 + lane / 16 + 4 * r, j = tj * 16 + lane % 16;
  if(i >= k || j >= m) return;
  C[i + (int64_t)j * ldc] = (red[0][threadIdx.x] + red[1][threadIdx.x]) +
                            (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

// inner[j] partials = V(:,j) . w for j < k (vector FMA; a GEMV is HBM bound).
__global__ __launch_bounds__(256) void dots_kernel(int64_t N, int k,
                                                    const double* __restrict__ V,
                                                    int64_t ldv,
                                                    const double* __restrict__ w,
                                                    double* __restrict__ partials) {
  double acc[kMaxK];
#pragma unroll
  for(int j = 0; j < kMaxK; ++j) acc[j] = 0.;
  for(int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N;
      n += (int64_t)gridDim.x * 256) {
    const double wn = w[n];
#pragma unroll
    for(int j = 0; j < kMaxK; ++j)
      if(j < k) acc[j] = fma(V[n + j * ldv], wn, acc[j]);
  }
  __shared__ double lds[4][kMaxK];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for(int j = 0; j < kMaxK; ++j) {
    if(j < k) {
      const double v = wave_sum(acc[j]);
      if(lane == 0) lds[wv][j] = v;
    }
  }
  __syncthreads();
  if(threadIdx.x < k)
    partials[(int64_t)blockIdx.x * kMaxK + threadIdx.x] =
        (lds[0][threadIdx.x] + lds[1][threadIdx.x]) + (lds[2][threadIdx.x] + lds[3][threadIdx.x]);
}

// out[i] = sum_b partials[b*stride + i]: one wavefront per output, lanes
// stride the blocks, fixed combination order (deterministic).
__global__ __launch_bounds__(64) void reduce_kernel(const double* __restrict__ partials,
                                                    int nblocks, int stride,
                                                    double* __restrict__ out) {
  const int i = blockIdx.x, lane = threadIdx.x;
  double s = 0.;
  for(int b = lane; b < nblocks; b += 64) s += partials[(int64_t)b * stride + i];
  s = wave_sum(s);
  if(lane == 0) out[i] = s;
}

// w = w - V(:, :k) * coef   (gram_schmidt gemm #2, davidson.hpp:116-117)
__global__ __launch_bounds__(256) void update_kernel(int64_t N, int k,
                                                      const double* __restrict__ V,
                                                      int64_t ldv,
                                                      const double* __restrict__ coef,
                                                      double* __restrict__ w) {
  __shared__ double c[kMaxK];
  if(threadIdx.x < k) c[threadIdx.x] = coef[threadIdx.x];
  __syncthreads();
  for(int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N;
      n += (int64_t)gridDim.x * 256) {
    double s = 0.;
#pragma unroll
    for(int j = 0; j < kMaxK; ++j)
      if(j < k) s = fma(V[n + j * ldv], c[j], s);
    w[n] = w[n] - s;
  }
}

struct RitzCoef {
  double c[kMaxK];
  double lam;
};

// X = V c; R = AV c - lam X; partial ||R||^2   (davidson.hpp:181-193)
__global__ __launch_bounds__(256) void residual_kernel(int64_t N, int k,
                                                        const double* __restrict__ V,
                                                        const double* __restrict__ AV,
                                                        int64_t ld, RitzCoef rc,
                                                        double* __restrict__ X,
                                                        double* __restrict__ R,
                                                        double* __restrict__ partials) {
  double acc = 0.;
  for(int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N;
      n += (int64_t)gridDim.x * 256) {
    double x = 0., r = 0.;
#pragma unroll
    for(int j = 0; j < kMaxK; ++j)
      if(j < k) {
        x = fma(V[n + j * ld], rc.c[j], x);
        r = fma(AV[n + j * ld], rc.c[j], r);
      }
    X[n] = x;
    const double rr = r + (-rc.lam) * x;
    R[n] = rr;
    acc = fma(rr, rr, acc);
  }
  __shared__ double lds[4];
  const double v = wave_sum(acc);
  if((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  if(threadIdx.x == 0) partials[blockIdx.x] = (lds[0] + lds[1]) + (lds[2] + lds[3]);
}

// Restart of the restarted variant, in place per row n:
//   [V(:,0..nb-1)]  = V(:, :k) Q,   [AV(:,0..nb-1)] = AV(:, :k) Q,
//   V(:,nb) = V(:,tcol), AV(:,nb) = AV(:,tcol)
// Each thread reads its row of V and AV before writing it, so the in-place
// update is race free (no other thread touches row n).
struct RestartCoef {
  double q[2][kMaxK];
};
__global__ __launch_bounds__(256) void restart_kernel(int64_t N, int k, int nb, RestartCoef Q,
                                                      double* V, double* AV, int64_t ld,
                                                      int64_t tcol) {
  for(int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N;
      n += (int64_t)gridDim.x * 256) {
    double v0 = 0., v1 = 0., a0 = 0., a1 = 0.;
#pragma unroll
    for(int j = 0; j < kMaxK; ++j)
      if(j < k) {
        const double x = V[n + j * ld], y = AV[n + j * ld];
        v0 = fma(x, Q.q[0][j], v0);
        a0 = fma(y, Q.q[0][j], a0);
        v1 = fma(x, Q.q[1][j], v1);
        a1 = fma(y, Q.q[1][j], a1);
      }
    const double t = V[n + tcol * ld], at = AV[n + tcol * ld];
    V[n] = v0;
    AV[n] = a0;
    if(nb > 1) {
      V[n + ld] = v1;
      AV[n + ld] = a1;
    }
    V[n + nb * ld] = t;
    AV[n + nb * ld] = at;
  }
}

// residual_kernel for subspaces wider than kMaxK: the Ritz coefficients
// c[0..k) come from device memory through LDS (k * 8 bytes of dynamic LDS).
__global__ __launch_bounds__(256) void residual_wide_kernel(int64_t N, int k,
                                                             const double* __restrict__ V,
                                                             const double* __restrict__ AV,
                                                             int64_t ld,
                                                             const double* __restrict__ c,
                                                             double lam, double* __restrict__ X,
                                                             double* __restrict__ R,
                                                             double* __restrict__ partials) {
  extern __shared__ double cs[];
  for(int j = threadIdx.x; j < k; j += 256) cs[j] = c[j];
  __syncthreads();
  double acc = 0.;
  for(int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N;
      n += (int64_t)gridDim.x * 256) {
    double x = 0., r = 0.;
    for(int j = 0; j < k; ++j) {
      x = fma(V[n + j * ld], cs[j], x);
      r = fma(AV[n + j * ld], cs[j], r);
    }
    X[n] = x;
    const double rr = r + (-lam) * x;
    R[n] = rr;
    acc = fma(rr, rr, acc);
  }
  __shared__ double lds[4];
  const double v = wave_sum(acc);
  if((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  if(threadIdx.x == 0) partials[blockIdx.x] = (lds[0] + lds[1]) + (lds[2] + lds[3]);
}

// restart_kernel for subspaces wider than kMaxK: q[0..k) and q[k..2k) (the
// two restart coefficient vectors) from device memory through LDS.
__global__ __launch_bounds__(256) void restart_wide_kernel(int64_t N, int k, int nb,
                                                            const double* __restrict__ q,
                                                            double* V, double* AV, int64_t ld,
                                                            int64_t tcol) {
  extern __shared__ double qs[];
  for(int j = threadIdx.x; j < 2 * k; j += 256) qs[j] = q[j];
  __syncthreads();
  for(int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N;
      n += (int64_t)gridDim.x * 256) {
    double v0 = 0., v1 = 0., a0 = 0., a1 = 0.;
    for(int j = 0; j < k; ++j) {
      const double x = V[n + j * ld], y = AV[n + j * ld];
      v0 = fma(x, qs[j], v0);
      a0 = fma(y, qs[j], a0);
      v1 = fma(x, qs[k + j], v1);
      a1 = fma(y, qs[k + j], a1);
    }
    const double t = V[n + tcol * ld], at = AV[n + tcol * ld];
    V[n] = v0;
    AV[n] = a0;
    if(nb > 1) {
      V[n + ld] = v1;
      AV[n + ld] = a1;
    }
    V[n + nb * ld] = t;
    AV[n + nb * ld] = at;
  }
}

// R = -R / (D - lam)   (davidson.hpp:212-214)
__global__ void precond_kernel(int64_t N, double* __restrict__ R,
                               const double* __restrict__ D, double lam) {
  for(int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N;
      n += (int64_t)gridDim.x * 256)
    R[n] = -R[n] / (D[n] - lam);
}

// w *= 1 / sqrt(nrm2[0])   (nrm2 + scal, davidson.hpp:124-125)
__global__ void scale_kernel(int64_t N, double* __restrict__ w,
                             const double* __restrict__ nrm2) {
  const double s = 1. / sqrt(nrm2[0]);
  for(int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N;
      n += (int64_t)gridDim.x * 256)
    w[n] *= s;
}

// Per-block first-index argmin of D and max |C|.
__global__ __launch_bounds__(256) void argmin_maxabs_kernel(
    int64_t N, const double* __restrict__ D, const double* __restrict__ C,
    double* __restrict__ pmin, int64_t* __restrict__ pidx,
    double* __restrict__ pmax) {
  double mn = INFINITY, mx = 0.;
  int64_t mi = INT64_MAX;
  for(int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N;
      n += (int64_t)gridDim.x * 256) {
    const double d = D[n];
    if(d < mn || (d == mn && n < mi)) {
      mn = d;
      mi = n;
    }
    if(C) mx = fmax(mx, fabs(C[n]));
  }
  __shared__ double smn[256], smx[256];
  __shared__ int64_t smi[256];
  smn[threadIdx.x] = mn;
  smi[threadIdx.x] = mi;
  smx[threadIdx.x] = mx;
  __syncthreads();
  for(int off = 128; off > 0; off >>= 1) {
    if(threadIdx.x < off) {
      const int o = threadIdx.x + off;
      if(smn[o] < smn[threadIdx.x] || (smn[o] == smn[threadIdx.x] && smi[o] < smi[threadIdx.x])) {
        smn[threadIdx.x] = smn[o];
        smi[threadIdx.x] = smi[o];
      }
      smx[threadIdx.x] = fmax(smx[threadIdx.x], smx[o]);
    }
    __syncthreads();
  }
  if(threadIdx.x == 0) {
    pmin[blockIdx.x] = smn[0];
    pidx[blockIdx.x] = smi[0];
    pmax[blockIdx.x] = smx[0];
  }
}

__global__ void set_unit_kernel(double* X, int64_t i) { X[i] = 1.; }

// ---- host-side helpers -----------------------------------------------------
struct Ops {
  mhip_context* ctx;
  int64_t N;
  int KS;        // stride of the small arrays: max(kMaxK, widest subspace)
  DevBuf part;   // reduction partials
  DevBuf small;  // reduced results (k inner products, norms, C)
  DevBuf wide;   // Ritz / restart coefficients of subspaces wider than kMaxK
  Ops(mhip_context* c, int64_t n, int kmax) : ctx(c), N(n), KS(std::max(kMaxK, kmax)) {
    part.alloc(sizeof(double) * std::max(kRedBlocks * kMaxK, kGramBlocks * 4 * 256));
    small.alloc(sizeof(double) * ((size_t)KS * KS + 4 * (size_t)KS));
    if(KS > kMaxK) wide.alloc(sizeof(double) * 2 * (size_t)KS);
  }
  hipStream_t st() const { return ctx->stream; }
  double* inner() { return small.as<double>(); }
  double* scalar() { return small.as<double>() + KS; }
  double* rn2() { return small.as<double>() + 2 * KS; }  // residual norm^2, then C
  double* Cdev() { return small.as<double>() + 4 * KS; }

  // out[j] = V(:,j) . w, j < k, in column blocks of kMaxK
  void dots(const double* V, int64_t ldv, int k, const double* w, double* out) {
    const int nb = red_blocks(N);
    for(int j0 = 0; j0 < k; j0 += kMaxK) {
      const int kc = std::min(kMaxK, k - j0);
      hipLaunchKernelGGL(dots_kernel, dim3(nb), dim3(256), 0, st(), N, kc, V + j0 * ldv, ldv, w,
                         part.as<double>());
      hipLaunchKernelGGL(reduce_kernel, dim3(kc), dim3(64), 0, st(), part.as<double>(), nb,
                         kMaxK, out + j0);
    }
  }
  // w -= V(:, :k) coef, in column blocks of kMaxK
  void update(const double* V, int64_t ldv, int k, const double* coef, double* w) {
    for(int j0 = 0; j0 < k; j0 += kMaxK)
      hipLaunchKernelGGL(update_kernel, dim3(red_blocks(N)), dim3(256), 0, st(), N,
                         std::min(kMaxK, k - j0), V + j0 * ldv, ldv, coef + j0, w);
  }
  // gram_schmidt (davidson.hpp:111-126): CGS2 against V(:, :k), normalise.
  void gram_schmidt(const double* V, int k, double* w) {
    for(int pass = 0; pass < 2; ++pass) {
      dots(V, N, k, w, inner());
      update(V, N, k, inner(), w);
    }
    normalize(w);
  }
  void normalize(double* w) {
    dots(w, N, 1, w, scalar());
    hipLaunchKernelGGL(scale_kernel, dim3(red_blocks(N)), dim3(256), 0, st(), N, w, scalar());
  }
  // C (k x k, ldc = k) = V^T AV on MFMA into Cdev() (not waited for); wider
  // subspaces in 32 x 32 tiles.
  void gram(const double* V, const double* AV, int k) {
    const int nb = gram_blocks(N);
    if(k <= 16) {
      hipLaunchKernelGGL((gram_mfma_kernel<1, 1>), dim3(nb), dim3(256), 0, st(), N, k, V, N, k,
                         AV, N, part.as<double>());
      hipLaunchKernelGGL((gram_finalize_kernel<1, 1>), dim3(4), dim3(256), 0, st(),
                         part.as<double>(), nb, k, k, Cdev(), k);
    } else {
      for(int i0 = 0; i0 < k; i0 += kMaxK)
        for(int j0 = 0; j0 < k; j0 += kMaxK) {
          const int ki = std::min(kMaxK, k - i0), kj = std::min(kMaxK, k - j0);
          hipLaunchKernelGGL((gram_mfma_kernel<2, 2>), dim3(nb), dim3(256), 0, st(), N, ki,
                             V + i0 * N, N, kj, AV + j0 * N, N, part.as<double>());
          hipLaunchKernelGGL((gram_finalize_kernel<2, 2>), dim3(16), dim3(256), 0, st(),
                             part.as<double>(), nb, ki, kj, Cdev() + i0 + (int64_t)j0 * k, k);
        }
    }
    MHIP_HIP_CHECK(hipGetLastError());
  }
  // X = V c; R = AV c - lam X; ||R||^2 into rn2() (not waited for).
  void residual(const double* V, const double* AV, int k, const double* c, double lam, double* X,
                double* R) {
    const int nb = red_blocks(N);
    if(k <= kMaxK) {
      RitzCoef rc;
      for(int j = 0; j < kMaxK; ++j) rc.c[j] = j < k ? c[j] : 0.;
      rc.lam = lam;
      hipLaunchKernelGGL(residual_kernel, dim3(nb), dim3(256), 0, st(), N, k, V, AV, N, rc, X, R,
                         part.as<double>());
    } else {
      MHIP_HIP_CHECK(hipMemcpyAsync(wide.p, c, sizeof(double) * k, hipMemcpyHostToDevice, st()));
      hipLaunchKernelGGL(residual_wide_kernel, dim3(nb), dim3(256), sizeof(double) * k, st(), N,
                         k, V, AV, N, wide.as<double>(), lam, X, R, part.as<double>());
    }
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(64), 0, st(), part.as<double>(), nb, 1,
                       rn2());
    MHIP_HIP_CHECK(hipGetLastError());
  }
  // In-place restart [V Q, t] (restart_kernel) with q0 / q1 of length k.
  void restart(int k, int nb, const double* q0, const double* q1, double* V, double* AV,
               int64_t tcol) {
    if(k <= kMaxK) {
      RestartCoef Q{};
      for(int j = 0; j < k; ++j) {
        Q.q[0][j] = q0[j];
        Q.q[1][j] = q1[j];
      }
      hipLaunchKernelGGL(restart_kernel, dim3(red_blocks(N)), dim3(256), 0, st(), N, k, nb, Q,
                         V, AV, N, tcol);
    } else {
      std::vector<double> q(2 * (size_t)k);
      std::copy_n(q0, k, q.begin());
      std::copy_n(q1, k, q.begin() + k);
      MHIP_HIP_CHECK(hipMemcpyAsync(wide.p, q.data(), sizeof(double) * 2 * k,
                                    hipMemcpyHostToDevice, st()));
      // the host vector must outlive the copy
      MHIP_HIP_CHECK(hipStreamSynchronize(st()));
      hipLaunchKernelGGL(restart_wide_kernel, dim3(red_blocks(N)), dim3(256),
                         sizeof(double) * 2 * k, st(), N, k, nb, wide.as<double>(), V, AV, N,
                         tcol);
    }
    MHIP_HIP_CHECK(hipGetLastError());
  }
  // One wait for the stream: the k x k Gram matrix and the last ||R||^2.
  void read_gram(int k, double* C_host, double* r2) {
    std::vector<double> buf(2 * (size_t)KS + (size_t)k * k);
    d2h(ctx, buf.data(), rn2(), sizeof(double) * buf.size());
    *r2 = buf[0];
    std::memcpy(C_host, buf.data() + 2 * KS, sizeof(double) * k * k);
  }
  double read_r2() {
    double r2 = 0.;
    d2h(ctx, &r2, rn2(), 8);
    return r2;
  }
};

// Start/stop event pairs of the SpMVs of one solve, read after the stream
// has been waited for (no wait per product).
struct EventPairs {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending, idle;
  ~EventPairs() {
    for(auto* v : {&pending, &idle})
      for(auto& e : *v) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
      }
  }
  std::pair<hipEvent_t, hipEvent_t> get() {
    std::pair<hipEvent_t, hipEvent_t> e{nullptr, nullptr};
    if(!idle.empty()) {
      e = idle.back();
      idle.pop_back();
    } else {
      MHIP_HIP_CHECK(hipEventCreate(&e.first));
      if(hipEventCreate(&e.second) != hipSuccess) {
        (void)hipEventDestroy(e.first);
        throw Error(MHIP_ERR_HIP, "hipEventCreate failed");
      }
    }
    pending.push_back(e);
    return e;
  }
  // sum of the elapsed times of the pending pairs (all complete)
  double drain() {
    double ms = 0.;
    for(auto& e : pending) {
      float t = 0.f;
      MHIP_HIP_CHECK(hipEventElapsedTime(&t, e.first, e.second));
      ms += t;
      idle.push_back(e);
    }
    pending.clear();
    return ms;
  }
};

}  // namespace

// Cyclic Jacobi on the lower triangle (Golub & Van Loan, Alg. 8.5.3);
// eigenvalues ascending, eigenvectors in the columns of A (replaces
// lapack::syev(Vec, Lower), rayleigh_ritz.hpp:75).
void syev_lower(int n, double* A, int lda, double* W) {
  std::vector<double> a(n * n), v(n * n, 0.);
  for(int j = 0; j < n; ++j)
    for(int i = j; i < n; ++i) a[i + n * j] = a[j + n * i] = A[i + lda * j];
  for(int i = 0; i < n; ++i) v[i + n * i] = 1.;
  for(int sweep = 0; sweep < 100; ++sweep) {
    double off = 0., tot = 0.;
    for(int j = 0; j < n; ++j)
      for(int i = 0; i < n; ++i) {
        tot += a[i + n * j] * a[i + n * j];
        if(i != j) off += a[i + n * j] * a[i + n * j];
      }
    if(off <= 1e-32 * tot || off == 0.) break;
    for(int p = 0; p < n - 1; ++p)
      for(int q = p + 1; q < n; ++q) {
        const double apq = a[p + n * q];
        if(apq == 0.) continue;
        // after four sweeps an element below the diagonals' rounding is set
        // to zero instead of rotated (Numerical Recipes, jacobi): without
        // it the off-diagonal floor of large-magnitude matrices never meets
        // the relative test and every call runs the sweep cap
        const double g = 100. * std::fabs(apq), app = std::fabs(a[p + n * p]),
                     aqq = std::fabs(a[q + n * q]);
        if(sweep > 3 && app + g == app && aqq + g == aqq) {
          a[p + n * q] = a[q + n * p] = 0.;
          continue;
        }
        const double tau = (a[q + n * q] - a[p + n * p]) / (2. * apq);
        const double t = (tau >= 0. ? 1. : -1.) / (std::fabs(tau) + std::sqrt(1. + tau * tau));
        const double c = 1. / std::sqrt(1. + t * t), s = t * c;
        for(int r = 0; r < n; ++r) {
          const double akp = a[r + n * p], akq = a[r + n * q];
          a[r + n * p] = c * akp - s * akq;
          a[r + n * q] = s * akp + c * akq;
        }
        for(int r = 0; r < n; ++r) {
          const double apk = a[p + n * r], aqk = a[q + n * r];
          a[p + n * r] = c * apk - s * aqk;
          a[q + n * r] = s * apk + c * aqk;
        }
        for(int r = 0; r < n; ++r) {
          const double vkp = v[r + n * p], vkq = v[r + n * q];
          v[r + n * p] = c * vkp - s * vkq;
          v[r + n * q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> ord(n);
  for(int i = 0; i < n; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(),
                   [&](int x, int y) { return a[x * (n + 1)] < a[y * (n + 1)]; });
  for(int j = 0; j < n; ++j) {
    W[j] = a[ord[j] * (n + 1)];
    for(int i = 0; i < n; ++i) A[i + lda * j] = v[i + n * ord[j]];
  }
}

// Restart coefficients (k x nb): q0 = c/|c|; for nb == 2, q1 = the previous
// Ritz coefficients (c_prev, 0) orthogonalised against q0 by two classical
// Gram-Schmidt passes and normalised, or the second Ritz vector's
// coefficients c2 treated the same way when c_prev is absent (the last two
// iterations did not share the basis) or its part orthogonal to q0 has norm
// <= 1e-10.
void restart_coefficients(int k, int nb, const double* c, const double* cprev, const double* c2,
                          double* q0, double* q1) {
  double cn = 0.;
  for(int j = 0; j < k; ++j) cn += c[j] * c[j];
  cn = 1. / std::sqrt(cn);
  for(int j = 0; j < k; ++j) q0[j] = c[j] * cn;
  if(nb < 2) return;
  for(int attempt = cprev ? 0 : 1; attempt < 2; ++attempt) {
    for(int j = 0; j < k; ++j) q1[j] = attempt ? c2[j] : (j < k - 1 ? cprev[j] : 0.);
    for(int pass = 0; pass < 2; ++pass) {
      double s = 0.;
      for(int j = 0; j < k; ++j) s += q0[j] * q1[j];
      for(int j = 0; j < k; ++j) q1[j] -= s * q0[j];
    }
    double pn = 0.;
    for(int j = 0; j < k; ++j) pn += q1[j] * q1[j];
    pn = std::sqrt(pn);
    if(pn > 1e-10) {
      for(int j = 0; j < k; ++j) q1[j] /= pn;
      return;
    }
  }
}

bool davidson(mhip_context* ctx, const mhip_csr* A, int64_t max_m,
              int64_t max_restarts, double tol, const double* d_D, double* d_X,
              double* E, int64_t* iters) {
  MHIP_REQUIRE(d_X, MHIP_ERR_NO_GUESS, "Davidson: No Guess Provided");
  MHIP_REQUIRE(A->m == A->n, MHIP_ERR_INVALID_ARG, "Davidson needs a square matrix");
  const int64_t N = A->m;
  mhip_timings& tm = ctx->timings;
  tm.dav_spmv_ms = 0.;
  tm.dav_spmv_calls = 0;
  tm.dav_iters = 0;
  tm.dav_restarts = 0;
  tm.dav_residual = 0.;
  tm.dav_layout_ms = 0.;
  tm.dav_sliced = 0;
  if(N == 0) {
    *E = 0.;
    if(iters) *iters = 0;
    return false;
  }
  max_m = std::min(max_m, N);  // davidson.hpp:140
  MHIP_REQUIRE(max_m >= 1 && max_m <= kMaxSubspace, MHIP_ERR_INVALID_ARG,
               "max_m must be in [1, 1024]");
  hipStream_t st = ctx->stream;
  Timer t_all(st);
  t_all.start();
  const int kmax = (int)max_m + 1;
  Ops ops(ctx, N, kmax);
  const int64_t ld = N;
  // subspace storage comes from (and goes back to) the context's kept
  // buffers: solves of the same size reuse it without remapping memory
  DevBuf Vb = take_buffer(ctx, sizeof(double) * N * (max_m + 1));
  DevBuf AVb = take_buffer(ctx, sizeof(double) * N * (max_m + 1));
  double* V = Vb.as<double>();
  double* AV = AVb.as<double>();
  DevBuf Dbuf;
  if(!d_D) {
    Dbuf.alloc(sizeof(double) * N);
    extract_diagonal(ctx, A, Dbuf.as<double>());
    d_D = Dbuf.as<double>();
  }
  // sliced copy of A for the SpMVs when there is room for it (made once per
  // matrix handle; the CSR kernel runs otherwise)
  {
    const bool had = A->sell || A->panels;
    const auto h0 = std::chrono::steady_clock::now();
    // (the x-tile copy is used when a caller attached it; on the
    // 2,000,000-determinant selected-CI matrix it measured 24.9 ms per
    // product against 14.5 ms for the 8-panel split)
    tm.dav_sliced = A->xtiles                ? 3
                    : prepare_sell(ctx, A)   ? 1
                    : prepare_panels(ctx, A) ? 2
                                             : 0;
    if(!had && tm.dav_sliced) MHIP_HIP_CHECK(hipStreamSynchronize(st));
    tm.dav_spmv_matrix_bytes =
        A->sell     ? A->sell->streamed_bytes()
        : A->xtiles ? 10. * (double)A->nnz + 4. * (double)A->xtiles->np * (double)(A->m + 1)
                    : 12. * (double)A->nnz + 8. * (double)(A->m + 1) +
                          (A->panels ? 4. * (double)A->m * A->panels->np : 0.);
    tm.dav_layout_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
  }
  EventPairs ev;
  auto op = [&](const double* x, double* y) {
    const auto e = ev.get();
    MHIP_HIP_CHECK(hipEventRecord(e.first, st));
    spmv(ctx, A, 1., x, 0., y);
    MHIP_HIP_CHECK(hipEventRecord(e.second, st));
    tm.dav_spmv_calls++;
  };
  MHIP_HIP_CHECK(hipMemsetAsync(AV, 0, sizeof(double) * N * (max_m + 1), st));
  MHIP_HIP_CHECK(hipMemcpyAsync(V, d_X, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
  op(V, AV);                                                               // :153
  MHIP_HIP_CHECK(hipMemcpyAsync(V + ld, AV, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
  ops.gram_schmidt(V, 1, V + ld);                                           // :157

  // Iteration i (k = i + 1 columns) as davidson.hpp:164-218. The host waits
  // once per iteration, for the Gram matrix: the residual norm of iteration
  // i is read together with the Gram matrix of iteration i + 1, so the
  // correction, its product and the next Gram matrix are queued before the
  // convergence test. When that test passes, the queued work (one product,
  // possibly a restart) is discarded: X and the energy are those of the
  // converged iteration, exactly as if the loop had stopped there.
  bool converged = false;
  int64_t iter = 1, restarts = 0, i0 = 1, i = 1;
  std::vector<double> Cv((size_t)kmax * kmax), LAMv(kmax), ccurv(kmax, 0.), cprevv(kmax, 0.),
      c2curv(kmax, 0.), q0(kmax, 0.), q1(kmax, 0.);
  double *C = Cv.data(), *LAM = LAMv.data(), *ccur = ccurv.data(), *cprev = cprevv.data(),
         *c2cur = c2curv.data();
  bool prev_valid = false, pending = false, pending_restart = false;
  int kcur = 0;
  double lam0 = 0., rnorm = 0.;
  if(max_m >= 2) op(V + ld, AV + ld);                                       // :166
  while(max_m >= 2) {
    const int k = (int)(i + 1);
    ops.gram(V, AV, k);                                                     // :172
    double r2 = 0.;
    ops.read_gram(k, C, &r2);
    tm.dav_spmv_ms += ev.drain();
    if(pending) {  // the previous iteration's residual norm (:193, :205)
      rnorm = std::sqrt(r2);
      if(!std::isfinite(rnorm)) break;
      if(rnorm < tol) {
        converged = true;
        if(pending_restart) --restarts;
        break;
      }
      ++iter;
    }
    syev_lower(k, C, k, LAM);
    lam0 = LAM[0];
    std::copy(ccurv.begin(), ccurv.end(), cprevv.begin());
    prev_valid = i > i0;  // cprev holds coefficients in this basis
    for(int j = 0; j < k; ++j) {
      ccur[j] = C[j];
      c2cur[j] = k > 1 ? C[j + k] : 0.;
    }
    kcur = k;
    double* R = V + (i + 1) * ld;
    ops.residual(V, AV, k, C, lam0, d_X, R);                                // :181-193
    pending = true;
    pending_restart = false;
    if(i + 1 < max_m || restarts < max_restarts) {
      hipLaunchKernelGGL(precond_kernel, dim3(red_blocks(N)), dim3(256), 0, st, N, R, d_D,
                         lam0);                                             // :212
      ops.gram_schmidt(V, k, R);                                            // :217
      op(R, AV + (i + 1) * ld);  // the next iteration's product (:166)
    }
    if(i + 1 < max_m) {
      ++i;
      continue;
    }
    if(restarts >= max_restarts) {  // subspace exhausted: this iteration decides
      rnorm = std::sqrt(ops.read_r2());
      converged = std::isfinite(rnorm) && rnorm < tol;
      if(!converged && std::isfinite(rnorm)) ++iter;
      break;
    }
    // Restart (labelled variant): basis [V c, V q1, t] with q1 from the
    // previous Ritz vector; AV follows by the same combinations (no SpMV).
    // t = V(:,max_m) is orthogonal to span(V) already; its product A t was
    // just formed and moves with it.
    ++restarts;
    pending_restart = true;
    const int nb = max_m >= 3 ? 2 : 1;
    std::fill(q0.begin(), q0.end(), 0.);
    std::fill(q1.begin(), q1.end(), 0.);
    restart_coefficients(kcur, nb, ccur, prev_valid ? cprev : nullptr, c2cur, q0.data(),
                         q1.data());
    ops.restart(kcur, nb, q0.data(), q1.data(), V, AV, max_m);
    i0 = i = nb;
  }
  MHIP_HIP_CHECK(hipStreamSynchronize(st));
  tm.dav_spmv_ms += ev.drain();
  t_all.stop();
  tm.dav_total_ms = t_all.ms();
  keep_buffer(ctx, std::move(Vb));
  keep_buffer(ctx, std::move(AVb));
  tm.dav_iters = iter;
  tm.dav_restarts = restarts;
  tm.dav_residual = rnorm;
  *E = lam0;
  if(iters) *iters = iter;
  return converged;
}

// Guess rule of serial_selected_ci_diag (selected_ci_diag.hpp:94-104).
void guess_rule(mhip_context* ctx, int64_t N, const double* d_D, double* d_C) {
  const int nb = red_blocks(N);
  DevBuf p(sizeof(double) * nb * 3);
  double* pmin = p.as<double>();
  int64_t* pidx = reinterpret_cast<int64_t*>(pmin + nb);
  double* pmax = pmin + 2 * nb;
  hipLaunchKernelGGL(argmin_maxabs_kernel, dim3(nb), dim3(256), 0, ctx->stream, N, d_D, d_C,
                     pmin, pidx, pmax);
  std::vector<double> h(3 * nb);
  d2h(ctx, h.data(), p.p, sizeof(double) * 3 * nb);
  double mn = INFINITY, mx = 0.;
  int64_t mi = 0;
  for(int b = 0; b < nb; ++b) {
    int64_t idx;
    std::memcpy(&idx, &h[nb + b], 8);
    if(h[b] < mn || (h[b] == mn && idx < mi)) {
      mn = h[b];
      mi = idx;
    }
    mx = std::max(mx, h[2 * nb + b]);
  }
  if(!(mx > 1. / (double)N))  // diagonal_guess (davidson.hpp:61-68)
    hipLaunchKernelGGL(set_unit_kernel, dim3(1), dim3(1), 0, ctx->stream, d_C, mi);
  MHIP_HIP_CHECK(hipGetLastError());
}

// Tall-skinny GEMM entry used by tests and the C++ layer: C = A^T B.
// C = A^T B through block partials in `part` (gram_blocks(N) * 1024
// doubles), waited for.
void gemm_tn(mhip_context* ctx, int64_t N, int k, const double* A, int64_t lda, int m,
             const double* B, int64_t ldb, double* d_C, int ldc, double* part) {
  const int nb = gram_blocks(N);
  hipLaunchKernelGGL((gram_mfma_kernel<2, 2>), dim3(nb), dim3(256), 0, ctx->stream, N, k, A,
                     lda, m, B, ldb, part);
  hipLaunchKernelGGL((gram_finalize_kernel<2, 2>), dim3(16), dim3(256), 0, ctx->stream,
                     part, nb, k, m, d_C, ldc);
  MHIP_HIP_CHECK(hipGetLastError());
  MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

void check_gemm_tn_args(int64_t N, int k, const double* A, int64_t lda, int m, const double* B,
                        int64_t ldb, const double* C, int ldc) {
  MHIP_REQUIRE(A && B && C && lda >= N && ldb >= N && ldc >= k,
               MHIP_ERR_INVALID_ARG, "bad arguments");
  MHIP_REQUIRE(k >= 1 && k <= 32 && m >= 1 && m <= 32, MHIP_ERR_INVALID_ARG,
               "k, m must be in [1, 32]");
}

}  // namespace mhip

using namespace mhip;

extern "C" {

int mhip_davidson(mhip_context* ctx, const mhip_csr* A, int64_t max_m,
                  int64_t max_restarts, double tol, const double* d_D,
                  double* d_X, double* E, int64_t* iters) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(A && E, MHIP_ERR_INVALID_ARG, "null argument");
    MHIP_REQUIRE(max_restarts >= 0, MHIP_ERR_INVALID_ARG, "negative max_restarts");
    if(!davidson(ctx, A, max_m, max_restarts, tol, d_D, d_X, E, iters))
      throw Error(MHIP_ERR_NOT_CONVERGED, "Davidson Did Not Converge!");
  });
}

int mhip_selected_ci_diag(mhip_context* ctx, int words_per_det,
                          const uint64_t* d_dets, int64_t ndets,
                          double h_el_tol, int64_t max_m, int64_t max_restarts,
                          double res_tol, double* d_C, double* E,
                          int64_t* nnz, mhip_csr** H_out) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(E && (ndets == 0 || d_C), MHIP_ERR_INVALID_ARG, "null argument");
    if(H_out) *H_out = nullptr;
    std::unique_ptr<mhip_csr> H(ham_build(ctx, words_per_det, d_dets, ndets, d_dets, ndets,
                                          h_el_tol));
    if(nnz) *nnz = H->nnz;
    DevBuf D(sizeof(double) * std::max<int64_t>(ndets, 1));
    extract_diagonal(ctx, H.get(), D.as<double>());
    if(ndets) guess_rule(ctx, ndets, D.as<double>(), d_C);
    bool ok = davidson(ctx, H.get(), max_m, max_restarts, res_tol, D.as<double>(), d_C, E,
                       nullptr);
    if(H_out) {
      *H_out = H.release();
    } else {  // keep the storage for the next build on this context
      keep_buffer(ctx, std::move(H->rowptr));
      keep_buffer(ctx, std::move(H->colind));
      keep_buffer(ctx, std::move(H->nzval));
      if(H->sell) {
        keep_buffer(ctx, std::move(H->sell->cols));
        keep_buffer(ctx, std::move(H->sell->vals));
        keep_buffer(ctx, std::move(H->sell->base));
      }
      if(H->xtiles) {
        keep_buffer(ctx, std::move(H->xtiles->cols));
        keep_buffer(ctx, std::move(H->xtiles->vals));
      }
    }
    if(!ok) throw Error(MHIP_ERR_NOT_CONVERGED, "Davidson Did Not Converge!");
  });
}

int mhip_gemm_tn(mhip_context* ctx, int64_t N, int k, const double* d_A,
                 int64_t lda, int m, const double* d_B, int64_t ldb, double* d_C,
                 int ldc) {
  return guarded([&] {
    use_device(ctx);
    check_gemm_tn_args(N, k, d_A, lda, m, d_B, ldb, d_C, ldc);
    DevBuf part(sizeof(double) * gram_blocks(N) * 4 * 256);
    gemm_tn(ctx, N, k, d_A, lda, m, d_B, ldb, d_C, ldc, part.as<double>());
  });
}

// The same product with its block partials in the device's static scratch
// array: no allocation. Calls that overlap on one GPU share that array and
// race.
int mhip_gemm_tn_noalloc(mhip_context* ctx, int64_t N, int k, const double* d_A, int64_t lda,
                         int m, const double* d_B, int64_t ldb, double* d_C, int ldc) {
  return guarded([&] {
    use_device(ctx);
    check_gemm_tn_args(N, k, d_A, lda, m, d_B, ldb, d_C, ldc);
    double* part = static_cast<double*>(noalloc_scratch(ctx->device));
    MHIP_REQUIRE(part, MHIP_ERR_HIP, "device scratch not set up");
    const size_t need = sizeof(double) * (size_t)gram_blocks(N) * 4 * 256;
    MHIP_REQUIRE(need <= kNoAllocScratchBytes, MHIP_ERR_INVALID_ARG,
                 "partials exceed the 2 MiB device scratch");
    gemm_tn(ctx, N, k, d_A, lda, m, d_B, ldb, d_C, ldc, part);
  });
}

}  // extern "C"
