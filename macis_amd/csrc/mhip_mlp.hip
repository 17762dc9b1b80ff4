// Learned ASCI pre-screen: an MLP 128 -> 512 -> 512 -> 1 (GELU) on the
// occupation bits of a determinant, predicting log|rv| (no counterpart in
// the reference; the plain search of determinant_search.hpp:394-663 stays
// the arbiter of what is kept and the fallback).
//
// Compute: BF16 operands on MFMA (v_mfma_f32_32x32x16_bf16, gfx950), FP32
// accumulation, FP32 master weights updated with Adam. One tiled GEMM
// kernel serves the forward pass (C = A B^T, both K-contiguous), and the
// backward products (A B with B N-contiguous, A^T B over the batch): tiles
// are staged through LDS in K-contiguous layout whatever the global layout.
// Epilogues fuse the bias, the GELU and the GELU derivative.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "mhip_device.hpp"
#include "mhip_internal.hpp"

namespace mhip {
namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr int kIn = 128, kHid = 512;
constexpr int BN = 128, BK = 32, LDK = BK + 8;
constexpr int kStageLd = 64 + 8;  // bf16 row pitch of the epilogue staging (16-byte aligned rows)

// GELU, tanh form: 0.5 z (1 + tanh u) with u = sqrt(2/pi) (z + 0.044715 z^3).
// Since 0.5 (1 + tanh u) = 1 / (1 + e^(-2u)) = s, gelu(z) = z s and
// gelu'(z) = s + z s (1 - s) 2 du/dz. s is formed from the hardware
// exp2 and reciprocal (a libm tanhf per element made the epilogues of the
// 2^20-row layers VALU bound); e^(-2u) -> inf gives s = 0, as it should.
__device__ __forceinline__ float gelu_s(float z) {
  constexpr float k2u = 2.f * 0.7978845608028654f * 1.4426950408889634f;  // 2 sqrt(2/pi) log2(e)
  // the library builds without FP contraction; these FMAs are explicit
  const float e = __builtin_amdgcn_exp2f(__builtin_fmaf(z * z, -k2u * 0.044715f, -k2u) * z);
  return __builtin_amdgcn_rcpf(1.f + e);
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_s(z); }
__device__ __forceinline__ float gelu_grad(float z) {
  constexpr float c = 2.f * 0.7978845608028654f;
  const float sg = gelu_s(z);
  const float w = __builtin_fmaf(-sg, sg, sg);                    // s (1 - s)
  const float du = __builtin_fmaf(z * z, 3.f * 0.044715f * c, c);  // 2 du/dz
  return __builtin_fmaf(z * w, du, sg);
}

enum Epi : int {
  EPI_F32 = 0,        // C (fp32)
  EPI_BIAS_GELU = 1,  // Z = C + bias (fp32), A = gelu(Z) (bf16)
  EPI_GELU_GRAD = 2,  // D = C * gelu'(Zaux) (bf16)
  EPI_BIAS_GELU_A = 3,  // A = gelu(C + bias) (bf16) only: inference
  // inference head folded into the last layer: zaux = w3, Cf = partial sums
  // part[2 * n_block + wave_column][row] = sum over the wave's 64 columns of
  // bf16(gelu(C + bias)) * w3 (ldc = row stride of part)
  EPI_HEAD = 4,
};

// C[m][n] = sum_k A(m,k) B(n,k); A(m,k) = AK ? A[m*lda+k] : A[k*lda+m],
// B(n,k) = BK ? B[n*ldb+k] : B[k*ldb+n]. Four waves in a 2 x 2 layout, each
// owning TI x 2 MFMA tiles of 32 x 32: block tile (64 TI) x 128, so M is a
// multiple of 64 TI, N of 128, K of 32. TI = 4 reads 3/4 of the LDS operand
// bytes per MFMA that TI = 2 does, at twice the accumulator registers.
// FW > 0 (with AK): A is not read from memory but formed from determinants
// of FW uint64 words (passed as A; lda = number of valid rows): A(m, k) =
// bit k of determinant m as bf16 0/1 (zero past 64 FW bits and for rows m >=
// lda) -- the featurisation folded into the first layer's loads.
template <bool AK, bool BKC, int EPI, int TI = 2, int FW = 0>
__global__ __launch_bounds__(256, TI == 4 ? 2 : 1) void gemm_bf16_kernel(
    int M, int N, int K, const __bf16* __restrict__ A, int lda, const __bf16* __restrict__ B,
    int ldb, const float* __restrict__ bias, const float* __restrict__ zaux,
    float* __restrict__ Cf, __bf16* __restrict__ Cb, int ldc) {
  // two buffers of A and B tiles (the next K step is staged while this one
  // feeds the MFMAs); after the K loop the same LDS stages bf16 outputs
  constexpr int BMt = 64 * TI;
  __shared__ __attribute__((aligned(16))) __bf16 smem[2 * (BMt + BN) * LDK];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // workgroups reach the 8 XCDs round-robin by linear id; hand each XCD a
  // contiguous run of tiles so that the N / BN tiles sharing one A row tile
  // run on the same XCD (one L2) at about the same time
  int bx = blockIdx.x, by = blockIdx.y;
  {
    const int nbx = gridDim.x, nb = nbx * gridDim.y;
    if((nb & 7) == 0) {
      const int b = bx + nbx * by, l = (b & 7) * (nb >> 3) + (b >> 3);
      bx = l % nbx;
      by = l / nbx;
    }
  }
  const int m0 = by * BMt, n0 = bx * BN;
  const int wm = (wv >> 1) * 32 * TI, wn = (wv & 1) * 64;
  // split K over gridDim.z (fp32 output only): slice z owns
  // [z*K/Z, (z+1)*K/Z) and writes its partial product at Cf + z*M*ldc
  const int kz = K / (int)gridDim.z, kbeg = (int)blockIdx.z * kz, kend = kbeg + kz;
  if constexpr(EPI == EPI_F32) Cf += (size_t)blockIdx.z * M * ldc;
  // lane l, register r of acc[i][j] holds C(row, col) with row = wm + 32 i +
  // l%32 and col = wn + 32 j + 8 (r/4) + 4 (l/32) + r%4 (the MFMAs take the
  // B tile as their first operand): each group of four registers is four
  // consecutive columns of one row. Epilogues with a bias start from it,
  // except the TI = 4 head, where the early bias loads cost register spills
  // and the epilogue adds it instead.
  auto col_of = [&](int j, int q) { return n0 + wn + j * 32 + 8 * q + 4 * (lane >> 5); };
  constexpr bool kBias = EPI == EPI_BIAS_GELU || EPI == EPI_BIAS_GELU_A;
  f32x16_t acc[TI][2];
#pragma unroll
  for(int j = 0; j < 2; ++j)
#pragma unroll
    for(int q = 0; q < 4; ++q) {
      float4 bv = {0.f, 0.f, 0.f, 0.f};
      if constexpr(kBias) bv = *reinterpret_cast<const float4*>(bias + col_of(j, q));
#pragma unroll
      for(int i = 0; i < TI; ++i) {
        acc[i][j][4 * q + 0] = bv.x;
        acc[i][j][4 * q + 1] = bv.y;
        acc[i][j][4 * q + 2] = bv.z;
        acc[i][j][4 * q + 3] = bv.w;
      }
    }

  // A and B pieces of a K step in registers: later steps' global loads are
  // in flight while this step's MFMAs run (one step ahead for TI = 2, two
  // for TI = 4, whose fewer waves per SIMD hide less latency).
  // 16-byte chunks: A tile BMt x 32 = TI per thread, B tile 128 x 32 = 2
  typedef bf16x8_t RegA[TI];
  typedef bf16x8_t RegB[2];
  RegA ra;
  RegB rb;
  auto gload = [&](int k0, RegA& ra, RegB& rb) {
#pragma unroll
    for(int it = 0; it < TI; ++it) {
      const int c = tid + 256 * it;
      if constexpr(AK && FW > 0) {
        const int64_t row = m0 + (c >> 2);
        const int kq = k0 + (c & 3) * 8;
        uint32_t bits = 0;
        if(row < lda && kq < 64 * FW)
          bits = (uint32_t)(reinterpret_cast<const uint64_t*>(A)[row * FW + (kq >> 6)] >>
                            (kq & 63)) & 0xffu;
#pragma unroll
        for(int e = 0; e < 8; ++e) ra[it][e] = (__bf16)(float)((bits >> e) & 1u);
      } else if constexpr(AK)
        ra[it] = *reinterpret_cast<const bf16x8_t*>(A + (size_t)(m0 + (c >> 2)) * lda + k0 +
                                                    (c & 3) * 8);
      else
        ra[it] = *reinterpret_cast<const bf16x8_t*>(A + (size_t)(k0 + (c & 31)) * lda + m0 +
                                                    (c >> 5) * 8);
    }
#pragma unroll
    for(int it = 0; it < 2; ++it) {
      const int c = tid + 256 * it;
      if constexpr(BKC)
        rb[it] = *reinterpret_cast<const bf16x8_t*>(B + (size_t)(n0 + (c >> 2)) * ldb + k0 +
                                                    (c & 3) * 8);
      else
        rb[it] = *reinterpret_cast<const bf16x8_t*>(B + (size_t)(k0 + (c & 31)) * ldb + n0 +
                                                    (c >> 5) * 8);
    }
  };
  // registers -> LDS buffer (K-contiguous whatever the global layout)
  auto lstore = [&](int buf, const RegA& ra, const RegB& rb) {
    __bf16(*As)[LDK] = reinterpret_cast<__bf16(*)[LDK]>(smem + buf * (BMt + BN) * LDK);
    __bf16(*Bs)[LDK] =
        reinterpret_cast<__bf16(*)[LDK]>(smem + buf * (BMt + BN) * LDK + BMt * LDK);
#pragma unroll
    for(int it = 0; it < TI; ++it) {
      const int c = tid + 256 * it;
      if constexpr(AK) {
        *reinterpret_cast<bf16x8_t*>(&As[c >> 2][(c & 3) * 8]) = ra[it];
      } else {
        const int kk = c & 31, mc = (c >> 5) * 8;
#pragma unroll
        for(int e = 0; e < 8; ++e) As[mc + e][kk] = ra[it][e];
      }
    }
#pragma unroll
    for(int it = 0; it < 2; ++it) {
      const int c = tid + 256 * it;
      if constexpr(BKC) {
        *reinterpret_cast<bf16x8_t*>(&Bs[c >> 2][(c & 3) * 8]) = rb[it];
      } else {
        const int kk = c & 31, nc = (c >> 5) * 8;
#pragma unroll
        for(int e = 0; e < 8; ++e) Bs[nc + e][kk] = rb[it][e];
      }
    }
  };
  // ---- MFMA on LDS buffer buf: TI x 2 tiles of 32 x 32, two K sub-steps of 16
  auto mfma_step = [&](int buf) {
    const __bf16(*As)[LDK] =
        reinterpret_cast<const __bf16(*)[LDK]>(smem + buf * (BMt + BN) * LDK);
    const __bf16(*Bs)[LDK] =
        reinterpret_cast<const __bf16(*)[LDK]>(smem + buf * (BMt + BN) * LDK + BMt * LDK);
#pragma unroll
    for(int ks = 0; ks < BK; ks += 16) {
      bf16x8_t a[TI], b[2];
      const int kq = ks + (lane >> 5) * 8;
#pragma unroll
      for(int i = 0; i < TI; ++i)
        a[i] = *reinterpret_cast<const bf16x8_t*>(&As[wm + i * 32 + (lane & 31)][kq]);
#pragma unroll
      for(int j = 0; j < 2; ++j)
        b[j] = *reinterpret_cast<const bf16x8_t*>(&Bs[wn + j * 32 + (lane & 31)][kq]);
#pragma unroll
      for(int i = 0; i < TI; ++i)
#pragma unroll
        for(int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[j], a[i], acc[i][j], 0, 0, 0);
    }
  };
  if constexpr(TI == 4) {
    // step k uses LDS buffer k % 2; while it runs, step k + 1 waits in
    // registers and step k + 2 is being loaded into the set step k used
    RegA ra1;
    RegB rb1;
    if(kbeg < kend) {
      gload(kbeg, ra, rb);
      if(kbeg + BK < kend) gload(kbeg + BK, ra1, rb1);
      lstore(0, ra, rb);
      __syncthreads();
    }
    for(int k0 = kbeg; k0 < kend;) {
      if(k0 + 2 * BK < kend) gload(k0 + 2 * BK, ra, rb);
      mfma_step(0);
      // buffer 1 was last read before the previous barrier
      if(k0 + BK < kend) lstore(1, ra1, rb1);
      __syncthreads();
      k0 += BK;
      if(k0 >= kend) break;
      if(k0 + 2 * BK < kend) gload(k0 + 2 * BK, ra1, rb1);
      mfma_step(1);
      if(k0 + BK < kend) lstore(0, ra, rb);
      __syncthreads();
      k0 += BK;
    }
  } else {
    int cur = 0;
    if(kbeg < kend) {
      gload(kbeg, ra, rb);
      lstore(0, ra, rb);
      __syncthreads();
    }
    for(int k0 = kbeg; k0 < kend; k0 += BK) {
      const bool more = k0 + BK < kend;
      if(more) gload(k0 + BK, ra, rb);
      mfma_step(cur);
      // the other buffer was last read before the previous barrier
      if(more) lstore(cur ^ 1, ra, rb);
      __syncthreads();
      cur ^= 1;
    }
  }
  // ---- epilogue
  const int rlane = m0 + wm + (lane & 31);
  if constexpr(EPI == EPI_HEAD) {
    // each bias / w3 piece serves all TI row tiles before the next is loaded
    float s[TI];
#pragma unroll
    for(int i = 0; i < TI; ++i) s[i] = 0.f;
#pragma unroll
    for(int j = 0; j < 2; ++j)
#pragma unroll
      for(int q = 0; q < 4; ++q) {
        const float4 bv = *reinterpret_cast<const float4*>(bias + col_of(j, q));
        const float4 wv3 = *reinterpret_cast<const float4*>(zaux + col_of(j, q));
#pragma unroll
        for(int i = 0; i < TI; ++i) {
          s[i] += (float)(__bf16)gelu(acc[i][j][4 * q + 0] + bv.x) * wv3.x;
          s[i] += (float)(__bf16)gelu(acc[i][j][4 * q + 1] + bv.y) * wv3.y;
          s[i] += (float)(__bf16)gelu(acc[i][j][4 * q + 2] + bv.z) * wv3.z;
          s[i] += (float)(__bf16)gelu(acc[i][j][4 * q + 3] + bv.w) * wv3.w;
        }
      }
#pragma unroll
    for(int i = 0; i < TI; ++i) {
      s[i] += __shfl_xor(s[i], 32, 64);  // the other half holds the row's other columns
      if(lane < 32) Cf[(size_t)(2 * bx + (wv & 1)) * ldc + rlane + i * 32] = s[i];
    }
    return;
  }
  if constexpr(EPI == EPI_BIAS_GELU_A || EPI == EPI_GELU_GRAD) {
    // bf16 outputs: each wave stages 32 x 64 of its tile in LDS (8-byte
    // writes) and stores whole 16-byte row pieces
    static_assert(4 * 32 * kStageLd <= 2 * (BMt + BN) * LDK, "staging exceeds the tile LDS");
    __bf16* stg = smem + wv * 32 * kStageLd;
#pragma unroll
    for(int i = 0; i < TI; ++i) {
      const int row = rlane + i * 32;
#pragma unroll
      for(int j = 0; j < 2; ++j)
#pragma unroll
        for(int q = 0; q < 4; ++q) {
          const int col = col_of(j, q);
          float4 v = {acc[i][j][4 * q], acc[i][j][4 * q + 1], acc[i][j][4 * q + 2],
                      acc[i][j][4 * q + 3]};
          if constexpr(EPI == EPI_BIAS_GELU_A) {
            v.x = gelu(v.x);
            v.y = gelu(v.y);
            v.z = gelu(v.z);
            v.w = gelu(v.w);
          } else {
            const float4 zv = *reinterpret_cast<const float4*>(zaux + (size_t)row * ldc + col);
            v.x *= gelu_grad(zv.x);
            v.y *= gelu_grad(zv.y);
            v.z *= gelu_grad(zv.z);
            v.w *= gelu_grad(zv.w);
          }
          bf16x4_t o = {(__bf16)v.x, (__bf16)v.y, (__bf16)v.z, (__bf16)v.w};
          *reinterpret_cast<bf16x4_t*>(stg + (lane & 31) * kStageLd + col - n0 - wn) = o;
        }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
#pragma unroll
      for(int q = 0; q < 4; ++q) {
        const int c = lane + 64 * q, lr = c >> 3, cc = (c & 7) * 8;
        *reinterpret_cast<bf16x8_t*>(Cb + (size_t)(m0 + wm + i * 32 + lr) * ldc + n0 + wn + cc) =
            *reinterpret_cast<const bf16x8_t*>(stg + lr * kStageLd + cc);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    return;
  }
#pragma unroll
  for(int i = 0; i < TI; ++i)
#pragma unroll
    for(int j = 0; j < 2; ++j)
#pragma unroll
      for(int q = 0; q < 4; ++q) {
        const int row = rlane + i * 32, col = col_of(j, q);
        const size_t o = (size_t)row * ldc + col;
        float4 v = {acc[i][j][4 * q], acc[i][j][4 * q + 1], acc[i][j][4 * q + 2],
                    acc[i][j][4 * q + 3]};
        if constexpr(EPI == EPI_BIAS_GELU) {
          bf16x4_t g = {(__bf16)gelu(v.x), (__bf16)gelu(v.y), (__bf16)gelu(v.z),
                        (__bf16)gelu(v.w)};
          *reinterpret_cast<bf16x4_t*>(Cb + o) = g;
        }
        *reinterpret_cast<float4*>(Cf + o) = v;
      }
}

// Occupation bits -> bf16 features (W words per det, 64 bits per word;
// features beyond 64*W are zero). A thread writes 8 features (16 bytes):
// 16 threads per row, 16 rows per block.
template <int W>
__global__ void features_kernel(const uint64_t* __restrict__ dets, int64_t n, int64_t npad,
                                __bf16* __restrict__ X) {
  const int64_t i = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
  const int f0 = (threadIdx.x & 15) * 8;
  if(i >= npad) return;
  uint32_t bits = 0;
  if(i < n && f0 < 64 * W) bits = (uint32_t)(dets[i * W + (f0 >> 6)] >> (f0 & 63)) & 0xffu;
  bf16x8_t v;
#pragma unroll
  for(int e = 0; e < 8; ++e) v[e] = (__bf16)(float)((bits >> e) & 1u);
  *reinterpret_cast<bf16x8_t*>(X + i * kIn + f0) = v;
}

// y = A2 . w3 + b3; with targets, dy = 2 (y - t) / B and partial losses.
static_assert(kHid == 64 * 8, "head_kernel maps 8 columns to each lane");
__global__ __launch_bounds__(256) void head_kernel(const __bf16* __restrict__ A2,
                                                   const float* __restrict__ w3,
                                                   const float* __restrict__ b3, int64_t n,
                                                   float* __restrict__ y,
                                                   const float* __restrict__ t,
                                                   float* __restrict__ dy, float inv_b,
                                                   float* __restrict__ loss_part) {
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  float l = 0.f;
  if(row < n) {
    // lane l: columns 8l .. 8l+7 (one 16-byte load; kHid = 512 = 64 x 8)
    const bf16x8_t a = *reinterpret_cast<const bf16x8_t*>(A2 + row * kHid + 8 * lane);
    float s = 0.f;
#pragma unroll
    for(int e = 0; e < 8; ++e) s += (float)a[e] * w3[8 * lane + e];
    s = wave_sum(s) + b3[0];
    if(lane == 0) {
      y[row] = s;
      if(t) {
        const float e = s - t[row];
        dy[row] = 2.f * e * inv_b;
        l = e * e;
      }
    }
  }
  if(t) {
    __shared__ float red[4];
    if(lane == 0) red[threadIdx.x >> 6] = l;
    __syncthreads();
    if(threadIdx.x == 0) loss_part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
  }
}

// Inference head: y = sum of the np column partials of the last layer's
// fused epilogue (fixed order) + b3.
__global__ void head_sum_kernel(const float* __restrict__ part, int np, int64_t ld,
                                const float* __restrict__ b3, int64_t n, float* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  float s = part[i];
  for(int p = 1; p < np; ++p) s += part[p * ld + i];
  y[i] = s + b3[0];
}

// dZ2 = (dy * w3) * gelu'(Z2) as bf16; a thread takes 8 consecutive
// elements of a row (two 16-byte loads of Z2, one 16-byte store).
__global__ void head_back_kernel(const float* __restrict__ dy, const float* __restrict__ w3,
                                 const float* __restrict__ Z2, int64_t n,
                                 __bf16* __restrict__ dZ2) {
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 8;
  if(i >= n * kHid) return;
  const int j = (int)(i % kHid);
  const float d = dy[i / kHid];
  const float4 z0 = *reinterpret_cast<const float4*>(Z2 + i);
  const float4 z1 = *reinterpret_cast<const float4*>(Z2 + i + 4);
  const float4 w0 = *reinterpret_cast<const float4*>(w3 + j);
  const float4 w1 = *reinterpret_cast<const float4*>(w3 + j + 4);
  const bf16x8_t o = {(__bf16)(d * w0.x * gelu_grad(z0.x)), (__bf16)(d * w0.y * gelu_grad(z0.y)),
                      (__bf16)(d * w0.z * gelu_grad(z0.z)), (__bf16)(d * w0.w * gelu_grad(z0.w)),
                      (__bf16)(d * w1.x * gelu_grad(z1.x)), (__bf16)(d * w1.y * gelu_grad(z1.y)),
                      (__bf16)(d * w1.z * gelu_grad(z1.z)), (__bf16)(d * w1.w * gelu_grad(z1.w))};
  *reinterpret_cast<bf16x8_t*>(dZ2 + i) = o;
}

// Column sums over row slices: part[s][j] = sum_{b in slice s} M[b][j] w[b]
// (bf16 rows of kHid columns; w = 1 for bias gradients, dy for dw3). Wave w
// of the block takes rows b0 + w, b0 + w + 4, ... (two in flight), lane l
// columns 8l .. 8l + 7 with one 16-byte load per row; the four waves' sums
// are added in a fixed order.
static_assert(kHid == 64 * 8, "colsum_part_kernel maps 8 columns to each lane");
__global__ __launch_bounds__(256) void colsum_part_kernel(const __bf16* __restrict__ Mt,
                                                          int64_t n,
                                                          const float* __restrict__ wrow,
                                                          float* __restrict__ part) {
  __shared__ float red[4][kHid];
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  const int64_t b0 = (int64_t)blockIdx.x * per, b1 = b0 + per < n ? b0 + per : n;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  float a0[8], a1[8];
#pragma unroll
  for(int e = 0; e < 8; ++e) a0[e] = a1[e] = 0.f;
  int64_t b = b0 + w;
  for(; b + 4 < b1; b += 8) {
    const bf16x8_t v0 = *reinterpret_cast<const bf16x8_t*>(Mt + b * kHid + 8 * l);
    const bf16x8_t v1 = *reinterpret_cast<const bf16x8_t*>(Mt + (b + 4) * kHid + 8 * l);
    const float w0 = wrow ? wrow[b] : 1.f, w1 = wrow ? wrow[b + 4] : 1.f;
#pragma unroll
    for(int e = 0; e < 8; ++e) {
      a0[e] += (float)v0[e] * w0;
      a1[e] += (float)v1[e] * w1;
    }
  }
  if(b < b1) {
    const bf16x8_t v0 = *reinterpret_cast<const bf16x8_t*>(Mt + b * kHid + 8 * l);
    const float w0 = wrow ? wrow[b] : 1.f;
#pragma unroll
    for(int e = 0; e < 8; ++e) a0[e] += (float)v0[e] * w0;
  }
#pragma unroll
  for(int e = 0; e < 8; ++e) red[w][8 * l + e] = a0[e] + a1[e];
  __syncthreads();
  for(int j = threadIdx.x; j < kHid; j += 256)
    part[(int64_t)blockIdx.x * kHid + j] = (red[0][j] + red[1][j]) + (red[2][j] + red[3][j]);
}

// out[j] = sum of part[z * stride + j] over the slices z in a fixed order:
// thread group g (of 4) takes z = g, g + 4, ... (two chains), the group sums
// are added as (g0 + g1) + (g2 + g3). 64 columns per block.
__global__ __launch_bounds__(256) void slice_sum_kernel(const float* __restrict__ part,
                                                        int nslice, int64_t stride,
                                                        int64_t count, float* __restrict__ out) {
  __shared__ float red[4][64];
  const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int64_t j = (int64_t)blockIdx.x * 64 + c;
  float s0 = 0.f, s1 = 0.f;
  if(j < count) {
    int z = g;
#pragma unroll 4
    for(; z + 4 < nslice; z += 8) {
      s0 += part[(int64_t)z * stride + j];
      s1 += part[(int64_t)(z + 4) * stride + j];
    }
    if(z < nslice) s0 += part[(int64_t)z * stride + j];
  }
  red[g][c] = s0 + s1;
  __syncthreads();
  if(g == 0 && j < count) out[j] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
}

// out[0] = sum of v[0 .. n) (one block of 1024 threads, four chains each,
// fixed order).
__global__ __launch_bounds__(1024) void sum_kernel(const float* __restrict__ v, int64_t n,
                                                   float* __restrict__ out) {
  __shared__ float red[1024];
  float s[4] = {0.f, 0.f, 0.f, 0.f};
  int64_t i = threadIdx.x;
  for(; i + 3 * 1024 < n; i += 4 * 1024) {
#pragma unroll
    for(int u = 0; u < 4; ++u) s[u] += v[i + u * 1024];
  }
  for(; i < n; i += 1024) s[0] += v[i];
  red[threadIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
  __syncthreads();
  for(int off = 512; off > 0; off >>= 1) {
    if((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if(threadIdx.x == 0) out[0] = red[0];
}

// Adam on fp32 master weights, bf16 shadow for compute.
__global__ void adam_kernel(float* __restrict__ w, __bf16* __restrict__ wb,
                            const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int64_t n, float lr, float b1, float b2,
                            float c1, float c2, float eps) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if(wb) wb[i] = (__bf16)wi;
}

__global__ void to_bf16_kernel(const float* __restrict__ a, __bf16* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) b[i] = (__bf16)a[i];
}

}  // namespace

// ---- host side -----------------------------------------------------------
struct Mlp {
  // parameter layout in one fp32 block: W1 (512x128) | b1 | W2 (512x512) |
  // b2 | w3 (512) | b3 (1)
  static constexpr int64_t oW1 = 0, ob1 = oW1 + kHid * kIn, oW2 = ob1 + kHid,
                           ob2 = oW2 + (int64_t)kHid * kHid, ow3 = ob2 + kHid, ob3 = ow3 + kHid,
                           nparam = ob3 + 1;
  DevBuf p, g, m, v, pb;  // fp32 params, grads, Adam moments, bf16 params
  int64_t step = 0;
  // activation workspace for up to `cap` rows
  int64_t cap = 0;
  DevBuf X, Z1, A1, Z2, A2, y, dy, dZ2, dZ1, dA1, loss;
  DevBuf wk;  // split-K partial weight gradients / column-sum slices
  void reserve(int64_t rows) {
    // split-K / column-sum slices: at most 32 x 512 x 512 or 256 x 512 floats
    wk.reserve(sizeof(float) * (size_t)32 * kHid * kHid);
    if(rows <= cap) return;
    cap = rows;
    X.alloc(2 * rows * kIn);
    Z1.alloc(4 * rows * kHid);
    A1.alloc(2 * rows * kHid);
    Z2.alloc(4 * rows * kHid);
    A2.alloc(2 * rows * kHid);
    y.alloc(4 * rows);
    dy.alloc(4 * rows);
    dZ2.alloc(2 * rows * kHid);
    dZ1.alloc(2 * rows * kHid);
    dA1.alloc(4 * rows * kHid);
    loss.alloc(4 * (rows / 4 + 2) + 16);  // per-block losses, then their total
  }
};

}  // namespace mhip

struct mhip_mlp {
  mhip::Mlp net;
  int device = 0;
};

namespace mhip {
namespace {

template <bool AK, bool BKC, int EPI, int TI = 2, int FW = 0>
void gemm(hipStream_t st, int M, int N, int K, const __bf16* A, int lda, const __bf16* B, int ldb,
          const float* bias, const float* zaux, float* Cf, __bf16* Cb, int ldc, int splits = 1) {
  MHIP_REQUIRE(M % (64 * TI) == 0 && N % BN == 0 && K % BK == 0, MHIP_ERR_INVALID_ARG,
               "gemm dims must be multiples of the block tile and of 32 in K");
  MHIP_REQUIRE(splits >= 1 && K % (splits * BK) == 0 && (splits == 1 || EPI == EPI_F32),
               MHIP_ERR_INVALID_ARG, "bad split-K");
  static_assert(FW == 0 || AK, "determinant-formed A needs the K-contiguous layout");
  hipLaunchKernelGGL((gemm_bf16_kernel<AK, BKC, EPI, TI, FW>), dim3(N / BN, M / (64 * TI), splits),
                     dim3(256), 0,
                     st, M, N, K, A, lda, B, ldb, bias, zaux, Cf, Cb, ldc);
  MHIP_HIP_CHECK(hipGetLastError());
}

// Weight gradient C (M x N, fp32) = A^T B over K batch rows, split over K
// into fixed slices summed in order (parallel over the batch).
void gemm_wgrad(hipStream_t st, Mlp& net, int M, int N, int K, const __bf16* A, int lda,
                const __bf16* B, int ldb, float* C) {
  // split K until the launch has 512 blocks (two per CU), each slice at
  // least 256 batch rows
  const int tiles = (M / 128) * (N / BN);
  int splits = 1;
  while(splits < 128 && tiles * splits < 512 && K % (2 * splits * BK) == 0 &&
        K / (2 * splits) >= 256)
    splits *= 2;
  net.wk.reserve(sizeof(float) * (size_t)splits * M * N);
  gemm<false, false, EPI_F32>(st, M, N, K, A, lda, B, ldb, nullptr, nullptr, net.wk.as<float>(),
                              nullptr, N, splits);
  const int64_t count = (int64_t)M * N;
  hipLaunchKernelGGL(slice_sum_kernel, dim3((count + 63) / 64), dim3(256), 0, st,
                     net.wk.as<float>(), splits, count, count, C);
  MHIP_HIP_CHECK(hipGetLastError());
}

// Column sums of a (rows x kHid) bf16 matrix, weighted by wrow if given.
void colsum(hipStream_t st, Mlp& net, const __bf16* Mt, int64_t rows, const float* wrow,
            float* out) {
  const int slices = (int)std::max<int64_t>(1, std::min<int64_t>(512, rows / 64));
  net.wk.reserve(sizeof(float) * (size_t)slices * kHid);
  hipLaunchKernelGGL(colsum_part_kernel, dim3(slices), dim3(256), 0, st, Mt, rows, wrow,
                     net.wk.as<float>());
  hipLaunchKernelGGL(slice_sum_kernel, dim3((kHid + 63) / 64), dim3(256), 0, st,
                     net.wk.as<float>(), slices, (int64_t)kHid, (int64_t)kHid, out);
  MHIP_HIP_CHECK(hipGetLastError());
}

const __bf16* bfp(const DevBuf& b, int64_t off = 0) { return b.as<__bf16>() + off; }

// Training forward pass for rows [0, npad) already featurised into net.X;
// keeps the pre-activations Z1, Z2 for the backward pass.
void forward(hipStream_t st, Mlp& net, int64_t npad) {
  const float* P = net.p.as<float>();
  gemm<true, true, EPI_BIAS_GELU>(st, (int)npad, kHid, kIn, bfp(net.X), kIn,
                                  bfp(net.pb, Mlp::oW1), kIn, P + Mlp::ob1, nullptr,
                                  net.Z1.as<float>(), net.A1.as<__bf16>(), kHid);
  gemm<true, true, EPI_BIAS_GELU>(st, (int)npad, kHid, kHid, bfp(net.A1), kHid,
                                  bfp(net.pb, Mlp::oW2), kHid, P + Mlp::ob2, nullptr,
                                  net.Z2.as<float>(), net.A2.as<__bf16>(), kHid);
}

// Inference for n determinants (rows padded to npad): the first layer forms
// its inputs from the determinant bits as it loads them, the head rides on
// the last layer's epilogue (per-wave column partials, 2 per 128-column
// tile, in A2's storage, which is far larger) and a small kernel adds them.
template <int W>
void infer(hipStream_t st, Mlp& net, const uint64_t* dets, int64_t n, int64_t npad, float* out) {
  const float* P = net.p.as<float>();
  gemm<true, true, EPI_BIAS_GELU_A, 2, W>(st, (int)npad, kHid, kIn,
                                          reinterpret_cast<const __bf16*>(dets), (int)n,
                                          bfp(net.pb, Mlp::oW1), kIn, P + Mlp::ob1, nullptr,
                                          nullptr, net.A1.as<__bf16>(), kHid);
  static_assert(2 * (kHid / BN) * sizeof(float) <= kHid * sizeof(__bf16), "head partials");
  gemm<true, true, EPI_HEAD, 4>(st, (int)npad, kHid, kHid, bfp(net.A1), kHid,
                                bfp(net.pb, Mlp::oW2), kHid, P + Mlp::ob2, P + Mlp::ow3,
                                net.A2.as<float>(), nullptr, (int)npad);
  hipLaunchKernelGGL(head_sum_kernel, dim3((n + 255) / 256), dim3(256), 0, st,
                     net.A2.as<float>(), 2 * (kHid / BN), npad, P + Mlp::ob3, n, out);
  MHIP_HIP_CHECK(hipGetLastError());
}

template <int W>
void featurize(hipStream_t st, Mlp& net, const uint64_t* dets, int64_t n, int64_t npad) {
  hipLaunchKernelGGL(features_kernel<W>, dim3((npad + 15) / 16), dim3(256), 0, st, dets, n, npad,
                     net.X.as<__bf16>());
}

}  // namespace
}  // namespace mhip

using namespace mhip;

extern "C" {

int mhip_mlp_create(mhip_context* ctx, uint64_t seed, mhip_mlp** out) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(out, MHIP_ERR_INVALID_ARG, "null output");
    std::unique_ptr<mhip_mlp> h(new mhip_mlp());
    h->device = ctx->device;
    Mlp& net = h->net;
    // He-style init from a fixed seed (host), biases zero.
    std::vector<float> p(Mlp::nparam, 0.f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> nd(0.f, 1.f);
    for(int64_t i = 0; i < kHid * kIn; ++i) p[Mlp::oW1 + i] = nd(rng) * std::sqrt(2.f / kIn);
    for(int64_t i = 0; i < (int64_t)kHid * kHid; ++i) p[Mlp::oW2 + i] = nd(rng) * std::sqrt(2.f / kHid);
    for(int64_t i = 0; i < kHid; ++i) p[Mlp::ow3 + i] = nd(rng) * std::sqrt(1.f / kHid);
    net.p.alloc(4 * Mlp::nparam);
    net.g.alloc(4 * Mlp::nparam);
    net.m.alloc(4 * Mlp::nparam);
    net.v.alloc(4 * Mlp::nparam);
    net.pb.alloc(2 * Mlp::nparam);
    MHIP_HIP_CHECK(hipMemcpyAsync(net.p.p, p.data(), 4 * Mlp::nparam, hipMemcpyHostToDevice, ctx->stream));
    MHIP_HIP_CHECK(hipMemsetAsync(net.m.p, 0, 4 * Mlp::nparam, ctx->stream));
    MHIP_HIP_CHECK(hipMemsetAsync(net.v.p, 0, 4 * Mlp::nparam, ctx->stream));
    hipLaunchKernelGGL(to_bf16_kernel, dim3((Mlp::nparam + 255) / 256), dim3(256), 0, ctx->stream,
                       net.p.as<float>(), net.pb.as<__bf16>(), Mlp::nparam);
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    *out = h.release();
  });
}

int mhip_mlp_destroy(mhip_mlp* mlp) {
  return guarded([&] {
    if(!mlp) return;
    (void)hipSetDevice(mlp->device);
    (void)hipDeviceSynchronize();
    delete mlp;
  });
}

int64_t mhip_mlp_num_params(void) { return Mlp::nparam; }

int mhip_mlp_get_params(mhip_context* ctx, const mhip_mlp* mlp, float* host) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(mlp && host, MHIP_ERR_INVALID_ARG, "null argument");
    MHIP_HIP_CHECK(hipMemcpyAsync(host, mlp->net.p.p, 4 * Mlp::nparam, hipMemcpyDeviceToHost, ctx->stream));
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

int mhip_mlp_set_params(mhip_context* ctx, mhip_mlp* mlp, const float* host) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(mlp && host, MHIP_ERR_INVALID_ARG, "null argument");
    MHIP_HIP_CHECK(hipMemcpyAsync(mlp->net.p.p, host, 4 * Mlp::nparam, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(to_bf16_kernel, dim3((Mlp::nparam + 255) / 256), dim3(256), 0, ctx->stream,
                       mlp->net.p.as<float>(), mlp->net.pb.as<__bf16>(), Mlp::nparam);
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

int mhip_mlp_get_grads(mhip_context* ctx, const mhip_mlp* mlp, float* host) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(mlp && host, MHIP_ERR_INVALID_ARG, "null argument");
    MHIP_HIP_CHECK(hipMemcpyAsync(host, mlp->net.g.p, 4 * Mlp::nparam, hipMemcpyDeviceToHost, ctx->stream));
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

int mhip_mlp_predict(mhip_context* ctx, mhip_mlp* mlp, int words_per_det, const uint64_t* d_dets,
                     int64_t n, float* d_out) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(mlp && (n == 0 || (d_dets && d_out)), MHIP_ERR_INVALID_ARG, "null argument");
    MHIP_REQUIRE(words_per_det == 1 || words_per_det == 2, MHIP_ERR_UNSUPPORTED, "bad W");
    Mlp& net = mlp->net;
    // rows padded to the 256-row tile of the last layer's GEMM
    const int64_t chunk = 1 << 20;
    net.reserve(std::min<int64_t>((n + 255) / 256 * 256, chunk));
    for(int64_t s = 0; s < n; s += chunk) {
      const int64_t nb = std::min(chunk, n - s), npad = (nb + 255) / 256 * 256;
      if(words_per_det == 1)
        infer<1>(ctx->stream, net, d_dets + s, nb, npad, d_out + s);
      else
        infer<2>(ctx->stream, net, d_dets + 2 * s, nb, npad, d_out + s);
    }
  });
}

// One optimisation step on a mini-batch (DEVICE dets and fp32 targets);
// returns the batch MSE. With lr == 0 only gradients are formed.
int mhip_mlp_train_step(mhip_context* ctx, mhip_mlp* mlp, int words_per_det,
                        const uint64_t* d_dets, const float* d_target, int64_t n, float lr,
                        float* loss_out) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(mlp && d_dets && d_target && n > 0, MHIP_ERR_INVALID_ARG, "bad argument");
    MHIP_REQUIRE(words_per_det == 1 || words_per_det == 2, MHIP_ERR_UNSUPPORTED, "bad W");
    Mlp& net = mlp->net;
    hipStream_t st = ctx->stream;
    const int64_t npad = (n + 127) / 128 * 128;
    MHIP_REQUIRE(npad <= (1 << 22), MHIP_ERR_INVALID_ARG, "batch too large");
    net.reserve(npad);
    if(words_per_det == 1)
      featurize<1>(st, net, d_dets, n, npad);
    else
      featurize<2>(st, net, d_dets, n, npad);
    forward(st, net, npad);
    float* P = net.p.as<float>();
    float* G = net.g.as<float>();
    // padded rows must not contribute: dy = 0 there
    MHIP_HIP_CHECK(hipMemsetAsync(net.dy.p, 0, 4 * npad, st));
    const int64_t nblk = (n + 3) / 4;
    hipLaunchKernelGGL(head_kernel, dim3(nblk), dim3(256), 0, st, net.A2.as<__bf16>(),
                       P + Mlp::ow3, P + Mlp::ob3, n, net.y.as<float>(), d_target,
                       net.dy.as<float>(), 1.f / (float)n, net.loss.as<float>());
    // head gradients
    colsum(st, net, net.A2.as<__bf16>(), npad, net.dy.as<float>(), G + Mlp::ow3);
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(1024), 0, st, net.dy.as<float>(), npad, G + Mlp::ob3);
    hipLaunchKernelGGL(head_back_kernel, dim3((npad * kHid / 8 + 255) / 256), dim3(256), 0, st,
                       net.dy.as<float>(), P + Mlp::ow3, net.Z2.as<float>(), npad,
                       net.dZ2.as<__bf16>());
    // layer 2: dW2 = dZ2^T A1, db2 = colsum dZ2, dZ1 = (dZ2 W2) * gelu'(Z1)
    gemm_wgrad(st, net, kHid, kHid, (int)npad, bfp(net.dZ2), kHid, bfp(net.A1), kHid,
               G + Mlp::oW2);
    colsum(st, net, net.dZ2.as<__bf16>(), npad, nullptr, G + Mlp::ob2);
    gemm<true, false, EPI_GELU_GRAD>(st, (int)npad, kHid, kHid, bfp(net.dZ2), kHid,
                                     bfp(net.pb, Mlp::oW2), kHid, nullptr, net.Z1.as<float>(),
                                     nullptr, net.dZ1.as<__bf16>(), kHid);
    // layer 1: dW1 = dZ1^T X, db1 = colsum dZ1
    gemm_wgrad(st, net, kHid, kIn, (int)npad, bfp(net.dZ1), kHid, bfp(net.X), kIn,
               G + Mlp::oW1);
    colsum(st, net, net.dZ1.as<__bf16>(), npad, nullptr, G + Mlp::ob1);
    if(lr > 0.f) {
      net.step++;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.f - std::pow(b1, (float)net.step), c2 = 1.f - std::pow(b2, (float)net.step);
      hipLaunchKernelGGL(adam_kernel, dim3((Mlp::nparam + 255) / 256), dim3(256), 0, st, P,
                         net.pb.as<__bf16>(), G, net.m.as<float>(), net.v.as<float>(),
                         Mlp::nparam, lr, b1, b2, c1, c2, 1e-8f);
    }
    MHIP_HIP_CHECK(hipGetLastError());
    if(loss_out) {  // the only wait of a step: callers timing steps pass NULL
      float* tot = net.loss.as<float>() + (net.loss.bytes / 4 - 4);
      hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(1024), 0, st, net.loss.as<float>(), nblk, tot);
      float l = 0.f;
      d2h(ctx, &l, tot, 4);
      *loss_out = l / (float)n;
    }
  });
}

// bf16 GEMM entry for parity tests: C (fp32, M x N, row-major) = A B^T with
// A (M x K) and B (N x K) row-major bf16 (raw 16-bit patterns).
int mhip_gemm_bf16_nt(mhip_context* ctx, int M, int N, int K, const uint16_t* d_A,
                      const uint16_t* d_B, float* d_C) {
  return guarded([&] {
    use_device(ctx);
    gemm<true, true, EPI_F32>(ctx->stream, M, N, K, reinterpret_cast<const __bf16*>(d_A), K,
                              reinterpret_cast<const __bf16*>(d_B), K, nullptr, nullptr, d_C,
                              nullptr, N);
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

}  // extern "C"
