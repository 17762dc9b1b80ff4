// Spin-summed one- and two-particle reduced density matrices on the device:
// form_rdms of DoubleLoopHamiltonianGenerator (double_loop.hpp:116-165) with
// the case dispatch of rdm_contributions (util/rdms.hpp:103-130).
//
// The pairs are those of the Hamiltonian build: popcount(bra ^ ket) <= 4
// between non-zero determinants. They come from the K1 passes (tiled scan or
// product-space tables) run with a negative threshold, so every connected
// pair is kept even where its matrix element is zero. One wavefront per bra
// row then takes 64 of the row's kets at a time and adds each pair's
// contributions with FP64 atomics (global_atomic_add_f64): the sums are those
// of the reference, in an order that depends on scheduling.
#include <memory>

#include "mhip_device.hpp"
#include "mhip_internal.hpp"

namespace mhip {
namespace {

// ordm(p,q) = ordm[p + n q], trdm(p,q,r,s) = trdm[p + n q + n^2 r + n^3 s]
// (column-major spans, types.hpp:24-33); trdm may be null.
struct RdmOut {
  int64_t n;
  double* ordm;
  double* trdm;
  __device__ void d1(uint32_t p, uint32_t q, double v) const {
    unsafeAtomicAdd(ordm + p + n * q, v);
  }
  __device__ void d2(uint32_t p, uint32_t q, uint32_t r, uint32_t s, double v) const {
    unsafeAtomicAdd(trdm + p + n * (q + n * (r + n * (int64_t)s)), v);
  }
};

// single_excitation_sign_indices (sd_operations.hpp:516-525)
__device__ __forceinline__ double single_sign_indices(uint64_t bra, uint64_t ket, uint64_t ex,
                                                      uint32_t& o, uint32_t& v) {
  o = ctz64(ket & ex);
  v = ctz64(bra & ex);
  return excitation_sign(ket, v, o);
}

// rdm_contributions_4 (util/rdms.hpp:15-26) with doubles_sign_indices
// (sd_operations.hpp:528-543)
__device__ void rdm_4(const RdmOut& R, uint64_t bra, uint64_t ket, uint64_t ex, double val) {
  uint32_t o1, v1, o2, v2;
  double sign = single_sign_indices(bra, ket, ex, o1, v1);
  const uint64_t f = (1ull << o1) | (1ull << v1);
  sign *= single_sign_indices(bra, ket ^ f, ex ^ f, o2, v2);
  val *= sign * 0.5;
  R.d2(v1, o1, v2, o2, val);
  R.d2(v2, o1, v1, o2, -val);
  R.d2(v1, o2, v2, o1, -val);
  R.d2(v2, o2, v1, o1, val);
}

// rdm_contributions_22 (util/rdms.hpp:28-42)
__device__ void rdm_22(const RdmOut& R, uint64_t bra_a, uint64_t ket_a, uint64_t ex_a,
                       uint64_t bra_b, uint64_t ket_b, uint64_t ex_b, double val) {
  uint32_t o1, v1, o2, v2;
  const double sa = single_sign_indices(bra_a, ket_a, ex_a, o1, v1);
  const double sb = single_sign_indices(bra_b, ket_b, ex_b, o2, v2);
  val *= sa * sb * 0.5;
  R.d2(v1, o1, v2, o2, val);
  R.d2(v2, o2, v1, o1, val);
}

// rdm_contributions_2 (util/rdms.hpp:44-67): `same` / `othr` are the bra's
// occupied orbitals of the excited spin and of the other spin.
__device__ void rdm_2(const RdmOut& R, uint64_t bra, uint64_t ket, uint64_t ex, uint64_t same,
                      uint64_t othr, double val) {
  uint32_t o1, v1;
  const double sign = single_sign_indices(bra, ket, ex, o1, v1);
  R.d1(v1, o1, sign * val);
  if(!R.trdm) return;
  val *= sign * 0.5;
  for(uint64_t m = same; m; m &= m - 1) {
    const uint32_t p = ctz64(m);
    R.d2(v1, o1, p, p, val);
    R.d2(p, p, v1, o1, val);
    R.d2(v1, p, p, o1, -val);
    R.d2(p, o1, v1, p, -val);
  }
  for(uint64_t m = othr; m; m &= m - 1) {
    const uint32_t p = ctz64(m);
    R.d2(v1, o1, p, p, val);
    R.d2(p, p, v1, o1, val);
  }
}

// rdm_contributions_diag (util/rdms.hpp:69-101)
__device__ void rdm_diag(const RdmOut& R, uint64_t oa, uint64_t ob, double val) {
  for(uint64_t m = oa; m; m &= m - 1) R.d1(ctz64(m), ctz64(m), val);
  for(uint64_t m = ob; m; m &= m - 1) R.d1(ctz64(m), ctz64(m), val);
  if(!R.trdm) return;
  val *= 0.5;
  for(uint64_t mq = oa; mq; mq &= mq - 1)
    for(uint64_t mp = oa; mp; mp &= mp - 1) {
      const uint32_t p = ctz64(mp), q = ctz64(mq);
      R.d2(p, p, q, q, val);
      R.d2(p, q, p, q, -val);
    }
  for(uint64_t mq = ob; mq; mq &= mq - 1)
    for(uint64_t mp = ob; mp; mp &= mp - 1) {
      const uint32_t p = ctz64(mp), q = ctz64(mq);
      R.d2(p, p, q, q, val);
      R.d2(p, q, p, q, -val);
    }
  for(uint64_t mq = ob; mq; mq &= mq - 1)
    for(uint64_t mp = oa; mp; mp &= mp - 1) {
      const uint32_t p = ctz64(mp), q = ctz64(mq);
      R.d2(p, p, q, q, val);
      R.d2(q, q, p, p, val);
    }
}

// One wavefront per bra row i over the row's connected kets j (CSR columns).
template <int W>
__global__ __launch_bounds__(256) void rdm_kernel(const uint64_t* __restrict__ bra, int64_t nbra,
                                                  const uint64_t* __restrict__ ket,
                                                  const int64_t* __restrict__ rowptr,
                                                  const int32_t* __restrict__ colind,
                                                  const double* __restrict__ C, RdmOut R) {
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if(i >= nbra) return;
  const uint32_t lane = threadIdx.x & 63;
  const Spins<W> b = load_spins<W>(bra, i);
  const double ci = C[i];
  const int64_t end = rowptr[i + 1];
  for(int64_t e = rowptr[i] + lane; e < end; e += 64) {
    const int64_t j = colind[e];
    const double val = ci * C[j];  // double_loop.hpp:149
    if(!(fabs(val) > 1e-16)) continue;
    const Spins<W> k = load_spins<W>(ket, j);
    const uint64_t exa = b.a ^ k.a, exb = b.b ^ k.b;
    const int ca = popc64(exa), cb = popc64(exb);
    if(ca == 4 && R.trdm)
      rdm_4(R, b.a, k.a, exa, val);
    else if(cb == 4 && R.trdm)
      rdm_4(R, b.b, k.b, exb, val);
    else if(ca == 2 && cb == 2 && R.trdm)
      rdm_22(R, b.a, k.a, exa, b.b, k.b, exb, val);
    else if(ca == 2)
      rdm_2(R, b.a, k.a, exa, b.a, b.b, val);
    else if(cb == 2)
      rdm_2(R, b.b, k.b, exb, b.b, b.a, val);
    else
      rdm_diag(R, b.a, b.b, val);
  }
}

}  // namespace

void form_rdms(mhip_context* ctx, int W, const uint64_t* d_bra, int64_t nbra,
               const uint64_t* d_ket, int64_t nket, const double* d_C, double* d_ordm,
               double* d_trdm) {
  use_device(ctx);
  MHIP_REQUIRE(W == 1 || W == 2, MHIP_ERR_INVALID_ARG, "words_per_det must be 1 or 2");
  MHIP_REQUIRE(ctx->ints.norb > 0, MHIP_ERR_NO_INTEGRALS,
               "mhip_set_integrals must be called before forming RDMs");
  MHIP_REQUIRE(d_ordm, MHIP_ERR_INVALID_ARG, "null 1-RDM");
  if(nbra == 0 || nket == 0) return;
  MHIP_REQUIRE(d_C, MHIP_ERR_INVALID_ARG, "null coefficients");
  // the connected pairs: the Hamiltonian pattern with nothing dropped (the
  // build's timings are those of the caller's last build, not this one)
  const mhip_timings saved = ctx->timings;
  std::unique_ptr<mhip_csr> P(ham_build(ctx, W, d_bra, nbra, d_ket, nket, -1.0));
  ctx->timings = saved;
  const RdmOut R{ctx->ints.norb, d_ordm, d_trdm};
  const dim3 grid((unsigned)((nbra + 3) / 4));
  if(W == 1)
    hipLaunchKernelGGL(rdm_kernel<1>, grid, dim3(256), 0, ctx->stream, d_bra, nbra, d_ket,
                       P->rowptr.as<int64_t>(), P->colind.as<int32_t>(), d_C, R);
  else
    hipLaunchKernelGGL(rdm_kernel<2>, grid, dim3(256), 0, ctx->stream, d_bra, nbra, d_ket,
                       P->rowptr.as<int64_t>(), P->colind.as<int32_t>(), d_C, R);
  MHIP_HIP_CHECK(hipGetLastError());
  MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

}  // namespace mhip

using namespace mhip;

extern "C" {

int mhip_form_rdms(mhip_context* ctx, int words_per_det, const uint64_t* d_bra, int64_t nbra,
                   const uint64_t* d_ket, int64_t nket, const double* d_C, double* d_ordm,
                   double* d_trdm) {
  return guarded(
      [&] { form_rdms(ctx, words_per_det, d_bra, nbra, d_ket, nket, d_C, d_ordm, d_trdm); });
}

}  // extern "C"
