// sparsexx::spblas::gespmbv for host csr_matrix arguments (restating
// src/sparsexx/include/sparsexx/spblas/spmbv.hpp:37-73):
// AV = ALPHA * A * V + BETA * AV over K columns. The product runs on the
// device (mhip_spmv on the default context): the host matrix and blocks are
// staged through HBM for the call. Device-resident callers use the
// DeviceCSR overload below or macis::HipSparseMatrixOperator.
#pragma once
#include <cstdint>
#include <vector>

#include <macis/hip/runtime.hpp>
#include <sparsexx/matrix_types/csr_matrix.hpp>

namespace sparsexx::spblas {

template <typename SpMatType, typename ALPHAT, typename BETAT>
void gespmbv(int64_t K, ALPHAT ALPHA, const SpMatType& A,
             const typename SpMatType::value_type* V, int64_t LDV, BETAT BETA,
             typename SpMatType::value_type* AV, int64_t LDAV) {
  static_assert(std::is_same_v<typename SpMatType::value_type, double>, "FP64 only");
  auto& ctx = macis::hip::default_context();
  std::vector<int64_t> rp(A.rowptr().begin(), A.rowptr().end());
  std::vector<int32_t> ci(A.colind().begin(), A.colind().end());
  mhip_csr* h = nullptr;
  macis::hip::check(mhip_csr_create_from_host(ctx.get(), A.m(), A.n(), A.nnz(), A.indexing(),
                                              rp.data(), ci.data(), A.nzval().data(), &h));
  macis::hip::DeviceCSR dA(h);
  if(!K || !A.m()) return;
  macis::hip::DeviceBuffer<double> dV(ctx, V, (K - 1) * LDV + A.n());
  macis::hip::DeviceBuffer<double> dAV(ctx, AV, (K - 1) * LDAV + A.m());
  macis::hip::check(mhip_spmv(ctx.get(), dA.get(), K, ALPHA, dV.data(), LDV, BETA, dAV.data(), LDAV));
  dAV.download(AV, (K - 1) * LDAV + A.m());
}

// The same product on a matrix already resident in HBM (a DeviceCSR from
// HipHamiltonianGenerator::device_csr or mhip_csr_create_from_host) with
// DEVICE blocks V and AV: nothing is staged, the call runs on the matrix's
// context and returns when the product is done.
template <typename ALPHAT, typename BETAT>
void gespmbv(int64_t K, ALPHAT ALPHA, const macis::hip::DeviceCSR& A, const double* V,
             int64_t LDV, BETAT BETA, double* AV, int64_t LDAV,
             macis::hip::Context& ctx = macis::hip::default_context()) {
  macis::hip::check(mhip_spmv(ctx.get(), A.get(), K, ALPHA, V, LDV, BETA, AV, LDAV));
  macis::hip::check(mhip_synchronize(ctx.get()));
}

}  // namespace sparsexx::spblas
