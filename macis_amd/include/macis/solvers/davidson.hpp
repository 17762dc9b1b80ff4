// Davidson API (restating include/macis/solvers/davidson.hpp:29-225). The
// eigensolver runs on the device: a host csr_matrix behind
// SparseMatrixOperator is uploaded once per solve, HipSparseMatrixOperator
// keeps the matrix in HBM. X is the guess in and the Ritz vector out.
#pragma once
#include <algorithm>
#include <utility>

#include <macis/hip/runtime.hpp>
#include <sparsexx/matrix_types/csr_matrix.hpp>
#include <sparsexx/spblas/spmbv.hpp>
#include <sparsexx/util/submatrix.hpp>

namespace macis {

// Operator over a host CSR (davidson.hpp:29-59).
template <typename SpMatType>
class SparseMatrixOperator {
 public:
  using index_type = typename SpMatType::index_type;
  explicit SparseMatrixOperator(const SpMatType& m) : m_matrix_(m) {}
  void operator_action(size_t m, double alpha, const double* V, size_t LDV, double beta,
                       double* AV, size_t LDAV) const {
    sparsexx::spblas::gespmbv(m, alpha, m_matrix_, V, LDV, beta, AV, LDAV);
  }
  const SpMatType& matrix() const { return m_matrix_; }

 private:
  const SpMatType& m_matrix_;
};

// Operator over a device-resident CSR (e.g. HipHamiltonianGenerator::device_csr).
class HipSparseMatrixOperator {
 public:
  HipSparseMatrixOperator(hip::Context& ctx, const hip::DeviceCSR& A) : ctx_(ctx), A_(A) {}
  // V / AV are DEVICE pointers here.
  void operator_action(size_t m, double alpha, const double* V, size_t LDV, double beta,
                       double* AV, size_t LDAV) const {
    hip::check(mhip_spmv(ctx_.get(), A_.get(), m, alpha, V, LDV, beta, AV, LDAV));
  }
  hip::Context& context() const { return ctx_; }
  const hip::DeviceCSR& matrix() const { return A_; }

 private:
  hip::Context& ctx_;
  const hip::DeviceCSR& A_;
};

// Unit vector at the first minimum of the diagonal (davidson.hpp:61-68).
template <typename SpMatType>
void diagonal_guess(size_t N, const SpMatType& A, double* X) {
  auto D = sparsexx::extract_diagonal_elements(A);
  auto it = std::min_element(D.begin(), D.end());
  X[std::distance(D.begin(), it)] = 1.;
}

namespace detail {
template <typename index_t>
hip::DeviceCSR upload_csr(hip::Context& ctx, const sparsexx::csr_matrix<double, index_t>& A) {
  std::vector<int64_t> rp(A.rowptr().begin(), A.rowptr().end());
  std::vector<int32_t> ci(A.colind().begin(), A.colind().end());
  mhip_csr* h = nullptr;
  hip::check(mhip_csr_create_from_host(ctx.get(), A.m(), A.n(), A.nnz(), A.indexing(),
                                       rp.data(), ci.data(), A.nzval().data(), &h));
  return hip::DeviceCSR(h);
}
}  // namespace detail

// davidson(N, max_m, op, D, tol, X) on the device with a device operator;
// D and X are DEVICE pointers. Returns (iterations, LAM(0)); throws
// "Davidson Did Not Converge!" like the reference.
inline std::pair<int64_t, double> davidson(int64_t N, int64_t max_m,
                                           const HipSparseMatrixOperator& op, const double* D,
                                           double tol, double* X) {
  (void)N;
  double E = 0.;
  int64_t it = 0;
  hip::check(mhip_davidson(op.context().get(), op.matrix().get(), max_m,
                           hip::davidson_options().max_restarts, tol, D, X, &E, &it));
  return {it, E};
}

// davidson over a host CSR operator: D and X are HOST pointers, the solve
// itself runs on the device (default context).
template <typename SpMatType>
std::pair<int64_t, double> davidson(int64_t N, int64_t max_m,
                                    const SparseMatrixOperator<SpMatType>& op, const double* D,
                                    double tol, double* X) {
  if(!X) throw std::runtime_error("Davidson: No Guess Provided");
  auto& ctx = hip::default_context();
  auto A = detail::upload_csr(ctx, op.matrix());
  hip::DeviceBuffer<double> dD(ctx, D, N), dX(ctx, X, N);
  double E = 0.;
  int64_t it = 0;
  hip::check(mhip_davidson(ctx.get(), A.get(), max_m, hip::davidson_options().max_restarts, tol,
                           dD.data(), dX.data(), &E, &it));
  dX.download(X, N);
  return {it, E};
}

}  // namespace macis
