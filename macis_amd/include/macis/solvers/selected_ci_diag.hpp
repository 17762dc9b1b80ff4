// selected_ci_diag (restating include/macis/solvers/selected_ci_diag.hpp
// :75-207, serial path). With a HipHamiltonianGenerator the Hamiltonian is
// built, the guess formed and Davidson run entirely on the device
// (mhip_selected_ci_diag); C_local is the guess in (reused when
// max|C| > 1/size) and the eigenvector out.
#pragma once
#include <vector>

#include <macis/csr_hamiltonian.hpp>
#include <macis/hip/hip_hamiltonian_generator.hpp>
#include <macis/solvers/davidson.hpp>

namespace macis {

template <typename SpMatType>
double serial_selected_ci_diag(const SpMatType& H, size_t davidson_max_m,
                               double davidson_res_tol, std::vector<double>& C) {
  C.resize(H.m(), 0);
  auto D = sparsexx::extract_diagonal_elements(H);
  double max_c = 0.;
  for(auto c : C) max_c = std::max(max_c, std::abs(c));
  if(!(max_c > 1. / C.size())) diagonal_guess(C.size(), H, C.data());
  auto [niter, E] = davidson(H.m(), davidson_max_m, SparseMatrixOperator(H), D.data(),
                             davidson_res_tol, C.data());
  (void)niter;
  return E;
}

template <size_t N, typename index_t = int32_t>
double selected_ci_diag(wavefunction_iterator_t<N> dets_begin, wavefunction_iterator_t<N> dets_end,
                        HamiltonianGenerator<N>& ham_gen, double h_el_tol, size_t davidson_max_m,
                        double davidson_res_tol, std::vector<double>& C_local,
                        const bool quiet = false) {
  (void)quiet;
  const size_t ndets = std::distance(dets_begin, dets_end);
  if(auto* hg = dynamic_cast<HipHamiltonianGenerator<N>*>(&ham_gen)) {
    auto& ctx = hg->context();
    C_local.resize(ndets, 0);
    auto dd = hip::upload_dets<N>(ctx, dets_begin, dets_end);
    hip::DeviceBuffer<double> dC(ctx, C_local.data(), std::max<size_t>(ndets, 1));
    double E = 0.;
    int64_t nnz = 0;
    hip::check(mhip_selected_ci_diag(ctx.get(), hip::words_per_det(N), dd.data(), ndets, h_el_tol,
                                     davidson_max_m, hip::davidson_options().max_restarts,
                                     davidson_res_tol, dC.data(), &E, &nnz, nullptr));
    dC.download(C_local.data(), ndets);
    return E;
  }
  // Any other generator: its own CSR build, then the device Davidson.
  auto H = make_csr_hamiltonian<index_t>(dets_begin, dets_end, ham_gen, h_el_tol);
  return serial_selected_ci_diag(H, davidson_max_m, davidson_res_tol, C_local);
}

}  // namespace macis
