// asci_grow (restating include/macis/asci/grow.hpp:16-173), with the
// natural-orbital rotation of grow_with_rot (:67-160): after an asci_iter
// that leaves at least rot_size_start determinants, the RDMs of the
// wavefunction are formed (form_rdms, on the device for
// HipHamiltonianGenerator), T and V are rotated in place to the natural
// orbitals (eigenvectors of -ordm: occupations descending), the integral
// intermediates are regenerated (which refreshes a device copy), and the same
// determinants are re-diagonalised; E0 keeps the asci_iter energy (:162).
#pragma once
#include <macis/asci/iteration.hpp>
#include <macis/solvers/selected_ci_diag.hpp>
#include <macis/util/syev.hpp>
#include <macis/util/transform.hpp>

namespace macis {

template <size_t N, typename index_t = int32_t>
auto asci_grow(ASCISettings asci_settings, MCSCFSettings mcscf_settings, double E0,
               std::vector<wfn_t<N>> wfn, std::vector<double> X, HamiltonianGenerator<N>& ham_gen,
               size_t norb) {
  while(wfn.size() < asci_settings.ntdets_max) {
    size_t ndets_new = std::min(std::max(asci_settings.ntdets_min,
                                         wfn.size() * asci_settings.grow_factor),
                                asci_settings.ntdets_max);
    double E;
    std::tie(E, wfn, X) = asci_iter<N, index_t>(asci_settings, mcscf_settings, ndets_new, E0,
                                                std::move(wfn), std::move(X), ham_gen, norb);
    if(ndets_new > wfn.size()) throw std::runtime_error("Wavefunction didn't grow enough...");
    if(asci_settings.grow_with_rot and wfn.size() >= asci_settings.rot_size_start) {
      std::vector<double> ordm(norb * norb, 0.0), trdm(norb * norb * norb * norb, 0.0);
      ham_gen.form_rdms(wfn.begin(), wfn.end(), wfn.begin(), wfn.end(), X.data(),
                        matrix_span<double>(ordm.data(), norb, norb),
                        rank4_span<double>(trdm.data(), norb, norb, norb, norb));
      std::vector<double> ONS(norb);
      for(auto& x : ordm) x *= -1.0;
      syev_lower(norb, ordm.data(), norb, ONS.data());
      for(auto& x : ONS) x *= -1.0;
      two_index_transform(norb, norb, ham_gen.T(), norb, ordm.data(), norb, ham_gen.T(), norb);
      four_index_transform(norb, norb, ham_gen.V(), norb, ordm.data(), norb, ham_gen.V(), norb);
      ham_gen.generate_integral_intermediates(ham_gen.V_pqrs_);
      std::vector<double> X_local;
      selected_ci_diag<N, index_t>(wfn.begin(), wfn.end(), ham_gen, mcscf_settings.ci_matel_tol,
                                   mcscf_settings.ci_max_subspace, mcscf_settings.ci_res_tol,
                                   X_local);
      X = std::move(X_local);
    }
    E0 = E;
  }
  return std::make_tuple(E0, wfn, X);
}

}  // namespace macis
