// asci_refine (restating include/macis/asci/refine.hpp:14-68).
#pragma once
#include <macis/asci/iteration.hpp>

namespace macis {

template <size_t N, typename index_t = int32_t>
auto asci_refine(ASCISettings asci_settings, MCSCFSettings mcscf_settings, double E0,
                 std::vector<wfn_t<N>> wfn, std::vector<double> X,
                 HamiltonianGenerator<N>& ham_gen, size_t norb) {
  const size_t ndets = wfn.size();
  bool converged = false;
  for(size_t iter = 0; iter < asci_settings.max_refine_iter; ++iter) {
    double E;
    std::tie(E, wfn, X) = asci_iter<N, index_t>(asci_settings, mcscf_settings, ndets, E0,
                                                std::move(wfn), std::move(X), ham_gen, norb);
    if(wfn.size() != ndets)
      throw std::runtime_error("Wavefunction size can't change in refinement");
    const double dE = E - E0;
    E0 = E;
    if(std::abs(dE) < asci_settings.refine_energy_tol) {
      converged = true;
      break;
    }
  }
  if(!converged) throw std::runtime_error("ACCI Refine did not converge");
  return std::make_tuple(E0, wfn, X);
}

}  // namespace macis
