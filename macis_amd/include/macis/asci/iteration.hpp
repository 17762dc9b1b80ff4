// asci_iter (restating include/macis/asci/iteration.hpp:16-67).
#pragma once
#include <tuple>

#include <macis/asci/determinant_search.hpp>
#include <macis/asci/determinant_sort.hpp>
#include <macis/solvers/selected_ci_diag.hpp>
#include <macis/util/mcscf.hpp>

namespace macis {

template <size_t N, typename index_t = int32_t>
auto asci_iter(ASCISettings asci_settings, MCSCFSettings mcscf_settings, size_t ndets_max,
               double E0, std::vector<wfn_t<N>> wfn, std::vector<double> X,
               HamiltonianGenerator<N>& ham_gen, size_t norb) {
  if(wfn.size() > 1) reorder_ci_on_coeff(wfn, X);
  const size_t nkeep = std::min(asci_settings.ncdets_max, wfn.size());
  wfn = asci_search(asci_settings, ndets_max, wfn.begin(), wfn.begin() + nkeep, E0, X, norb,
                    ham_gen.T(), ham_gen.G_red(), ham_gen.V_red(), ham_gen.G(), ham_gen.V(),
                    ham_gen);
  std::vector<double> X_local;  // fresh guess
  auto E = selected_ci_diag<N, index_t>(wfn.begin(), wfn.end(), ham_gen, mcscf_settings.ci_matel_tol,
                                        mcscf_settings.ci_max_subspace, mcscf_settings.ci_res_tol,
                                        X_local);
  X = std::move(X_local);
  return std::make_tuple(E, wfn, X);
}

}  // namespace macis
