// asci_grow (restating include/macis/asci/grow.hpp:16-173). Natural-orbital
// rotation (grow_with_rot) needs RDMs, which are outside the accelerated
// path, and is rejected.
#pragma once
#include <macis/asci/iteration.hpp>

namespace macis {

template <size_t N, typename index_t = int32_t>
auto asci_grow(ASCISettings asci_settings, MCSCFSettings mcscf_settings, double E0,
               std::vector<wfn_t<N>> wfn, std::vector<double> X, HamiltonianGenerator<N>& ham_gen,
               size_t norb) {
  if(asci_settings.grow_with_rot)
    throw std::runtime_error("asci_grow: grow_with_rot is not supported");
  while(wfn.size() < asci_settings.ntdets_max) {
    size_t ndets_new = std::min(std::max(asci_settings.ntdets_min,
                                         wfn.size() * asci_settings.grow_factor),
                                asci_settings.ntdets_max);
    double E;
    std::tie(E, wfn, X) = asci_iter<N, index_t>(asci_settings, mcscf_settings, ndets_new, E0,
                                                std::move(wfn), std::move(X), ham_gen, norb);
    if(ndets_new > wfn.size()) throw std::runtime_error("Wavefunction didn't grow enough...");
    E0 = E;
  }
  return std::make_tuple(E0, wfn, X);
}

}  // namespace macis
