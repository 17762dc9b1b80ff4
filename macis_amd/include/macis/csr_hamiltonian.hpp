// make_csr_hamiltonian[_block] (restating include/macis/csr_hamiltonian.hpp
// :22-44, serial path).
#pragma once
#include <macis/hamiltonian_generator.hpp>

namespace macis {

template <typename index_t, size_t N>
sparsexx::csr_matrix<double, index_t> make_csr_hamiltonian_block(
    wavefunction_iterator_t<N> bra_begin, wavefunction_iterator_t<N> bra_end,
    wavefunction_iterator_t<N> ket_begin, wavefunction_iterator_t<N> ket_end,
    HamiltonianGenerator<N>& ham_gen, double H_thresh) {
  const size_t nbra = std::distance(bra_begin, bra_end);
  const size_t nket = std::distance(ket_begin, ket_end);
  if(nbra && nket)
    return ham_gen.template make_csr_hamiltonian_block<index_t>(bra_begin, bra_end, ket_begin,
                                                                ket_end, H_thresh);
  return sparsexx::csr_matrix<double, index_t>(nbra, nket, 0, 0);
}

template <typename index_t, size_t N>
sparsexx::csr_matrix<double, index_t> make_csr_hamiltonian(wavefunction_iterator_t<N> sd_begin,
                                                           wavefunction_iterator_t<N> sd_end,
                                                           HamiltonianGenerator<N>& ham_gen,
                                                           double H_thresh) {
  return make_csr_hamiltonian_block<index_t>(sd_begin, sd_end, sd_begin, sd_end, ham_gen,
                                             H_thresh);
}

}  // namespace macis
