// RDM contributions of one determinant pair (restating the reference's
// include/macis/util/rdms.hpp:15-132). `val` is C_bra * C_ket; the spin-summed
// 1-RDM ordm(p,q) and 2-RDM trdm(p,q,r,s) are column-major spans that these
// functions add into. A null 2-RDM span (data_handle() == nullptr) forms
// the 1-RDM alone; the dispatcher then sends same-spin doubles on to the
// diagonal branch, as the reference's branch order does.
#pragma once
#include <bitset>
#include <cstdint>
#include <vector>

#include <macis/sd_operations.hpp>
#include <macis/types.hpp>

namespace macis {

// Same-spin double (o1,v1,o2,v2): the four antisymmetric placements.
template <typename T, size_t N>
inline void rdm_contributions_4(std::bitset<N> bra, std::bitset<N> ket, std::bitset<N> ex, T val,
                                rank4_span<T> trdm) {
  const auto [o1, v1, o2, v2, sign] = doubles_sign_indices(bra, ket, ex);
  const T w = val * (sign * 0.5);
  trdm(v1, o1, v2, o2) += w;
  trdm(v2, o1, v1, o2) -= w;
  trdm(v1, o2, v2, o1) -= w;
  trdm(v2, o2, v1, o1) += w;
}

// Opposite-spin double: alpha (o1,v1), beta (o2,v2).
template <typename T, size_t N>
inline void rdm_contributions_22(std::bitset<N> bra_alpha, std::bitset<N> ket_alpha,
                                 std::bitset<N> ex_alpha, std::bitset<N> bra_beta,
                                 std::bitset<N> ket_beta, std::bitset<N> ex_beta, T val,
                                 rank4_span<T> trdm) {
  const auto [o1, v1, sa] = single_excitation_sign_indices(bra_alpha, ket_alpha, ex_alpha);
  const auto [o2, v2, sb] = single_excitation_sign_indices(bra_beta, ket_beta, ex_beta);
  const T w = val * ((sa * sb) * 0.5);
  trdm(v1, o1, v2, o2) += w;
  trdm(v2, o2, v1, o1) += w;
}

// Single o1 -> v1 in one spin; bra_occ_alpha is the occupied list of that
// spin, bra_occ_beta of the other.
template <typename T, size_t N, typename IndexType>
inline void rdm_contributions_2(std::bitset<N> bra, std::bitset<N> ket, std::bitset<N> ex,
                                const IndexType& bra_occ_alpha, const IndexType& bra_occ_beta,
                                T val, matrix_span<T> ordm, rank4_span<T> trdm) {
  const auto [o1, v1, sign] = single_excitation_sign_indices(bra, ket, ex);
  ordm(v1, o1) += sign * val;
  if(!trdm.data_handle()) return;
  const T w = val * (sign * 0.5);
  for(auto p : bra_occ_alpha) {
    trdm(v1, o1, p, p) += w;
    trdm(p, p, v1, o1) += w;
    trdm(v1, p, p, o1) -= w;
    trdm(p, o1, v1, p) -= w;
  }
  for(auto p : bra_occ_beta) {
    trdm(v1, o1, p, p) += w;
    trdm(p, p, v1, o1) += w;
  }
}

// Diagonal pair (bra == ket).
template <typename T, typename IndexType>
inline void rdm_contributions_diag(const IndexType& occ_alpha, const IndexType& occ_beta, T val,
                                   matrix_span<T> ordm, rank4_span<T> trdm) {
  for(auto p : occ_alpha) ordm(p, p) += val;
  for(auto p : occ_beta) ordm(p, p) += val;
  if(!trdm.data_handle()) return;
  const T w = val * 0.5;
  for(auto q : occ_alpha)
    for(auto p : occ_alpha) {
      trdm(p, p, q, q) += w;
      trdm(p, q, p, q) -= w;
    }
  for(auto q : occ_beta)
    for(auto p : occ_beta) {
      trdm(p, p, q, q) += w;
      trdm(p, q, p, q) -= w;
    }
  for(auto q : occ_beta)
    for(auto p : occ_alpha) {
      trdm(p, p, q, q) += w;
      trdm(q, q, p, p) += w;
    }
}

// Dispatch on the excitation counts (util/rdms.hpp:100-132).
template <typename T, size_t N, typename IndexType>
inline void rdm_contributions(std::bitset<N> bra_alpha, std::bitset<N> ket_alpha,
                              std::bitset<N> ex_alpha, std::bitset<N> bra_beta,
                              std::bitset<N> ket_beta, std::bitset<N> ex_beta,
                              const IndexType& bra_occ_alpha, const IndexType& bra_occ_beta,
                              T val, matrix_span<T> ordm, rank4_span<T> trdm) {
  const size_t ca = ex_alpha.count(), cb = ex_beta.count();
  if(ca + cb > 4) return;
  const bool two = trdm.data_handle() != nullptr;
  if(ca == 4 && two)
    rdm_contributions_4(bra_alpha, ket_alpha, ex_alpha, val, trdm);
  else if(cb == 4 && two)
    rdm_contributions_4(bra_beta, ket_beta, ex_beta, val, trdm);
  else if(ca == 2 && cb == 2 && two)
    rdm_contributions_22(bra_alpha, ket_alpha, ex_alpha, bra_beta, ket_beta, ex_beta, val, trdm);
  else if(ca == 2)
    rdm_contributions_2(bra_alpha, ket_alpha, ex_alpha, bra_occ_alpha, bra_occ_beta, val, ordm,
                        trdm);
  else if(cb == 2)
    rdm_contributions_2(bra_beta, ket_beta, ex_beta, bra_occ_beta, bra_occ_alpha, val, ordm, trdm);
  else
    rdm_contributions_diag(bra_occ_alpha, bra_occ_beta, val, ordm, trdm);
}

}  // namespace macis
