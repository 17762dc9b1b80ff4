// HamiltonianGenerator<N> (restating include/macis/hamiltonian_generator.hpp
// :17-188 and hamiltonian_generator/impl.hpp:14-70, matrix_elements.hpp,
// fast_diagonals.hpp, rdms.hpp). The integral intermediates live here on the
// host so the accessors T(), G_red(), V_red(), G(), V(), the public spans and
// the per-pair members (matrix elements, fast diagonals, RDM contributions)
// behave as in the reference; CSR construction and form_rdms are the virtual
// seam a device generator overrides.
#pragma once
#include <algorithm>
#include <bitset>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <vector>

#include <macis/bitset_operations.hpp>
#include <macis/sd_operations.hpp>
#include <macis/types.hpp>
#include <macis/util/rdms.hpp>
#include <macis/util/syev.hpp>
#include <macis/util/transform.hpp>
#include <sparsexx/matrix_types/csr_matrix.hpp>

namespace macis {

template <size_t N = 64>
class HamiltonianGenerator {
  static_assert(N % 2 == 0, "N Must Be Even");

 public:
  constexpr static size_t nbits = N;
  using full_det_t = std::bitset<N>;
  using spin_det_t = std::bitset<N / 2>;
  template <typename index_t>
  using sparse_matrix_type = sparsexx::csr_matrix<double, index_t>;
  using full_det_iterator = wavefunction_iterator_t<N>;
  using matrix_span_t = matrix_span<double>;
  using rank3_span_t = rank3_span<double>;
  using rank4_span_t = rank4_span<double>;

  inline spin_det_t alpha_string(full_det_t str) { return bitset_lo_word(str); }
  inline spin_det_t beta_string(full_det_t str) { return bitset_hi_word(str); }

  size_t norb_ = 0, norb2_ = 0, norb3_ = 0;
  matrix_span_t T_pq_;
  rank4_span_t V_pqrs_;
  // G(i,j,k,l) = V(i,j,k,l) - V(i,l,k,j)
  std::vector<double> G_pqrs_data_;
  rank4_span_t G_pqrs_;
  // G_red(k,i,j) = G(k,k,i,j)
  std::vector<double> G_red_data_;
  rank3_span_t G_red_;
  // V_red(k,i,j) = V(k,k,i,j)
  std::vector<double> V_red_data_;
  rank3_span_t V_red_;
  // G2_red(i,j) = G(i,i,j,j) / 2
  std::vector<double> G2_red_data_;
  matrix_span_t G2_red_;
  // V2_red(i,j) = V(i,i,j,j)
  std::vector<double> V2_red_data_;
  matrix_span_t V2_red_;

  HamiltonianGenerator(matrix_span_t T, rank4_span_t V)
      : norb_(T.extent(0)), norb2_(norb_ * norb_), norb3_(norb2_ * norb_), T_pq_(T), V_pqrs_(V) {
    generate_integral_intermediates(V_pqrs_);
  }
  virtual ~HamiltonianGenerator() noexcept = default;

  // The intermediates above from V (impl.hpp:25-70). Calls
  // integrals_changed() so device copies refresh.
  void generate_integral_intermediates(rank4_span_t V) {
    if(V.extent(0) != norb_ || V.extent(1) != norb_ || V.extent(2) != norb_ ||
       V.extent(3) != norb_)
      throw std::runtime_error("V has incorrect dimensions");
    const size_t n = norb_, n2 = norb2_, n3 = norb3_;
    G_pqrs_data_.assign(n3 * n, 0.);
    for(size_t l = 0; l < n; ++l)
      for(size_t k = 0; k < n; ++k)
        for(size_t j = 0; j < n; ++j)
          for(size_t i = 0; i < n; ++i) G_pqrs_data_[i + n * j + n2 * k + n3 * l] = V(i, j, k, l) - V(i, l, k, j);
    G_red_data_.assign(n3, 0.);
    V_red_data_.assign(n3, 0.);
    for(size_t j = 0; j < n; ++j)
      for(size_t i = 0; i < n; ++i)
        for(size_t k = 0; k < n; ++k) {
          G_red_data_[k + n * i + n2 * j] = G_pqrs_data_[k + n * k + n2 * i + n3 * j];
          V_red_data_[k + n * i + n2 * j] = V(k, k, i, j);
        }
    G2_red_data_.assign(n2, 0.);
    V2_red_data_.assign(n2, 0.);
    for(size_t j = 0; j < n; ++j)
      for(size_t i = 0; i < n; ++i) {
        G2_red_data_[i + n * j] = 0.5 * G_pqrs_data_[i + n * i + n2 * j + n3 * j];
        V2_red_data_[i + n * j] = V(i, i, j, j);
      }
    G_pqrs_ = rank4_span_t(G_pqrs_data_.data(), n, n, n, n);
    G_red_ = rank3_span_t(G_red_data_.data(), n, n, n);
    V_red_ = rank3_span_t(V_red_data_.data(), n, n, n);
    G2_red_ = matrix_span_t(G2_red_data_.data(), n, n);
    V2_red_ = matrix_span_t(V2_red_data_.data(), n, n);
    integrals_changed();
  }

  const double* T() const { return T_pq_.data_handle(); }
  double* T() { return T_pq_.data_handle(); }
  const double* G_red() const { return G_red_data_.data(); }
  const double* V_red() const { return V_red_data_.data(); }
  const double* G() const { return G_pqrs_data_.data(); }
  const double* V() const { return V_pqrs_.data_handle(); }
  double* V() { return V_pqrs_.data_handle(); }
  size_t norb() const { return norb_; }

  // ---- Slater-Condon rules (matrix_elements.hpp:14-135) -------------------
  // Same-spin double: sign * G(v1,o1,v2,o2).
  double matrix_element_4(spin_det_t bra, spin_det_t ket, spin_det_t ex) const {
    const auto [o1, v1, o2, v2, sign] = doubles_sign_indices(bra, ket, ex);
    return sign * G_pqrs_data_[v1 + norb_ * o1 + norb2_ * v2 + norb3_ * o2];
  }

  // Opposite-spin double: sign_a * sign_b * V(v1,o1,v2,o2).
  double matrix_element_22(spin_det_t bra_alpha, spin_det_t ket_alpha, spin_det_t ex_alpha,
                           spin_det_t bra_beta, spin_det_t ket_beta, spin_det_t ex_beta) const {
    const auto [o1, v1, sa] = single_excitation_sign_indices(bra_alpha, ket_alpha, ex_alpha);
    const auto [o2, v2, sb] = single_excitation_sign_indices(bra_beta, ket_beta, ex_beta);
    return (sa * sb) * V_pqrs_(v1, o1, v2, o2);
  }

  // Single in the spin of bra/ket: sign * (T(v,o) + sum over that spin's
  // occupied p of G_red(p,v,o) + sum over the other spin's of V_red(p,v,o)).
  double matrix_element_2(spin_det_t bra, spin_det_t ket, spin_det_t ex,
                          const std::vector<uint32_t>& bra_occ_alpha,
                          const std::vector<uint32_t>& bra_occ_beta) const {
    const auto [o1, v1, sign] = single_excitation_sign_indices(bra, ket, ex);
    double h = T_pq_(v1, o1);
    const double* g = G_red_data_.data() + norb_ * v1 + norb2_ * o1;
    for(auto p : bra_occ_alpha) h += g[p];
    const double* v = V_red_data_.data() + norb_ * v1 + norb2_ * o1;
    for(auto p : bra_occ_beta) h += v[p];
    return sign * h;
  }

  // Diagonal: one-body terms, then alpha-alpha and beta-beta G2_red, then
  // alpha-beta V2_red, in that order.
  double matrix_element_diag(const std::vector<uint32_t>& occ_alpha,
                             const std::vector<uint32_t>& occ_beta) const {
    double h = 0.;
    for(auto p : occ_alpha) h += T_pq_(p, p);
    for(auto p : occ_beta) h += T_pq_(p, p);
    for(auto q : occ_alpha)
      for(auto p : occ_alpha) h += G2_red_data_[p + norb_ * q];
    for(auto q : occ_beta)
      for(auto p : occ_beta) h += G2_red_data_[p + norb_ * q];
    for(auto q : occ_beta)
      for(auto p : occ_alpha) h += V2_red_data_[p + norb_ * q];
    return h;
  }

  // Dispatch on the excitation counts of the two spin strings.
  double matrix_element(spin_det_t bra_alpha, spin_det_t ket_alpha, spin_det_t ex_alpha,
                        spin_det_t bra_beta, spin_det_t ket_beta, spin_det_t ex_beta,
                        const std::vector<uint32_t>& bra_occ_alpha,
                        const std::vector<uint32_t>& bra_occ_beta) const {
    const size_t ca = ex_alpha.count(), cb = ex_beta.count();
    if(ca + cb > 4) return 0.;
    if(ca == 4) return matrix_element_4(bra_alpha, ket_alpha, ex_alpha);
    if(cb == 4) return matrix_element_4(bra_beta, ket_beta, ex_beta);
    if(ca == 2 && cb == 2)
      return matrix_element_22(bra_alpha, ket_alpha, ex_alpha, bra_beta, ket_beta, ex_beta);
    if(ca == 2) return matrix_element_2(bra_alpha, ket_alpha, ex_alpha, bra_occ_alpha, bra_occ_beta);
    if(cb == 2) return matrix_element_2(bra_beta, ket_beta, ex_beta, bra_occ_beta, bra_occ_alpha);
    return matrix_element_diag(bra_occ_alpha, bra_occ_beta);
  }

  double matrix_element(full_det_t bra, full_det_t ket) const {
    const auto ba = bitset_lo_word(bra), ka = bitset_lo_word(ket);
    const auto bb = bitset_hi_word(bra), kb = bitset_hi_word(ket);
    return matrix_element(ba, ka, ba ^ ka, bb, kb, bb ^ kb, bits_to_indices(ba),
                          bits_to_indices(bb));
  }

  // ---- orbital energies and fast diagonals (fast_diagonals.hpp:14-127) ---
  // e_p = T(p,p) + sum_q in ss_occ (G2_red(p,q) + G2_red(q,p)) - G2_red(p,p)
  //       + sum_q in os_occ V2_red(p,q)
  double single_orbital_en(uint32_t orb, const std::vector<uint32_t>& ss_occ,
                           const std::vector<uint32_t>& os_occ) const {
    double e = T_pq_(orb, orb);
    for(auto q : ss_occ) e += G2_red_(orb, q) + G2_red_(q, orb);
    e -= G2_red_(orb, orb);
    for(auto q : os_occ) e += V2_red_(orb, q);
    return e;
  }

  std::vector<double> single_orbital_ens(size_t norb, const std::vector<uint32_t>& ss_occ,
                                         const std::vector<uint32_t>& os_occ) const {
    std::vector<double> ens(norb);
    for(size_t i = 0; i < norb; ++i) ens[i] = single_orbital_en(uint32_t(i), ss_occ, os_occ);
    return ens;
  }

  // Diagonal of the determinant with orb_hol -> orb_par applied, from the
  // parent's diagonal orig_det_Hii and orbital energies.
  double fast_diag_single(double hol_en, double par_en, uint32_t orb_hol, uint32_t orb_par,
                          double orig_det_Hii) const {
    return orig_det_Hii + par_en - hol_en - G2_red_(orb_par, orb_hol) - G2_red_(orb_hol, orb_par);
  }
  double fast_diag_single(const std::vector<uint32_t>& ss_occ,
                          const std::vector<uint32_t>& os_occ, uint32_t orb_hol,
                          uint32_t orb_par, double orig_det_Hii) const {
    return fast_diag_single(single_orbital_en(orb_hol, ss_occ, os_occ),
                            single_orbital_en(orb_par, ss_occ, os_occ), orb_hol, orb_par,
                            orig_det_Hii);
  }

  // Same-spin double (h1, h2) -> (p1, p2).
  double fast_diag_ss_double(double en_hol1, double en_hol2, double en_par1, double en_par2,
                             uint32_t h1, uint32_t h2, uint32_t p1, uint32_t p2,
                             double orig_det_Hii) const {
    const auto& g = G2_red_;
    return orig_det_Hii + en_par1 + en_par2 - en_hol1 - en_hol2 + g(h1, h2) + g(h2, h1) +
           g(p1, p2) + g(p2, p1) - g(p1, h1) - g(h1, p1) - g(p2, h1) - g(h1, p2) - g(p1, h2) -
           g(h2, p1) - g(p2, h2) - g(h2, p2);
  }
  double fast_diag_ss_double(const std::vector<uint32_t>& ss_occ,
                             const std::vector<uint32_t>& os_occ, uint32_t h1, uint32_t h2,
                             uint32_t p1, uint32_t p2, double orig_det_Hii) const {
    const double e_h1 = single_orbital_en(h1, ss_occ, os_occ);
    const double e_h2 = single_orbital_en(h2, ss_occ, os_occ);
    const double e_p1 = single_orbital_en(p1, ss_occ, os_occ);
    const double e_p2 = single_orbital_en(p2, ss_occ, os_occ);
    return fast_diag_ss_double(e_h1, e_h2, e_p1, e_p2, h1, h2, p1, p2, orig_det_Hii);
  }

  // Opposite-spin double: up hu -> pu, down hd -> pd.
  double fast_diag_os_double(double en_holu, double en_hold, double en_paru, double en_pard,
                             uint32_t hu, uint32_t hd, uint32_t pu, uint32_t pd,
                             double orig_det_Hii) const {
    const auto &g = G2_red_, &v = V2_red_;
    return orig_det_Hii + en_paru + en_pard - en_holu - en_hold + v(hu, hd) + v(pu, pd) -
           g(pu, hu) - g(hu, pu) - g(pd, hd) - g(hd, pd) - v(pu, hd) - v(hu, pd);
  }
  double fast_diag_os_double(const std::vector<uint32_t>& up_occ,
                             const std::vector<uint32_t>& do_occ, uint32_t hu, uint32_t hd,
                             uint32_t pu, uint32_t pd, double orig_det_Hii) const {
    // the down orbitals see the down string as their own spin
    const double e_hu = single_orbital_en(hu, up_occ, do_occ);
    const double e_hd = single_orbital_en(hd, do_occ, up_occ);
    const double e_pu = single_orbital_en(pu, up_occ, do_occ);
    const double e_pd = single_orbital_en(pd, do_occ, up_occ);
    return fast_diag_os_double(e_hu, e_hd, e_pu, e_pd, hu, hd, pu, pd, orig_det_Hii);
  }

  // ---- RDM contributions of one pair (hamiltonian_generator.hpp:157-177,
  // implemented by util/rdms.hpp:15-132) -----------------------------------
  void rdm_contributions_4(spin_det_t bra, spin_det_t ket, spin_det_t ex, double val,
                           rank4_span_t trdm) {
    macis::rdm_contributions_4(bra, ket, ex, val, trdm);
  }
  void rdm_contributions_22(spin_det_t bra_alpha, spin_det_t ket_alpha, spin_det_t ex_alpha,
                            spin_det_t bra_beta, spin_det_t ket_beta, spin_det_t ex_beta,
                            double val, rank4_span_t trdm) {
    macis::rdm_contributions_22(bra_alpha, ket_alpha, ex_alpha, bra_beta, ket_beta, ex_beta, val,
                                trdm);
  }
  void rdm_contributions_2(spin_det_t bra, spin_det_t ket, spin_det_t ex,
                           const std::vector<uint32_t>& bra_occ_alpha,
                           const std::vector<uint32_t>& bra_occ_beta, double val,
                           matrix_span_t ordm, rank4_span_t trdm) {
    macis::rdm_contributions_2(bra, ket, ex, bra_occ_alpha, bra_occ_beta, val, ordm, trdm);
  }
  void rdm_contributions_diag(const std::vector<uint32_t>& occ_alpha,
                              const std::vector<uint32_t>& occ_beta, double val,
                              matrix_span_t ordm, rank4_span_t trdm) {
    macis::rdm_contributions_diag(occ_alpha, occ_beta, val, ordm, trdm);
  }
  void rdm_contributions(spin_det_t bra_alpha, spin_det_t ket_alpha, spin_det_t ex_alpha,
                         spin_det_t bra_beta, spin_det_t ket_beta, spin_det_t ex_beta,
                         const std::vector<uint32_t>& bra_occ_alpha,
                         const std::vector<uint32_t>& bra_occ_beta, double val,
                         matrix_span_t ordm, rank4_span_t trdm) {
    macis::rdm_contributions(bra_alpha, ket_alpha, ex_alpha, bra_beta, ket_beta, ex_beta,
                             bra_occ_alpha, bra_occ_beta, val, ordm, trdm);
  }

  // Rotates T and V in place to the natural orbitals of the 1-RDM `ordm`
  // (norb x norb, column-major) and regenerates the intermediates
  // (hamiltonian_generator/rdms.hpp:19-107). The reference takes the left
  // singular vectors of ordm in descending singular-value order; for the
  // symmetric ordm these are its eigenvectors ordered by descending
  // |eigenvalue|, taken here from the Jacobi eigensolver (vector signs may
  // differ from LAPACK's, which no energy depends on).
  void rotate_hamiltonian_ordm(const double* ordm) {
    const size_t n = norb_;
    std::vector<double> A(ordm, ordm + n * n), w(n);
    syev_lower(n, A.data(), n, w.data());
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), size_t(0));
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return std::abs(w[a]) > std::abs(w[b]); });
    std::vector<double> U(n * n);
    for(size_t c = 0; c < n; ++c)
      std::copy_n(A.data() + order[c] * n, n, U.data() + c * n);
    two_index_transform(n, n, T(), n, U.data(), n, T(), n);
    four_index_transform(n, n, V(), n, U.data(), n, V(), n);
    generate_integral_intermediates(V_pqrs_);
  }

  // Front of the virtual CSR seam (hamiltonian_generator.hpp:142-156).
  template <typename index_t>
  sparse_matrix_type<index_t> make_csr_hamiltonian_block(full_det_iterator bra_begin,
                                                         full_det_iterator bra_end,
                                                         full_det_iterator ket_begin,
                                                         full_det_iterator ket_end,
                                                         double H_thresh) {
    if constexpr(std::is_same_v<index_t, int32_t>)
      return make_csr_hamiltonian_block_32bit_(bra_begin, bra_end, ket_begin, ket_end, H_thresh);
    else if constexpr(std::is_same_v<index_t, int64_t>)
      return make_csr_hamiltonian_block_64bit_(bra_begin, bra_end, ket_begin, ket_end, H_thresh);
    else
      throw std::runtime_error("Unsupported index_t");
  }

  virtual sparse_matrix_type<int32_t> make_csr_hamiltonian_block_32bit_(
      full_det_iterator, full_det_iterator, full_det_iterator, full_det_iterator, double) = 0;
  virtual sparse_matrix_type<int64_t> make_csr_hamiltonian_block_64bit_(
      full_det_iterator, full_det_iterator, full_det_iterator, full_det_iterator, double) = 0;
  virtual void form_rdms(full_det_iterator, full_det_iterator, full_det_iterator,
                         full_det_iterator, double* C, matrix_span_t ordm,
                         rank4_span_t trdm) = 0;
  virtual void SetJustSingles(bool) {}
  virtual bool GetJustSingles() { return false; }
  virtual size_t GetNimp() const { return N / 2; }

 protected:
  // Hook for generators that mirror the integrals elsewhere.
  virtual void integrals_changed() {}
};

}  // namespace macis
