// HamiltonianGenerator<N> (restating include/macis/hamiltonian_generator.hpp
// :17-188 and hamiltonian_generator/impl.hpp:14-70, matrix_elements.hpp,
// fast_diagonals.hpp). The integral intermediates live here on the host so
// the accessors T(), G_red(), V_red(), G(), V() behave as in the reference;
// CSR construction is the virtual seam a device generator overrides.
#pragma once
#include <bitset>
#include <cmath>
#include <stdexcept>
#include <vector>

#include <macis/bitset_operations.hpp>
#include <macis/types.hpp>
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

  size_t norb_ = 0, norb2_ = 0, norb3_ = 0;
  matrix_span_t T_pq_;
  rank4_span_t V_pqrs_;
  std::vector<double> G_pqrs_data_, G_red_data_, V_red_data_, G2_red_data_, V2_red_data_;

  HamiltonianGenerator(matrix_span_t T, rank4_span_t V)
      : norb_(T.extent(0)), norb2_(norb_ * norb_), norb3_(norb2_ * norb_), T_pq_(T), V_pqrs_(V) {
    generate_integral_intermediates(V_pqrs_);
  }
  virtual ~HamiltonianGenerator() noexcept = default;

  // G(i,j,k,l) = V(i,j,k,l) - V(i,l,k,j); G_red(k,i,j) = G(k,k,i,j);
  // V_red(k,i,j) = V(k,k,i,j); G2_red = G(i,i,j,j)/2; V2_red = V(i,i,j,j)
  // (impl.hpp:25-70). Calls integrals_changed() so device copies refresh.
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

  // <bra|H|ket> with the Slater-Condon rules (matrix_elements.hpp:14-135).
  double matrix_element(full_det_t bra, full_det_t ket) const {
    const auto ba = bitset_lo_word(bra), ka = bitset_lo_word(ket);
    const auto bb = bitset_hi_word(bra), kb = bitset_hi_word(ket);
    const auto ea = ba ^ ka, eb = bb ^ kb;
    const size_t ca = ea.count(), cb = eb.count();
    if(ca + cb > 4) return 0.;
    const size_t n = norb_, n2 = norb2_, n3 = norb3_;
    auto low = [](auto s) { return ffs(s) - 1; };
    auto sign1 = [](auto state, uint32_t p, uint32_t q) {
      const uint32_t lo = std::min(p, q) + 1, hi = std::max(p, q);
      size_t c = 0;
      for(uint32_t i = lo; i < hi; ++i) c += state[i];
      return (c % 2) ? -1. : 1.;
    };
    auto dbl = [&](auto b, auto k, auto e) {
      const uint32_t o1 = low(k & e), v1 = low(b & e);
      double s = sign1(k, v1, o1);
      k.flip(o1).flip(v1);
      e.flip(o1).flip(v1);
      const uint32_t o2 = low(k & e), v2 = low(b & e);
      s *= sign1(k, v2, o2);
      return s * G_pqrs_data_[v1 + n * o1 + n2 * v2 + n3 * o2];
    };
    auto sgl = [&](auto b, auto k, auto e, auto same, auto othr) {
      const uint32_t o1 = low(k & e), v1 = low(b & e);
      const double s = sign1(k, v1, o1);
      double h = T_pq_(v1, o1);
      for(auto p : bits_to_indices(same)) h += G_red_data_[p + n * v1 + n2 * o1];
      for(auto p : bits_to_indices(othr)) h += V_red_data_[p + n * v1 + n2 * o1];
      return s * h;
    };
    if(ca == 4) return dbl(ba, ka, ea);
    if(cb == 4) return dbl(bb, kb, eb);
    if(ca == 2 && cb == 2) {
      const uint32_t o1 = low(ka & ea), v1 = low(ba & ea);
      const uint32_t o2 = low(kb & eb), v2 = low(bb & eb);
      const double s = sign1(ka, v1, o1) * sign1(kb, v2, o2);
      return s * V_pqrs_(v1, o1, v2, o2);
    }
    if(ca == 2) return sgl(ba, ka, ea, ba, bb);
    if(cb == 2) return sgl(bb, kb, eb, bb, ba);
    const auto oa = bits_to_indices(ba), ob = bits_to_indices(bb);
    double h = 0.;
    for(auto p : oa) h += T_pq_(p, p);
    for(auto p : ob) h += T_pq_(p, p);
    for(auto q : oa)
      for(auto p : oa) h += G2_red_data_[p + n * q];
    for(auto q : ob)
      for(auto p : ob) h += G2_red_data_[p + n * q];
    for(auto q : ob)
      for(auto p : oa) h += V2_red_data_[p + n * q];
    return h;
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
