// Determinant and integral-view types of the MACIS API (restating
// include/macis/types.hpp:22-89 of the reference) without the kokkos/mdspan
// dependency: col_major_span is a small non-owning column-major view.
#pragma once
#include <array>
#include <bitset>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace macis {

template <typename T, size_t rank>
class col_major_span {
 public:
  col_major_span() = default;
  template <typename... Ext>
  explicit col_major_span(T* p, Ext... ext) : p_(p), ext_{static_cast<size_t>(ext)...} {
    static_assert(sizeof...(Ext) == rank, "wrong number of extents");
  }
  T* data_handle() const { return p_; }
  size_t extent(size_t i) const { return ext_[i]; }
  size_t size() const {
    size_t s = 1;
    for(auto e : ext_) s *= e;
    return s;
  }
  template <typename... Idx>
  T& operator()(Idx... idx) const {
    static_assert(sizeof...(Idx) == rank, "wrong number of indices");
    const size_t ii[] = {static_cast<size_t>(idx)...};
    size_t off = 0, stride = 1;
    for(size_t d = 0; d < rank; ++d) {
      off += ii[d] * stride;
      stride *= ext_[d];
    }
    return p_[off];
  }

 private:
  T* p_ = nullptr;
  std::array<size_t, rank> ext_{};
};

template <typename T>
using matrix_span = col_major_span<T, 2>;
template <typename T>
using rank3_span = col_major_span<T, 3>;
template <typename T>
using rank4_span = col_major_span<T, 4>;

template <size_t N>
using wfn_t = std::bitset<N>;

template <size_t N>
using wavefunction_iterator_t = typename std::vector<std::bitset<N>>::iterator;

}  // namespace macis
