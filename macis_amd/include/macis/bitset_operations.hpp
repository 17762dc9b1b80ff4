// Bit primitives of the MACIS API used on the hot path (restating
// include/macis/bitset_operations.hpp:108-307). std::bitset<64> and
// std::bitset<128> are read through their word representation, exactly as
// the reference does (libstdc++ stores word 0 first).
#pragma once
#include <bitset>
#include <cstdint>
#include <cstring>
#include <vector>

#include <macis/types.hpp>

namespace macis {

// Words of a 64- or 128-bit determinant (alpha word first).
template <size_t N>
inline void bitset_words(const std::bitset<N>& b, uint64_t* w) {
  static_assert(N == 64 || N == 128, "wfn_t<64> or wfn_t<128> only");
  std::memcpy(w, &b, N / 8);
}
template <size_t N>
inline std::bitset<N> bitset_from_words(const uint64_t* w) {
  static_assert(N == 64 || N == 128, "wfn_t<64> or wfn_t<128> only");
  std::bitset<N> b;
  std::memcpy(static_cast<void*>(&b), w, N / 8);
  return b;
}

// Low (alpha) and high (beta) halves (bitset_operations.hpp:353-375).
template <size_t N>
inline std::bitset<N / 2> bitset_lo_word(std::bitset<N> bits) {
  uint64_t w[2] = {0, 0};
  bitset_words(bits, w);
  if constexpr(N == 128) return std::bitset<64>(w[0]);
  else return std::bitset<32>(w[0] & 0xffffffffull);
}
template <size_t N>
inline std::bitset<N / 2> bitset_hi_word(std::bitset<N> bits) {
  uint64_t w[2] = {0, 0};
  bitset_words(bits, w);
  if constexpr(N == 128) return std::bitset<64>(w[1]);
  else return std::bitset<32>(w[0] >> 32);
}

// Numeric order of the whole word (beta most significant), :377-395.
template <size_t N>
inline bool bitset_less(std::bitset<N> x, std::bitset<N> y) {
  uint64_t a[2] = {0, 0}, b[2] = {0, 0};
  bitset_words(x, a);
  bitset_words(y, b);
  if constexpr(N == 128) {
    if(a[1] != b[1]) return a[1] < b[1];
  }
  return a[0] < b[0];
}

template <size_t N>
struct bitset_less_comparator {
  bool operator()(std::bitset<N> x, std::bitset<N> y) const { return bitset_less(x, y); }
};

// full_mask<N>(i): i set low bits (:225-230).
template <size_t N>
inline std::bitset<N> full_mask(size_t i) {
  std::bitset<N> m;
  for(size_t k = 0; k < i && k < N; ++k) m.set(k);
  return m;
}

// 1-based first set bit, 0 if none (:232-261).
template <size_t N>
inline uint32_t ffs(std::bitset<N> bits) {
  for(size_t i = 0; i < N; ++i)
    if(bits[i]) return static_cast<uint32_t>(i + 1);
  return 0;
}

// 0-based last set bit (:263-292).
template <size_t N>
inline uint32_t fls(std::bitset<N> bits) {
  for(size_t i = N; i-- > 0;)
    if(bits[i]) return static_cast<uint32_t>(i);
  return static_cast<uint32_t>(-1);
}

// Ascending indices of set bits (:294-314).
template <size_t N>
inline void bits_to_indices(std::bitset<N> bits, std::vector<uint32_t>& idx) {
  idx.clear();
  for(size_t i = 0; i < N; ++i)
    if(bits[i]) idx.push_back(static_cast<uint32_t>(i));
}
template <size_t N>
inline std::vector<uint32_t> bits_to_indices(std::bitset<N> bits) {
  std::vector<uint32_t> v;
  bits_to_indices(bits, v);
  return v;
}

}  // namespace macis
