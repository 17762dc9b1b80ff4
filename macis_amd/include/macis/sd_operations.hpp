// Determinant-space helpers of the MACIS API on the hot path (restating
// include/macis/sd_operations.hpp:31-97, 421-476).
#pragma once
#include <algorithm>
#include <bitset>
#include <string>
#include <tuple>
#include <vector>

#include <macis/bitset_operations.hpp>

namespace macis {

// Lowest nalpha alpha and nbeta beta orbitals occupied (:31-37).
template <size_t N>
std::bitset<N> canonical_hf_determinant(uint32_t nalpha, uint32_t nbeta) {
  static_assert(N % 2 == 0, "N Must Be Even");
  std::bitset<N> s;
  for(uint32_t i = 0; i < nalpha; ++i) s.set(i);
  for(uint32_t i = 0; i < nbeta; ++i) s.set(i + N / 2);
  return s;
}

// Occupied / first (norb - nocc) unoccupied orbitals (:82-97).
template <size_t N>
void bitset_to_occ_vir(size_t norb, std::bitset<N> state, std::vector<uint32_t>& occ,
                       std::vector<uint32_t>& vir) {
  bits_to_indices(state, occ);
  vir.clear();
  const size_t nvir = norb - occ.size();
  for(size_t i = 0; i < N && vir.size() < nvir; ++i)
    if(!state[i]) vir.push_back(static_cast<uint32_t>(i));
}

// All nset-subsets of nbits bits in the std::prev_permutation order of the
// reference (:440-458), i.e. lexicographic order of the occupied indices.
template <size_t N>
std::vector<std::bitset<N>> generate_combs(uint64_t nbits, uint64_t nset) {
  std::vector<std::bitset<N>> out;
  std::vector<uint32_t> idx(nset);
  for(uint32_t i = 0; i < nset; ++i) idx[i] = i;
  if(nset > nbits) return out;
  for(;;) {
    std::bitset<N> b;
    for(auto i : idx) b.set(i);
    out.push_back(b);
    int64_t k = static_cast<int64_t>(nset) - 1;
    while(k >= 0 && idx[k] == nbits - nset + k) --k;
    if(k < 0) break;
    ++idx[k];
    for(uint64_t j = k + 1; j < nset; ++j) idx[j] = idx[j - 1] + 1;
  }
  return out;
}

// FCI space, alpha strings outer, beta inner (:460-476).
template <size_t N>
std::vector<std::bitset<N>> generate_hilbert_space(size_t norbs, size_t nalpha, size_t nbeta) {
  auto a = generate_combs<N>(norbs, nalpha);
  auto b = generate_combs<N>(norbs, nbeta);
  std::vector<std::bitset<N>> out;
  out.reserve(a.size() * b.size());
  for(auto& x : a)
    for(auto& y : b) out.push_back(x | (y << (N / 2)));
  return out;
}

// CISD space: HF, alpha singles, beta singles, alpha doubles, beta doubles,
// alpha x beta singles (:345-438).
template <size_t N>
std::vector<std::bitset<N>> generate_cisd_hilbert_space(size_t norb, std::bitset<N> state) {
  auto spin = [&](std::bitset<N / 2> s, std::vector<std::bitset<N / 2>>& sing,
                  std::vector<std::bitset<N / 2>>& dbl) {
    std::vector<uint32_t> occ, vir;
    bitset_to_occ_vir(norb, s, occ, vir);
    for(size_t a = 0; a < vir.size(); ++a)
      for(size_t i = 0; i < occ.size(); ++i) {
        auto t = s;
        t.flip(occ[i]).flip(vir[a]);
        sing.push_back(t);
      }
    for(size_t a = 0; a < vir.size(); ++a)
      for(size_t i = 0; i < occ.size(); ++i)
        for(size_t b = a + 1; b < vir.size(); ++b)
          for(size_t j = i + 1; j < occ.size(); ++j) {
            auto t = s;
            t.flip(occ[i]).flip(vir[a]).flip(occ[j]).flip(vir[b]);
            dbl.push_back(t);
          }
  };
  auto lo = bitset_lo_word(state), hi = bitset_hi_word(state);
  std::vector<std::bitset<N / 2>> sa, da, sb, db;
  spin(lo, sa, da);
  spin(hi, sb, db);
  auto full = [](std::bitset<N / 2> a, std::bitset<N / 2> b) {
    std::bitset<N> r;
    for(size_t i = 0; i < N / 2; ++i) {
      if(a[i]) r.set(i);
      if(b[i]) r.set(i + N / 2);
    }
    return r;
  };
  std::vector<std::bitset<N>> out{state};
  for(auto& x : sa) out.push_back(full(x, hi));
  for(auto& x : sb) out.push_back(full(lo, x));
  for(auto& x : da) out.push_back(full(x, hi));
  for(auto& x : db) out.push_back(full(lo, x));
  for(auto& x : sa)
    for(auto& y : sb) out.push_back(full(x, y));
  return out;
}

// Excitation primitives (sd_operations.hpp:497-551). For bra and ket that
// differ by one or two electron moves within one spin string (ex = bra ^
// ket): the hole o is the lowest flipped orbital occupied in the ket, the
// particle v the lowest flipped orbital occupied in the bra, and the sign is
// the parity of the ket's occupied orbitals strictly between them; a double
// is two such moves, the second applied to the updated ket.
template <size_t N>
uint32_t first_occupied_flipped(std::bitset<N> state, std::bitset<N> ex) {
  return ffs(state & ex) - 1u;
}

template <size_t N>
double single_excitation_sign(std::bitset<N> state, unsigned p, unsigned q) {
  const unsigned lo = std::min(p, q), hi = std::max(p, q);
  const auto between = full_mask<N>(hi) ^ full_mask<N>(lo + 1);
  return ((state & between).count() % 2) ? -1. : 1.;
}

template <size_t N>
inline auto single_excitation_sign_indices(std::bitset<N> bra, std::bitset<N> ket,
                                           std::bitset<N> ex) {
  const uint32_t o1 = first_occupied_flipped(ket, ex);
  const uint32_t v1 = first_occupied_flipped(bra, ex);
  return std::make_tuple(o1, v1, single_excitation_sign(ket, v1, o1));
}

template <size_t N>
inline auto doubles_sign_indices(std::bitset<N> bra, std::bitset<N> ket, std::bitset<N> ex) {
  const uint32_t o1 = first_occupied_flipped(ket, ex);
  const uint32_t v1 = first_occupied_flipped(bra, ex);
  double sign = single_excitation_sign(ket, v1, o1);
  ket.flip(o1).flip(v1);
  ex.flip(o1).flip(v1);
  const uint32_t o2 = first_occupied_flipped(ket, ex);
  const uint32_t v2 = first_occupied_flipped(bra, ex);
  sign *= single_excitation_sign(ket, v2, o2);
  return std::make_tuple(o1, v1, o2, v2, sign);
}

template <size_t N>
inline double doubles_sign(std::bitset<N> bra, std::bitset<N> ket, std::bitset<N> ex) {
  return std::get<4>(doubles_sign_indices(bra, ket, ex));
}

// to_canonical_string / from_canonical_string (sd_operations.hpp:596-630):
// one character per spatial orbital of the spin-string width, '2' doubly
// occupied, 'u' alpha only, 'd' beta only, '0' empty; a shorter string
// leaves the remaining orbitals empty.
template <size_t N>
std::string to_canonical_string(std::bitset<N> state) {
  static_assert((N % 2) == 0, "N Odd");
  std::string str;
  for(size_t i = 0; i < N / 2; ++i) {
    const bool a = state[i], b = state[i + N / 2];
    str.push_back(a && b ? '2' : a ? 'u' : b ? 'd' : '0');
  }
  return str;
}

template <size_t N>
std::bitset<N> from_canonical_string(std::string str) {
  // both halves are built N bits wide and combined as alpha | beta << N/2,
  // as the reference does (a character past N/2 lands in the beta half or,
  // past N, throws std::out_of_range from bitset::set)
  std::bitset<N> alpha, beta;
  for(size_t i = 0; i < str.length(); ++i) {
    if(str[i] == '2' || str[i] == 'u') alpha.set(i);
    if(str[i] == '2' || str[i] == 'd') beta.set(i);
  }
  return alpha | (beta << (N / 2));
}

}  // namespace macis
