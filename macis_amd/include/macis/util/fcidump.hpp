// FCIDUMP reading and writing (restating src/macis/fcidump.cxx:78-207):
// 5-token lines, either "p q r s value" or "value p q r s", 1-based indices,
// all-zero indices = core energy, (p,q,0,0) one-body, four nonzero two-body
// with 8-fold unpacking. Arrays are column-major with leading dimension
// norb, or col_major_span views of any extents.
#pragma once
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include <macis/types.hpp>

namespace macis {
namespace detail {
inline bool fcidump_is_float(const std::string& s) {
  return std::any_of(s.begin(), s.end(), [](char c) { return std::isalpha((unsigned char)c) || c == '.'; });
}
template <typename F>
void fcidump_for_each(const std::string& fname, F&& f) {
  std::ifstream file(fname);
  if(!file) throw std::runtime_error("cannot open " + fname);
  std::string line;
  while(std::getline(file, line)) {
    std::istringstream ss(line);
    std::vector<std::string> tok;
    std::string t;
    while(ss >> t) tok.push_back(t);
    if(tok.size() != 5) continue;
    const bool idx_first = fcidump_is_float(tok.back()), val_first = fcidump_is_float(tok.front());
    if(idx_first && val_first) throw std::runtime_error("Invalid FCIDUMP Line");
    int p, q, r, s;
    double v;
    if(idx_first) {
      p = std::stoi(tok[0]); q = std::stoi(tok[1]); r = std::stoi(tok[2]); s = std::stoi(tok[3]);
      v = std::stod(tok[4]);
    } else {
      p = std::stoi(tok[1]); q = std::stoi(tok[2]); r = std::stoi(tok[3]); s = std::stoi(tok[4]);
      v = std::stod(tok[0]);
    }
    if(p < 0 || q < 0 || r < 0 || s < 0) throw std::runtime_error("Invalid Orb Idx");
    f(p, q, r, s, v);
  }
}
}  // namespace detail

inline uint32_t read_fcidump_norb(std::string fname) {
  int m = 0;
  detail::fcidump_for_each(fname, [&](int p, int q, int r, int s, double) {
    m = std::max({m, p, q, r, s});
  });
  return m;
}

inline double read_fcidump_core(std::string fname) {
  double e = 0.;
  bool found = false;
  detail::fcidump_for_each(fname, [&](int p, int q, int r, int s, double v) {
    if(!found && !(p || q || r || s)) {
      e = v;
      found = true;
    }
  });
  return e;
}

inline void read_fcidump_1body(std::string fname, double* T, size_t LDT) {
  detail::fcidump_for_each(fname, [&](int p, int q, int r, int s, double v) {
    if((p || q || r || s) && !(p && q && r && s)) {
      T[(p - 1) + LDT * (q - 1)] = v;
      T[(q - 1) + LDT * (p - 1)] = v;
    }
  });
}

inline void read_fcidump_2body(std::string fname, double* V, size_t LDV) {
  const size_t L = LDV, L2 = L * L, L3 = L2 * L;
  detail::fcidump_for_each(fname, [&](int p, int q, int r, int s, double v) {
    if(p && q && r && s) {
      --p, --q, --r, --s;
      for(auto [a, b, c, d] : {std::tuple{p, q, r, s}, {p, q, s, r}, {q, p, r, s}, {q, p, s, r},
                               {r, s, p, q}, {s, r, p, q}, {r, s, q, p}, {s, r, q, p}})
        V[a + L * b + L2 * c + L3 * d] = v;
    }
  });
}

// Span variants (fcidump.hpp:54-71): the entries land at the view's own
// column-major positions.
inline void read_fcidump_1body(std::string fname, col_major_span<double, 2> T) {
  detail::fcidump_for_each(fname, [&](int p, int q, int r, int s, double v) {
    if((p || q || r || s) && !(p && q && r && s)) {
      T(p - 1, q - 1) = v;
      T(q - 1, p - 1) = v;
    }
  });
}

inline void read_fcidump_2body(std::string fname, col_major_span<double, 4> V) {
  detail::fcidump_for_each(fname, [&](int p, int q, int r, int s, double v) {
    if(p && q && r && s) {
      --p, --q, --r, --s;
      for(auto [a, b, c, d] : {std::tuple{p, q, r, s}, {p, q, s, r}, {q, p, r, s}, {q, p, s, r},
                               {r, s, p, q}, {s, r, p, q}, {r, s, q, p}, {s, r, q, p}})
        V(a, b, c, d) = v;
    }
  });
}

// write_fcidump (fcidump.cxx:184-207): every two-body entry (i,j,k,l) in
// the reference's loop order (i outermost, l innermost), then every one-body
// (i,j,0,0), then the core line (0,0,0,0), each as four width-8 integers and
// a width-25, 16-digit scientific value. Lines are appended to the file, as
// the reference's logger sink does.
inline void write_fcidump(std::string fname, size_t norb, const double* T, size_t LDT,
                          const double* V, size_t LDV, double E_core) {
  std::FILE* f = std::fopen(fname.c_str(), "a");
  if(!f) throw std::runtime_error("cannot open " + fname);
  auto line = [&](size_t p, size_t q, size_t r, size_t s, double v) {
    std::fprintf(f, "%8zu %8zu %8zu %8zu %25.16e\n", p, q, r, s, v);
  };
  for(size_t i = 0; i < norb; ++i)
    for(size_t j = 0; j < norb; ++j)
      for(size_t k = 0; k < norb; ++k)
        for(size_t l = 0; l < norb; ++l)
          line(i + 1, j + 1, k + 1, l + 1, V[i + j * LDV + k * LDV * LDV + l * LDV * LDV * LDV]);
  for(size_t i = 0; i < norb; ++i)
    for(size_t j = 0; j < norb; ++j) line(i + 1, j + 1, 0, 0, T[i + j * LDT]);
  line(0, 0, 0, 0, E_core);
  if(std::fclose(f) != 0) throw std::runtime_error("cannot write " + fname);
}

}  // namespace macis
