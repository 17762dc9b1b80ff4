// Orbital transforms of the reference's util/transform.hpp
// (src/macis/transform.cxx:22-115), restated as host loops: one contraction
// ("quarter") at a time, each moving the contracted index from the front of
// the tensor to a new index at the back. C is norb_old x norb_new
// (leading dimension LDC); Y may alias X, as in asci_grow's in-place
// rotation (grow.hpp:107-110).
#pragma once
#include <cstddef>
#include <vector>

namespace macis {
namespace detail {

// out(c, p) = sum_i in(i, c) C(i, p), in: no x ncols dense, out: ncols x nn
inline void transform_quarter(size_t no, size_t nn, size_t ncols, const double* in,
                              const double* C, size_t LDC, double* out) {
  for(size_t p = 0; p < nn; ++p)
    for(size_t c = 0; c < ncols; ++c) {
      double acc = 0.;
      const double* x = in + no * c;
      const double* u = C + LDC * p;
      for(size_t i = 0; i < no; ++i) acc += x[i] * u[i];
      out[c + ncols * p] = acc;
    }
}

}  // namespace detail

// Y(p,q) = C(i,p) * X(i,j) * C(j,q)
inline void two_index_transform(size_t norb_old, size_t norb_new, const double* X, size_t LDX,
                                const double* C, size_t LDC, double* Y, size_t LDY) {
  std::vector<double> x(norb_old * norb_old), t(norb_old * norb_new), y(norb_new * norb_new);
  for(size_t j = 0; j < norb_old; ++j)
    for(size_t i = 0; i < norb_old; ++i) x[i + norb_old * j] = X[i + LDX * j];
  detail::transform_quarter(norb_old, norb_new, norb_old, x.data(), C, LDC, t.data());  // t(j,p)
  detail::transform_quarter(norb_old, norb_new, norb_new, t.data(), C, LDC, y.data());  // y(p,q)
  for(size_t q = 0; q < norb_new; ++q)
    for(size_t p = 0; p < norb_new; ++p) Y[p + LDY * q] = y[p + norb_new * q];
}

// Y(p,q,r,s) = X(i,j,k,l) * C(i,p) * C(j,q) * C(k,r) * C(l,s)
inline void four_index_transform(size_t norb_old, size_t norb_new, const double* X, size_t LDX,
                                 const double* C, size_t LDC, double* Y, size_t LDY) {
  const size_t o = norb_old, n = norb_new;
  std::vector<double> a(o * o * o * o), b(o * o * o * n > 0 ? o * o * o * n : 1);
  for(size_t l = 0; l < o; ++l)
    for(size_t k = 0; k < o; ++k)
      for(size_t j = 0; j < o; ++j)
        for(size_t i = 0; i < o; ++i)
          a[i + o * (j + o * (k + o * l))] = X[i + LDX * (j + LDX * (k + LDX * l))];
  // (i,j,k,l) -> (j,k,l,p) -> (k,l,p,q) -> (l,p,q,r) -> (p,q,r,s)
  detail::transform_quarter(o, n, o * o * o, a.data(), C, LDC, b.data());
  detail::transform_quarter(o, n, o * o * n, b.data(), C, LDC, a.data());
  detail::transform_quarter(o, n, o * n * n, a.data(), C, LDC, b.data());
  detail::transform_quarter(o, n, n * n * n, b.data(), C, LDC, a.data());
  for(size_t s = 0; s < n; ++s)
    for(size_t r = 0; r < n; ++r)
      for(size_t q = 0; q < n; ++q)
        for(size_t p = 0; p < n; ++p)
          Y[p + LDY * (q + LDY * (r + LDY * s))] = a[p + n * (q + n * (r + n * s))];
}

}  // namespace macis
