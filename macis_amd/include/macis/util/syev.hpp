// Symmetric eigensolver for the small (norb x norb) matrices of the host
// loops (asci_grow), in place of lapack::syev(Job::Vec, Uplo::Lower, ...) as
// asci_grow calls it (grow.hpp:90-92): cyclic Jacobi rotations on a full
// copy of the lower triangle, eigenvalues ascending in W, eigenvectors
// overwriting the columns of A. Eigenvector signs follow from the
// rotations (LAPACK fixes them differently); everything built on them here
// is invariant under a column's sign.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <numeric>
#include <vector>

namespace macis {

inline void syev_lower(size_t n, double* A, size_t lda, double* W) {
  std::vector<double> a(n * n), v(n * n, 0.);
  for(size_t j = 0; j < n; ++j)
    for(size_t i = j; i < n; ++i) a[i + n * j] = a[j + n * i] = A[i + lda * j];
  for(size_t i = 0; i < n; ++i) v[i * (n + 1)] = 1.;
  for(int sweep = 0; sweep < 100; ++sweep) {
    double off = 0., tot = 0.;
    for(size_t k = 0; k < n * n; ++k) {
      tot += a[k] * a[k];
      if(k % (n + 1)) off += a[k] * a[k];
    }
    if(off == 0. || off <= 1e-32 * tot) break;
    for(size_t p = 0; p + 1 < n; ++p)
      for(size_t q = p + 1; q < n; ++q) {
        const double apq = a[p + n * q];
        if(apq == 0.) continue;
        // past the first sweeps, drop what is below the diagonals' rounding
        const double g = 100. * std::abs(apq);
        if(sweep > 3 && std::abs(a[p * (n + 1)]) + g == std::abs(a[p * (n + 1)]) &&
           std::abs(a[q * (n + 1)]) + g == std::abs(a[q * (n + 1)])) {
          a[p + n * q] = a[q + n * p] = 0.;
          continue;
        }
        // rotation angle from cot(2 theta) = (a_qq - a_pp) / (2 a_pq)
        const double th = (a[q * (n + 1)] - a[p * (n + 1)]) / (2. * apq);
        const double t = (th >= 0. ? 1. : -1.) / (std::abs(th) + std::sqrt(1. + th * th));
        const double c = 1. / std::sqrt(1. + t * t), s = t * c;
        for(size_t k = 0; k < n; ++k) {
          const double x = a[k + n * p], y = a[k + n * q];
          a[k + n * p] = c * x - s * y;
          a[k + n * q] = s * x + c * y;
        }
        for(size_t k = 0; k < n; ++k) {
          const double x = a[p + n * k], y = a[q + n * k];
          a[p + n * k] = c * x - s * y;
          a[q + n * k] = s * x + c * y;
        }
        for(size_t k = 0; k < n; ++k) {
          const double x = v[k + n * p], y = v[k + n * q];
          v[k + n * p] = c * x - s * y;
          v[k + n * q] = s * x + c * y;
        }
      }
  }
  std::vector<size_t> ord(n);
  std::iota(ord.begin(), ord.end(), size_t(0));
  std::stable_sort(ord.begin(), ord.end(),
                   [&](size_t x, size_t y) { return a[x * (n + 1)] < a[y * (n + 1)]; });
  for(size_t j = 0; j < n; ++j) {
    W[j] = a[ord[j] * (n + 1)];
    for(size_t i = 0; i < n; ++i) A[i + lda * j] = v[i + n * ord[j]];
  }
}

}  // namespace macis
