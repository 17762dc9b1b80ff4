// reorder_ci_on_coeff (restating include/macis/asci/determinant_sort.hpp
// :18-37): sort determinants by |C| descending (ties keep their order).
#pragma once
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace macis {
template <typename WfnT>
void reorder_ci_on_coeff(std::vector<WfnT>& dets, std::vector<double>& C) {
  std::vector<size_t> idx(C.size());
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(),
                   [&](size_t i, size_t j) { return std::abs(C[i]) > std::abs(C[j]); });
  std::vector<double> rc(C.size());
  std::vector<WfnT> rd(dets.size());
  for(size_t i = 0; i < idx.size(); ++i) {
    rc[i] = C[idx[i]];
    rd[i] = dets[idx[i]];
  }
  C = std::move(rc);
  dets = std::move(rd);
}
}  // namespace macis
