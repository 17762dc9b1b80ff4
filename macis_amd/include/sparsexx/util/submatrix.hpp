// sparsexx::extract_diagonal_elements for host CSR (restating
// src/sparsexx/include/sparsexx/util/submatrix.hpp:264-295): first stored
// column equal to the row, 0 when absent.
#pragma once
#include <vector>

namespace sparsexx {

template <typename SpMatType>
std::vector<typename SpMatType::value_type> extract_diagonal_elements(const SpMatType& A) {
  std::vector<typename SpMatType::value_type> D(A.m(), 0);
  const auto base = A.indexing();
  for(int64_t i = 0; i < A.m(); ++i)
    for(auto j = A.rowptr()[i] - base; j < A.rowptr()[i + 1] - base; ++j)
      if(A.colind()[j] == i + base) {
        D[i] = A.nzval()[j];
        break;
      }
  return D;
}

}  // namespace sparsexx
