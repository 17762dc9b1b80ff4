// sparsexx::csr_matrix (restating src/sparsexx/include/sparsexx/matrix_types/
// csr_matrix.hpp:27-224): host CSR with one index type for rowptr and
// colind, indexing base taken from rowptr[0].
#pragma once
#include <cstdint>
#include <memory>
#include <vector>

namespace sparsexx {

template <typename T, typename index_t = int64_t, typename Alloc = std::allocator<T>>
class csr_matrix {
 public:
  using value_type = T;
  using index_type = index_t;
  using size_type = int64_t;
  using allocator_type = Alloc;

  csr_matrix() = default;
  // Zero-filled storage with the given sizes (indexing default 1, :65-72).
  csr_matrix(size_type m, size_type n, size_type nnz, size_type indexing = 1)
      : m_(m), n_(n), nnz_(nnz), indexing_(indexing), nzval_(nnz), colind_(nnz), rowptr_(m + 1) {}
  // Adopts the arrays; indexing = rowptr[0] (:74-82).
  csr_matrix(size_type m, size_type n, std::vector<index_t>&& rowptr,
             std::vector<index_t>&& colind, std::vector<T>&& nzval)
      : m_(m),
        n_(n),
        nnz_(static_cast<size_type>(nzval.size())),
        indexing_(rowptr.empty() ? 0 : rowptr[0]),
        nzval_(std::move(nzval)),
        colind_(std::move(colind)),
        rowptr_(std::move(rowptr)) {}

  size_type m() const { return m_; }
  size_type n() const { return n_; }
  size_type nnz() const { return nnz_; }
  size_type indexing() const { return indexing_; }
  auto& nzval() { return nzval_; }
  auto& colind() { return colind_; }
  auto& rowptr() { return rowptr_; }
  const auto& nzval() const { return nzval_; }
  const auto& colind() const { return colind_; }
  const auto& rowptr() const { return rowptr_; }

  void set_indexing(index_type idx) {
    if(idx == indexing_) return;
    for(auto& i : colind_) i -= (indexing_ - idx);
    for(auto& i : rowptr_) i -= (indexing_ - idx);
    indexing_ = idx;
  }
  bool operator==(const csr_matrix& o) const noexcept {
    return m_ == o.m_ && n_ == o.n_ && indexing_ == o.indexing_ && colind_ == o.colind_ &&
           rowptr_ == o.rowptr_ && nzval_ == o.nzval_;
  }
  bool operator!=(const csr_matrix& o) const noexcept { return !(*this == o); }
  size_type mem_footprint() const noexcept {
    return nzval_.capacity() * sizeof(T) + colind_.capacity() * sizeof(index_t) +
           rowptr_.capacity() * sizeof(index_t);
  }

 protected:
  size_type m_ = 0, n_ = 0, nnz_ = 0, indexing_ = 0;
  std::vector<T> nzval_;
  std::vector<index_t> colind_;
  std::vector<index_t> rowptr_;
};

}  // namespace sparsexx
