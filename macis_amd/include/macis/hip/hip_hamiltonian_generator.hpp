// HipHamiltonianGenerator<N>: the MI355X drop-in for the reference's
// HamiltonianGenerator seam (hamiltonian_generator.hpp:66-72), replacing
// DoubleLoopHamiltonianGenerator::make_csr_hamiltonian_block_
// (double_loop.hpp:30-114). The CSR is built on the device by
// mhip_ham_build; the *_bit_ overrides copy it back into the reference's
// host csr_matrix, device_csr() keeps it in HBM.
#pragma once
#include <algorithm>
#include <limits>

#include <macis/hamiltonian_generator.hpp>
#include <macis/hip/runtime.hpp>

namespace macis {

template <size_t N = 64>
class HipHamiltonianGenerator : public HamiltonianGenerator<N> {
 public:
  using base_type = HamiltonianGenerator<N>;
  using typename base_type::full_det_iterator;
  using typename base_type::matrix_span_t;
  using typename base_type::rank4_span_t;
  template <typename index_t>
  using sparse_matrix_type = sparsexx::csr_matrix<double, index_t>;
  static constexpr int W = hip::words_per_det(N);

  HipHamiltonianGenerator(matrix_span_t T, rank4_span_t V, int device = 0)
      : base_type(T, V), ctx_(std::make_unique<hip::Context>(device)) {
    upload();
  }

  hip::Context& context() { return *ctx_; }

  // Device-resident CSR over [bra) x [ket).
  hip::DeviceCSR device_csr(full_det_iterator bra_begin, full_det_iterator bra_end,
                            full_det_iterator ket_begin, full_det_iterator ket_end,
                            double H_thresh) {
    auto db = hip::upload_dets<N>(*ctx_, bra_begin, bra_end);
    auto dk = hip::upload_dets<N>(*ctx_, ket_begin, ket_end);
    mhip_csr* h = nullptr;
    hip::check(mhip_ham_build(ctx_->get(), W, db.data(), std::distance(bra_begin, bra_end),
                              dk.data(), std::distance(ket_begin, ket_end), H_thresh, &h));
    return hip::DeviceCSR(h);
  }

  sparse_matrix_type<int32_t> make_csr_hamiltonian_block_32bit_(
      full_det_iterator bra_begin, full_det_iterator bra_end, full_det_iterator ket_begin,
      full_det_iterator ket_end, double H_thresh) override {
    auto H = device_csr(bra_begin, bra_end, ket_begin, ket_end, H_thresh);
    std::vector<int32_t> rp(H.m() + 1), ci(H.nnz());
    std::vector<double> nz(H.nnz());
    hip::check(mhip_csr_copy_to_host_i32(ctx_->get(), H.get(), rp.data(), ci.data(), nz.data()));
    return sparse_matrix_type<int32_t>(H.m(), H.n(), std::move(rp), std::move(ci), std::move(nz));
  }

  sparse_matrix_type<int64_t> make_csr_hamiltonian_block_64bit_(
      full_det_iterator bra_begin, full_det_iterator bra_end, full_det_iterator ket_begin,
      full_det_iterator ket_end, double H_thresh) override {
    auto H = device_csr(bra_begin, bra_end, ket_begin, ket_end, H_thresh);
    std::vector<int64_t> rp(H.m() + 1), ci(H.nnz());
    std::vector<double> nz(H.nnz());
    hip::check(mhip_csr_copy_to_host_i64(ctx_->get(), H.get(), rp.data(), ci.data(), nz.data()));
    return sparse_matrix_type<int64_t>(H.m(), H.n(), std::move(rp), std::move(ci), std::move(nz));
  }

  // form_rdms (double_loop.hpp:116-165) on the device through mhip_form_rdms:
  // the spans are added into, as the reference does, and a null 2-RDM handle
  // forms the 1-RDM alone. C weights bra i with C[i] and ket j with C[j].
  // Spans wider than norb (any extents >= norb, as the reference indexes
  // them only below norb) are staged through dense norb-sized copies.
  void form_rdms(full_det_iterator bra_begin, full_det_iterator bra_end,
                 full_det_iterator ket_begin, full_det_iterator ket_end, double* C,
                 matrix_span_t ordm, rank4_span_t trdm) override {
    const size_t n = this->norb_;
    const size_t nb = std::distance(bra_begin, bra_end), nk = std::distance(ket_begin, ket_end);
    if(!nb || !nk) return;
    if(ordm.extent(0) < n || ordm.extent(1) < n)
      throw std::runtime_error("form_rdms: the 1-RDM span is smaller than norb x norb");
    const bool two = trdm.data_handle() != nullptr;
    bool dense = ordm.extent(0) == n && ordm.extent(1) == n;
    if(two)
      for(size_t d = 0; d < 4; ++d) {
        if(trdm.extent(d) < n)
          throw std::runtime_error("form_rdms: the 2-RDM span is smaller than norb^4");
        dense = dense && trdm.extent(d) == n;
      }
    std::vector<double> h1, h2;
    double *p1 = ordm.data_handle(), *p2 = trdm.data_handle();
    if(!dense) {  // gather the norb-sized corners, added into below
      h1.resize(n * n);
      for(size_t q = 0; q < n; ++q)
        for(size_t p = 0; p < n; ++p) h1[p + n * q] = ordm(p, q);
      p1 = h1.data();
      if(two) {
        h2.resize(n * n * n * n);
        for(size_t l = 0; l < n; ++l)
          for(size_t k = 0; k < n; ++k)
            for(size_t j = 0; j < n; ++j)
              for(size_t i = 0; i < n; ++i) h2[i + n * (j + n * (k + n * l))] = trdm(i, j, k, l);
        p2 = h2.data();
      }
    }
    auto db = hip::upload_dets<N>(*ctx_, bra_begin, bra_end);
    auto dk = hip::upload_dets<N>(*ctx_, ket_begin, ket_end);
    hip::DeviceBuffer<double> dC(*ctx_, C, std::max(nb, nk));
    hip::DeviceBuffer<double> d1(*ctx_, p1, n * n);
    hip::DeviceBuffer<double> d2;
    if(two) d2 = hip::DeviceBuffer<double>(*ctx_, p2, n * n * n * n);
    hip::check(mhip_form_rdms(ctx_->get(), W, db.data(), nb, dk.data(), nk, dC.data(), d1.data(),
                              two ? d2.data() : nullptr));
    d1.download(p1, n * n);
    if(two) d2.download(p2, n * n * n * n);
    if(!dense) {
      for(size_t q = 0; q < n; ++q)
        for(size_t p = 0; p < n; ++p) ordm(p, q) = h1[p + n * q];
      if(two)
        for(size_t l = 0; l < n; ++l)
          for(size_t k = 0; k < n; ++k)
            for(size_t j = 0; j < n; ++j)
              for(size_t i = 0; i < n; ++i) trdm(i, j, k, l) = h2[i + n * (j + n * (k + n * l))];
    }
  }

 protected:
  void integrals_changed() override {
    if(ctx_) upload();
  }

 private:
  void upload() {
    hip::check(mhip_set_integrals(ctx_->get(), this->norb_, this->T_pq_.data_handle(),
                                  this->V_pqrs_.data_handle()));
  }
  std::unique_ptr<hip::Context> ctx_;
};

}  // namespace macis
