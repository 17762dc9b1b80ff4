// Thin C++ layer over the C-ABI (include/macis_hip.h): RAII handles, status
// -> std::runtime_error with the reference's messages, and device uploads
// of determinant lists.
#pragma once
#include <macis_hip.h>

#include <array>
#include <bitset>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include <macis/bitset_operations.hpp>

namespace macis::hip {

inline void check(int rc) {
  if(rc == MHIP_SUCCESS) return;
  // The reference signals failures with std::runtime_error only.
  if(rc == MHIP_ERR_NOT_CONVERGED) throw std::runtime_error("Davidson Did Not Converge!");
  if(rc == MHIP_ERR_NO_GUESS) throw std::runtime_error("Davidson: No Guess Provided");
  if(rc == MHIP_ERR_UNSUPPORTED) throw std::runtime_error("Unsupported index_t");
  std::string msg = mhip_last_error();
  if(msg.empty()) msg = mhip_error_string(rc);
  throw std::runtime_error("macis_hip: " + msg);
}

class Context {
 public:
  explicit Context(int device = 0) { check(mhip_context_create(device, &h_)); }
  ~Context() {
    if(h_) mhip_context_destroy(h_);
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  mhip_context* get() const { return h_; }
  mhip_timings timings() const {
    mhip_timings t;
    check(mhip_get_timings(h_, &t));
    return t;
  }

 private:
  mhip_context* h_ = nullptr;
};

// Process-wide context on device 0 for API calls that carry no generator.
inline Context& default_context() {
  static Context ctx(0);
  return ctx;
}

template <typename T>
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(Context& c, size_t n) : ctx_(c.get()), n_(n) {
    void* p = nullptr;
    check(mhip_malloc(ctx_, n * sizeof(T), &p));
    p_ = static_cast<T*>(p);
  }
  DeviceBuffer(Context& c, const T* host, size_t n) : DeviceBuffer(c, n) { upload(host, n); }
  ~DeviceBuffer() {
    if(p_) mhip_free(ctx_, p_);
  }
  DeviceBuffer(DeviceBuffer&& o) noexcept : ctx_(o.ctx_), p_(o.p_), n_(o.n_) { o.p_ = nullptr; }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    std::swap(ctx_, o.ctx_);
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
    return *this;
  }
  DeviceBuffer(const DeviceBuffer&) = delete;
  void upload(const T* host, size_t n) { check(mhip_memcpy_h2d(ctx_, p_, host, n * sizeof(T))); }
  void download(T* host, size_t n) const { check(mhip_memcpy_d2h(ctx_, host, p_, n * sizeof(T))); }
  void zero() { check(mhip_memset(ctx_, p_, 0, n_ * sizeof(T))); }
  T* data() const { return p_; }
  size_t size() const { return n_; }

 private:
  mhip_context* ctx_ = nullptr;
  T* p_ = nullptr;
  size_t n_ = 0;
};

// Device-resident CSR handle.
class DeviceCSR {
 public:
  DeviceCSR() = default;
  explicit DeviceCSR(mhip_csr* h) : h_(h) {}
  ~DeviceCSR() {
    if(h_) mhip_csr_destroy(h_);
  }
  DeviceCSR(DeviceCSR&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
  DeviceCSR& operator=(DeviceCSR&& o) noexcept {
    std::swap(h_, o.h_);
    return *this;
  }
  DeviceCSR(const DeviceCSR&) = delete;
  mhip_csr* get() const { return h_; }
  int64_t m() const { return info()[0]; }
  int64_t n() const { return info()[1]; }
  int64_t nnz() const { return info()[2]; }

 private:
  std::array<int64_t, 3> info() const {
    std::array<int64_t, 3> r{};
    check(mhip_csr_info(h_, &r[0], &r[1], &r[2]));
    return r;
  }
  mhip_csr* h_ = nullptr;
};

constexpr int words_per_det(size_t N) { return N == 64 ? 1 : 2; }

// Determinant range -> packed uint64 words -> device.
template <size_t N, typename It>
DeviceBuffer<uint64_t> upload_dets(Context& ctx, It begin, It end) {
  static_assert(N == 64 || N == 128, "wfn_t<64> or wfn_t<128> only");
  constexpr int W = words_per_det(N);
  const size_t n = std::distance(begin, end);
  std::vector<uint64_t> w(std::max<size_t>(n, 1) * W);
  size_t i = 0;
  for(It it = begin; it != end; ++it, ++i) bitset_words(*it, w.data() + W * i);
  DeviceBuffer<uint64_t> d(ctx, std::max<size_t>(n, 1) * W);
  if(n) d.upload(w.data(), n * W);
  return d;
}

template <size_t N>
std::vector<std::bitset<N>> download_dets(const DeviceBuffer<uint64_t>& d, size_t n) {
  constexpr int W = words_per_det(N);
  std::vector<uint64_t> w(std::max<size_t>(n, 1) * W);
  if(n) d.download(w.data(), n * W);
  std::vector<std::bitset<N>> out(n);
  for(size_t i = 0; i < n; ++i) out[i] = bitset_from_words<N>(w.data() + W * i);
  return out;
}

// Options of the device Davidson beyond the reference signature.
struct DavidsonOptions {
  // 0 = the reference algorithm (no restart). > 0 selects the labelled
  // restarted variant.
  int64_t max_restarts = 0;
};
inline DavidsonOptions& davidson_options() {
  static thread_local DavidsonOptions o;
  return o;
}

}  // namespace macis::hip
