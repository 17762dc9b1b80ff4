// Context, device memory helpers, integral upload and error reporting of
// the C-ABI (include/macis_hip.h).
#include <algorithm>
#include <cstring>
#include <map>
#include <cstdlib>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "mhip_internal.hpp"

// Device scratch of the allocation-free entry points (mhip_gemm_tn_noalloc):
// one array per GPU, held until the process ends. Host code never uses this
// name as an address: the per-device address comes from hipGetSymbolAddress
// at context creation and is kept in g_scratch_addr.
static __device__ __align__(256) unsigned char scratch[mhip::kNoAllocScratchBytes];

namespace mhip {

namespace {
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {nullptr};

// Looks up this device's copy of `scratch` (the device must be current).
void init_noalloc_scratch(int device) {
  MHIP_REQUIRE(device >= 0 && device < kMaxDevices, MHIP_ERR_INVALID_ARG,
               "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if(g_scratch_addr[device]) return;
  void* p = nullptr;
  MHIP_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  MHIP_REQUIRE(p, MHIP_ERR_HIP, "hipGetSymbolAddress returned no address");
  g_scratch_addr[device] = p;
}
}  // namespace

namespace {
// Every block is kept for reuse: a hipMalloc + hipFree pair costs ~70 us,
// which the ~10 small buffers of a Hamiltonian build had added up to
// ~0.7 ms of a 6 ms build. (A reused block may be up to 1/8 larger than
// asked; device_cache_trim returns them all when an allocation fails.)
constexpr size_t kCacheMinBytes = 0;
struct BlockCache {
  std::mutex mu;
  std::multimap<size_t, void*> idle[kMaxDevices];  // size -> block, per device
  struct Live {
    size_t bytes;
    int device;
  };
  std::unordered_map<void*, Live> live;  // cacheable blocks in use
};
// Never destroyed: blocks still cached at process exit go back with the
// process, not through hipFree calls from static destructors.
BlockCache& block_cache() {
  static BlockCache* c = new BlockCache;
  return *c;
}
int current_device() {
  int d = 0;
  (void)hipGetDevice(&d);
  return d;
}
}  // namespace

namespace {
// MHIP_ALLOC_POISON=1 (a debugging aid): every block handed out is filled
// with 0xff bytes first, so a read of memory no kernel has written shows up
// as NaNs / -1 indices on every run instead of depending on what the block
// held before.
void poison(void* p, size_t bytes) {
  const char* e = std::getenv("MHIP_ALLOC_POISON");
  if(e && e[0] == '1' && p && bytes) {
    (void)hipMemset(p, 0xff, bytes);
    (void)hipDeviceSynchronize();
  }
}
}  // namespace

void* device_alloc(size_t bytes) {
  void* p = device_alloc_raw(bytes);
  poison(p, bytes);
  return p;
}

void* device_alloc_raw(size_t bytes) {
  if(!bytes) return nullptr;
  void* p = nullptr;
  if(bytes < kCacheMinBytes) {
    hipError_t e = hipMalloc(&p, bytes);
    if(e == hipErrorOutOfMemory) {  // kept blocks may hold the memory
      (void)hipGetLastError();
      device_cache_trim(current_device());
      e = hipMalloc(&p, bytes);
    }
    MHIP_HIP_CHECK(e);
    return p;
  }
  BlockCache& c = block_cache();
  const int dev = current_device();
  MHIP_REQUIRE(dev >= 0 && dev < kMaxDevices, MHIP_ERR_INVALID_ARG, "device index out of range");
  {
    std::lock_guard<std::mutex> lock(c.mu);
    auto& idle = c.idle[dev];
    auto it = idle.lower_bound(bytes);
    if(it != idle.end() && it->first <= bytes + bytes / 8) {
      p = it->second;
      c.live[p] = {it->first, dev};
      idle.erase(it);
      return p;
    }
  }
  hipError_t e = hipMalloc(&p, bytes);
  if(e == hipErrorOutOfMemory) {
    (void)hipGetLastError();  // not a pending error: retried after the trim
    device_cache_trim(dev);
    e = hipMalloc(&p, bytes);
  }
  MHIP_HIP_CHECK(e);
  std::lock_guard<std::mutex> lock(c.mu);
  c.live[p] = {bytes, dev};
  return p;
}

void device_free(void* p) noexcept {
  if(!p) return;
  BlockCache& c = block_cache();
  {
    std::lock_guard<std::mutex> lock(c.mu);
    auto it = c.live.find(p);
    if(it != c.live.end()) {
      const BlockCache::Live l = it->second;
      c.live.erase(it);
      const int cur = current_device();
      if(cur != l.device) (void)hipSetDevice(l.device);
      (void)hipDeviceSynchronize();  // no kernel may still use the block
      if(cur != l.device) (void)hipSetDevice(cur);
      c.idle[l.device].emplace(l.bytes, p);
      return;
    }
  }
  (void)hipFree(p);
}

void device_cache_trim(int device) noexcept {
  if(device < 0 || device >= kMaxDevices) return;
  std::vector<void*> blocks;
  {
    BlockCache& c = block_cache();
    std::lock_guard<std::mutex> lock(c.mu);
    for(auto& kv : c.idle[device]) blocks.push_back(kv.second);
    c.idle[device].clear();
  }
  if(blocks.empty()) return;
  const int cur = current_device();
  if(cur != device) (void)hipSetDevice(device);
  for(void* b : blocks) (void)hipFree(b);
  if(cur != device) (void)hipSetDevice(cur);
}

void device_mem_info(size_t* free_b, size_t* total_b) {
  MHIP_HIP_CHECK(hipMemGetInfo(free_b, total_b));
  const int dev = current_device();
  if(dev < 0 || dev >= kMaxDevices) return;
  BlockCache& c = block_cache();
  std::lock_guard<std::mutex> lock(c.mu);
  for(auto& kv : c.idle[dev]) *free_b += kv.first;
}

void* noalloc_scratch(int device) {
  if(device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_addr[device];
}

namespace {
// Trivially destructible on purpose: contexts may be destroyed from static
// destructors at process exit, after the thread's thread_local objects with
// destructors are gone, and every C-ABI call records its status here.
constexpr size_t kMaxErr = 1024;
thread_local char g_last_error[kMaxErr] = {0};
}  // namespace

void set_last_error(const std::string& msg) {
  const size_t n = std::min(msg.size(), kMaxErr - 1);
  std::memcpy(g_last_error, msg.data(), n);
  g_last_error[n] = '\0';
}

void use_device(mhip_context* ctx) {
  MHIP_REQUIRE(ctx, MHIP_ERR_INVALID_ARG, "null context");
  MHIP_HIP_CHECK(hipSetDevice(ctx->device));
}

namespace {
// Copies bytes (multiple of 4 plus a tail) into host-pinned memory with a
// kernel: small reads then need neither a copy engine nor a blit launch
// from the runtime's memcpy path.
__global__ void stage_copy_kernel(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src,
                                  size_t words) {
  for(size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
      i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace

void d2h(mhip_context* ctx, void* dst, const void* src, size_t bytes) {
  if(!bytes) return;
  const bool aligned = ((uintptr_t)src % 4) == 0 && bytes % 4 == 0;
  if(bytes <= (size_t(4) << 20) && aligned) {
    ctx->pinned.reserve(std::max<size_t>(bytes, 64 << 10));
    const size_t words = bytes / 4;
    const unsigned blocks = (unsigned)std::min<size_t>(1024, (words + 255) / 256);
    hipLaunchKernelGGL(stage_copy_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                       static_cast<uint32_t*>(ctx->pinned.p), static_cast<const uint32_t*>(src),
                       words);
    MHIP_HIP_CHECK(hipGetLastError());
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    std::memcpy(dst, ctx->pinned.p, bytes);
  } else {
    MHIP_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  }
}

DevBuf take_buffer(mhip_context* ctx, size_t bytes) {
  bytes = std::max<size_t>(bytes, 1);
  int best = -1;
  for(int i = 0; i < (int)ctx->spare.size(); ++i) {
    const size_t c = ctx->spare[i].bytes;
    if(c >= bytes && c <= 2 * bytes && (best < 0 || c < ctx->spare[best].bytes)) best = i;
  }
  if(best >= 0) {
    DevBuf b = std::move(ctx->spare[best]);
    ctx->spare.erase(ctx->spare.begin() + best);
    return b;
  }
  DevBuf b;
  try {
    b.alloc(bytes);
  } catch(const Error& e) {
    if(e.code != MHIP_ERR_OUT_OF_MEMORY || ctx->spare.empty()) throw;
    (void)hipGetLastError();  // the failed hipMalloc is not a pending error
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->spare.clear();
    b.alloc(bytes);
  }
  return b;
}

void keep_buffer(mhip_context* ctx, DevBuf&& b) {
  if(!b.p) return;
  ctx->spare.push_back(std::move(b));
  // at most 8 buffers and a quarter of the device's memory, oldest dropped
  // first: kept Hamiltonians must not starve the next, differently sized one
  size_t free_b = 0, total_b = 0;
  MHIP_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  size_t held = 0;
  for(auto& k : ctx->spare) held += k.bytes;
  while(ctx->spare.size() > 1 && (ctx->spare.size() > 8 || held > total_b / 4)) {
    held -= ctx->spare.front().bytes;
    ctx->spare.erase(ctx->spare.begin());
  }
}

}  // namespace mhip

using namespace mhip;

extern "C" {

int mhip_get_version(void) { return MHIP_VERSION; }

const char* mhip_error_string(int code) {
  switch(code) {
    case MHIP_SUCCESS: return "success";
    case MHIP_ERR_INVALID_ARG: return "invalid argument";
    case MHIP_ERR_HIP: return "HIP runtime error";
    case MHIP_ERR_OUT_OF_MEMORY: return "out of memory";
    case MHIP_ERR_NO_INTEGRALS: return "no integrals set on the context";
    case MHIP_ERR_NOT_CONVERGED: return "Davidson Did Not Converge!";
    case MHIP_ERR_NO_GUESS: return "Davidson: No Guess Provided";
    case MHIP_ERR_UNSUPPORTED: return "Unsupported index_t";
    case MHIP_ERR_INDEX_OVERFLOW: return "index type overflow";
    case MHIP_ERR_NO_DEVICE: return "no HIP device";
    default: return "unknown error";
  }
}

const char* mhip_last_error(void) { return g_last_error; }

int mhip_context_create(int device, mhip_context** out) {
  return guarded([&] {
    MHIP_REQUIRE(out, MHIP_ERR_INVALID_ARG, "null output pointer");
    *out = nullptr;
    int ndev = 0;
    if(hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
      throw Error(MHIP_ERR_NO_DEVICE, "no HIP device visible");
    MHIP_REQUIRE(device >= 0 && device < ndev, MHIP_ERR_INVALID_ARG,
                 "device index out of range");
    auto* ctx = new mhip_context();
    ctx->device = device;
    try {
      MHIP_HIP_CHECK(hipSetDevice(device));
      init_noalloc_scratch(device);
      // Default priority stream (non-blocking with respect to stream 0).
      MHIP_HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    } catch(...) {
      delete ctx;
      throw;
    }
    *out = ctx;
  });
}

int mhip_context_destroy(mhip_context* ctx) {
  return guarded([&] {
    if(!ctx) return;
    (void)hipSetDevice(ctx->device);
    if(ctx->stream) {
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipStreamDestroy(ctx->stream);
    }
    const int device = ctx->device;
    delete ctx;
    device_cache_trim(device);
  });
}

int mhip_context_get_stream(mhip_context* ctx, void** stream) {
  return guarded([&] {
    MHIP_REQUIRE(ctx && stream, MHIP_ERR_INVALID_ARG, "null argument");
    *stream = (void*)ctx->stream;
  });
}

int mhip_context_set_build_path(mhip_context* ctx, int path) {
  return guarded([&] {
    MHIP_REQUIRE(ctx, MHIP_ERR_INVALID_ARG, "null context");
    MHIP_REQUIRE(path == MHIP_BUILD_AUTO || path == MHIP_BUILD_TILED ||
                     path == MHIP_BUILD_PRODUCT || path == MHIP_BUILD_SPIN_GROUPS,
                 MHIP_ERR_INVALID_ARG, "unknown build path");
    ctx->build_path = path;
  });
}

int mhip_context_release_cached(mhip_context* ctx) {
  return guarded([&] {
    use_device(ctx);
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->spare.clear();
  });
}

int mhip_synchronize(mhip_context* ctx) {
  return guarded([&] {
    use_device(ctx);
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

int mhip_get_timings(mhip_context* ctx, mhip_timings* out) {
  return guarded([&] {
    MHIP_REQUIRE(ctx && out, MHIP_ERR_INVALID_ARG, "null argument");
    *out = ctx->timings;
  });
}

int mhip_malloc(mhip_context* ctx, size_t bytes, void** dptr) {
  return guarded([&] {
    MHIP_REQUIRE(dptr, MHIP_ERR_INVALID_ARG, "null output pointer");
    use_device(ctx);
    *dptr = nullptr;
    if(bytes) *dptr = device_alloc(bytes);
  });
}

int mhip_free(mhip_context* ctx, void* dptr) {
  return guarded([&] {
    use_device(ctx);
    if(dptr) {
      MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
      device_free(dptr);
    }
  });
}

int mhip_memcpy_h2d(mhip_context* ctx, void* dst, const void* src,
                    size_t bytes) {
  return guarded([&] {
    use_device(ctx);
    if(!bytes) return;
    MHIP_REQUIRE(dst && src, MHIP_ERR_INVALID_ARG, "null pointer");
    MHIP_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

int mhip_memcpy_d2h(mhip_context* ctx, void* dst, const void* src,
                    size_t bytes) {
  return guarded([&] {
    use_device(ctx);
    if(!bytes) return;
    MHIP_REQUIRE(dst && src, MHIP_ERR_INVALID_ARG, "null pointer");
    d2h(ctx, dst, src, bytes);
  });
}

int mhip_memcpy_d2d(mhip_context* ctx, void* dst, const void* src,
                    size_t bytes) {
  return guarded([&] {
    use_device(ctx);
    if(!bytes) return;
    MHIP_REQUIRE(dst && src, MHIP_ERR_INVALID_ARG, "null pointer");
    MHIP_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  });
}

int mhip_memset(mhip_context* ctx, void* dst, int value, size_t bytes) {
  return guarded([&] {
    use_device(ctx);
    if(!bytes) return;
    MHIP_REQUIRE(dst, MHIP_ERR_INVALID_ARG, "null pointer");
    MHIP_HIP_CHECK(hipMemsetAsync(dst, value, bytes, ctx->stream));
  });
}

int mhip_set_integrals(mhip_context* ctx, int64_t norb, const double* T,
                       const double* V) {
  return guarded([&] {
    use_device(ctx);
    MHIP_REQUIRE(T && V, MHIP_ERR_INVALID_ARG, "null integral pointer");
    MHIP_REQUIRE(norb > 0 && norb <= 64, MHIP_ERR_INVALID_ARG,
                 "norb must be in [1, 64]");
    const size_t n = (size_t)norb, n2 = n * n, n3 = n2 * n, n4 = n3 * n;
    // generate_integral_intermediates (impl.hpp:25-70). Only subtraction
    // and exact scaling by 0.5 are involved, so the tables equal the
    // reference's bit for bit.
    std::vector<double> G(n4), G_red(n3), V_red(n3), G2(n2), V2(n2);
    auto idx = [&](size_t i, size_t j, size_t k, size_t l) {
      return i + n * j + n2 * k + n3 * l;
    };
    for(size_t l = 0; l < n; ++l)
      for(size_t k = 0; k < n; ++k)
        for(size_t j = 0; j < n; ++j)
          for(size_t i = 0; i < n; ++i)
            G[idx(i, j, k, l)] = V[idx(i, j, k, l)] - V[idx(i, l, k, j)];
    for(size_t j = 0; j < n; ++j)
      for(size_t i = 0; i < n; ++i)
        for(size_t k = 0; k < n; ++k) {
          G_red[k + n * i + n2 * j] = G[idx(k, k, i, j)];
          V_red[k + n * i + n2 * j] = V[idx(k, k, i, j)];
        }
    for(size_t j = 0; j < n; ++j)
      for(size_t i = 0; i < n; ++i) {
        G2[i + n * j] = 0.5 * G[idx(i, i, j, j)];
        V2[i + n * j] = V[idx(i, i, j, j)];
      }
    // V(pq|rs) stored equal to V(rs|pq) bit for bit (real integrals as the
    // FCIDUMP reader and the seeded generator produce them): element passes
    // may then read the pair-swapped entry where it is the contiguous one
    bool pair_sym = true;
    for(size_t l = 0; l < n && pair_sym; ++l)
      for(size_t k = 0; k < n && pair_sym; ++k)
        for(size_t j = 0; j < n && pair_sym; ++j)
          for(size_t i = 0; i < n; ++i)
            if(std::memcmp(&V[idx(i, j, k, l)], &V[idx(k, l, i, j)], sizeof(double))) {
              pair_sym = false;
              break;
            }
    Integrals& I = ctx->ints;
    I.norb = 0;
    I.pair_sym = pair_sym;
    I.T.alloc(n2 * 8);
    I.V.alloc(n4 * 8);
    I.G.alloc(n4 * 8);
    I.G_red.alloc(n3 * 8);
    I.V_red.alloc(n3 * 8);
    I.G2_red.alloc(n2 * 8);
    I.V2_red.alloc(n2 * 8);
    auto up = [&](DevBuf& d, const double* h, size_t cnt) {
      MHIP_HIP_CHECK(hipMemcpyAsync(d.p, h, cnt * 8, hipMemcpyHostToDevice, ctx->stream));
    };
    up(I.T, T, n2);
    up(I.V, V, n4);
    up(I.G, G.data(), n4);
    up(I.G_red, G_red.data(), n3);
    up(I.V_red, V_red.data(), n3);
    up(I.G2_red, G2.data(), n2);
    up(I.V2_red, V2.data(), n2);
    MHIP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    I.norb = norb;
  });
}

}  // extern "C"
