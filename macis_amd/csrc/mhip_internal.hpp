// Internal declarations shared by the HIP translation units of libmacis_hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "macis_hip.h"

namespace mhip {

// Error carrying one of the MHIP_ERR_* codes.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define MHIP_HIP_CHECK(expr)                                                 \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if(_e != hipSuccess) {                                                   \
      throw ::mhip::Error(_e == hipErrorOutOfMemory ? MHIP_ERR_OUT_OF_MEMORY \
                                                    : MHIP_ERR_HIP,          \
                          std::string(#expr) + ": " + hipGetErrorString(_e)); \
    }                                                                        \
  } while(0)

#define MHIP_REQUIRE(cond, code, msg)       \
  do {                                      \
    if(!(cond)) throw ::mhip::Error(code, msg); \
  } while(0)

// Runs `f` and converts exceptions to status codes (the C-ABI contract).
template <typename F>
int guarded(F&& f) {
  try {
    f();
    set_last_error("");
    return MHIP_SUCCESS;
  } catch(const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch(const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return MHIP_ERR_OUT_OF_MEMORY;
  } catch(const std::exception& e) {
    set_last_error(e.what());
    return MHIP_ERR_HIP;
  }
}

// Device memory of the library. Blocks of 64 MiB and more are not handed
// back to the driver on release but kept per device and reused for later
// requests of up to 1/8 less: the driver clears released VRAM before it can
// be mapped again, and a search or build that frees and re-allocates tens of
// GB per call would otherwise run behind that clearing (measured: the same
// 1.1e9-candidate ASCI search 0.65 s on fresh memory, 1.9-4.1 s on the next
// calls). Kept blocks are freed when an allocation fails (which is then
// retried) and when a context of the device is destroyed. device_free waits
// for the device, as hipFree does, before a block can be handed out again.
void* device_alloc(size_t bytes);
void* device_alloc_raw(size_t bytes);
void device_free(void* p) noexcept;
void device_cache_trim(int device) noexcept;
// Free device memory of the current device as hipMemGetInfo reports it plus
// the kept blocks, which an allocation would release: what sizing decisions
// (staging pools, layout copies) may plan with.
void device_mem_info(size_t* free_b, size_t* total_b);

// Owning device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b) { alloc(b); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if(this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t b) {
    release();
    if(b) p = device_alloc(b);
    bytes = b;
  }
  // Grow-only reuse.
  void reserve(size_t b) {
    if(b > bytes) alloc(b);
  }
  void release() {
    if(p) device_free(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Pinned host staging buffer (grow-only) for small device-to-host reads:
// pageable copies go through the runtime's own staging and have been seen
// to stall for milliseconds.
struct PinnedBuf {
  void* p = nullptr;
  size_t bytes = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() {
    if(p) (void)hipHostFree(p);
  }
  void reserve(size_t b) {
    if(b <= bytes) return;
    if(p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    MHIP_HIP_CHECK(hipHostMalloc(&p, b, hipHostMallocDefault));
    bytes = b;
  }
};

// Integral tables on the device (column-major, see macis_hip.h).
struct Integrals {
  int64_t norb = 0;
  bool pair_sym = false;  // V(pq|rs) == V(rs|pq) bit for bit
  DevBuf T, V, G, G_red, V_red, G2_red, V2_red;
};

// Event pair timer on a stream.
struct Timer {
  hipEvent_t a = nullptr, b = nullptr;
  hipStream_t s = nullptr;
  explicit Timer(hipStream_t st) : s(st) {
    MHIP_HIP_CHECK(hipEventCreate(&a));
    MHIP_HIP_CHECK(hipEventCreate(&b));
  }
  ~Timer() {
    if(a) (void)hipEventDestroy(a);
    if(b) (void)hipEventDestroy(b);
  }
  void start() { MHIP_HIP_CHECK(hipEventRecord(a, s)); }
  void stop() { MHIP_HIP_CHECK(hipEventRecord(b, s)); }
  double ms() {
    MHIP_HIP_CHECK(hipEventSynchronize(b));
    float t = 0.f;
    MHIP_HIP_CHECK(hipEventElapsedTime(&t, a, b));
    return t;
  }
};

}  // namespace mhip

namespace mhip {
// Sliced (SELL-64) copy of a square CSR matrix for the single-vector SpMV:
// slice s holds rows [64 s, 64 s + 64), its entries stored slot-major
// (value of entry j of row 64 s + l at soff[s] + 64 j + l), every row
// padded to the slice's longest row with value 0; columns 0-based.
// A slice whose columns at every slot lie within 65,536 of each other is
// stored narrow: per slot a base column (the smallest at that slot) and per
// entry a 16-bit offset from it, two slots packed per 32-bit word (word
// soff[s] + 64 p + l holds slot 2p in its low half and slot 2p + 1 in its
// high half for row 64 s + l), padding at offset 0. Other slices keep
// 32-bit columns at soff[s] + 64 j + l, padding at column = row.
struct SellCopy {
  int64_t nsl = 0, entries = 0;
  int64_t narrow_entries = 0;  // entries of the narrow slices
  DevBuf meta;        // int64 soff[nsl], total, then int32 slen[nsl]
  DevBuf cols, vals;  // column words (4 bytes reserved per entry) / double [entries]
  DevBuf narrow;      // int32[nsl]: 1 when slice s is stored narrow
  DevBuf base;        // int32[entries / 64]: slot bases, slot j of slice s at soff[s] / 64 + j
  const int64_t* soff() const { return meta.as<int64_t>(); }
  const int32_t* slen() const { return reinterpret_cast<const int32_t*>(soff() + nsl + 1); }
  // bytes one single-vector SpMV streams from the copy: values, column
  // words and the narrow slices' slot bases
  double streamed_bytes() const {
    const double nw = (double)narrow_entries;
    return 8. * (double)entries + 4. * ((double)entries - nw) + 2. * nw + nw / 16.;
  }
};
// Column-panel split of a CSR for single-vector SpMVs whose x vector
// outgrows one XCD's L2: columns [bounds[p], bounds[p+1]) form panel p
// (np panels of about equal nonzero counts); off holds, per row, the
// start of each panel's entries relative to the row start (columns are
// ascending, so a panel's entries are one contiguous run of the row).
constexpr int kMaxPanels = 32;
struct PanelSplit {
  int np = 8;                   // panels, a multiple of 8 (XCD x takes x, x + 8, ...)
  std::vector<int64_t> bounds;  // np + 1 column bounds
  DevBuf off;                   // int32[m * np]
  DevBuf part;                  // double[np * m] partial sums
};
// Panel-major copy of a square CSR for single-vector SpMVs whose x vector
// is too large for an XCD's L2: panel p holds columns [p T, (p + 1) T) with
// T = kTileCols, its entries stored contiguously row after row (columns
// ascending within a row) as 16-bit local columns and values; rowoff[p][r]
// (m + 1 per panel) is the first entry of row r in panel p relative to the
// panel's first entry base[p]. A workgroup stages a panel's share of x in
// LDS and streams that panel's entries of its rows.
constexpr int kTileCols = 12288;  // 96 KB of x per panel
struct XTiles {
  int np = 0;
  int64_t entries = 0;
  std::vector<int64_t> base_h;  // host copy of base (np + 1)
  DevBuf base;                  // int64[np + 1]
  DevBuf rowoff;                // uint32[np * (m + 1)]
  DevBuf cols;                  // uint16[entries]
  DevBuf vals;                  // double[entries]
};
}  // namespace mhip

struct mhip_context {
  int device = 0;
  hipStream_t stream = nullptr;
  mhip::Integrals ints;
  mhip_timings timings{};
  // Reusable scratch (partials, flags, temp storage of device primitives).
  mhip::DevBuf scratch_a, scratch_b, scratch_c;
  std::vector<char> host_scratch;
  mhip::PinnedBuf pinned;  // staging of small device-to-host reads
  // device buffers kept for reuse (storage of the last internal Hamiltonian
  // of selected_ci_diag, Davidson subspaces): freeing and re-mapping large
  // buffers on every solve stalls the next blocking call for milliseconds
  std::vector<mhip::DevBuf> spare;
  int build_path = 0;  // MHIP_BUILD_*
};

struct mhip_csr {
  int device = 0;
  int64_t m = 0, n = 0, nnz = 0;
  int64_t indexing = 0;
  mhip::DevBuf rowptr;  // int64[m+1]
  mhip::DevBuf colind;  // int32[nnz]
  mhip::DevBuf nzval;   // double[nnz]
  // optional sliced copy used by single-vector SpMVs (prepare_sell)
  mutable std::unique_ptr<mhip::SellCopy> sell;
  // optional column-panel split used by single-vector SpMVs (prepare_panels)
  mutable std::unique_ptr<mhip::PanelSplit> panels;
  // optional panel-major copy with LDS-staged x tiles (prepare_xtiles)
  mutable std::unique_ptr<mhip::XTiles> xtiles;
};

namespace mhip {
// Makes the context's device current.
void use_device(mhip_context* ctx);
// Scratch array of the allocation-free entry points: its size, the devices
// the address table covers, and this device's address of it (nullptr until a
// context was created on the device).
constexpr size_t kNoAllocScratchBytes = size_t(2) << 20;
constexpr int kMaxDevices = 64;
void* noalloc_scratch(int device);
// Device-to-host copy on the context stream, waited for; copies up to 4 MiB
// are staged through the context's pinned buffer.
void d2h(mhip_context* ctx, void* dst, const void* src, size_t bytes);
// A device buffer of at least `bytes`: a kept one of capacity in
// [bytes, 2 bytes] if the context holds one, else a new allocation (kept
// buffers are dropped first when the device is out of memory).
DevBuf take_buffer(mhip_context* ctx, size_t bytes);
// Keeps `b` on the context for take_buffer (at most 8 buffers).
void keep_buffer(mhip_context* ctx, DevBuf&& b);

// Hamiltonian construction shared by mhip_ham_build and selected_ci_diag.
mhip_csr* ham_build(mhip_context* ctx, int W, const uint64_t* d_bra,
                    int64_t nbra, const uint64_t* d_ket, int64_t nket,
                    double thresh);
void ham_diagonal(mhip_context* ctx, int W, const uint64_t* d_dets,
                  int64_t ndets, double* d_diag);
// y = alpha*A*x + beta*y on the context stream (one column).
void spmv(mhip_context* ctx, const mhip_csr* A, double alpha, const double* x,
          double beta, double* y);
void extract_diagonal(mhip_context* ctx, const mhip_csr* A, double* d_D);
// Attaches the sliced copy to a square matrix when the device has room for
// it and (unless any_padding) padding adds at most a quarter of the
// nonzeros; true if the matrix has one afterwards; spmv() then uses it.
bool prepare_sell(mhip_context* ctx, const mhip_csr* A, bool any_padding = false);
// Attaches the column-panel split to a square matrix (when it has no sliced
// copy); with any_size false only when x exceeds kPanelMinCols entries.
// True if the matrix has one afterwards; spmv() then uses it.
constexpr int64_t kPanelMinCols = 262144;
bool prepare_panels(mhip_context* ctx, const mhip_csr* A, bool any_size = false);
// Attaches the panel-major copy with LDS-staged x tiles to a square matrix
// (no sliced copy) when the device has room for it (10 bytes per nonzero
// and 4 per row and panel, plus 1 GiB); with any_size false only when x
// exceeds kPanelMinCols entries. True if the matrix has one afterwards;
// spmv() then uses it.
bool prepare_xtiles(mhip_context* ctx, const mhip_csr* A, bool any_size = false);
// y = A x on the sliced copy (A must have one); with conv, rows r with
// conv[row_inst[r]] set are skipped (y keeps its value there).
void spmv_sliced_masked(mhip_context* ctx, const mhip_csr* A, const double* x, double* y,
                        const int32_t* row_inst, const int32_t* conv);
// Davidson core; returns convergence flag.
bool davidson(mhip_context* ctx, const mhip_csr* A, int64_t max_m,
              int64_t max_restarts, double tol, const double* d_D, double* d_X,
              double* E, int64_t* iters);
void validate_dets(mhip_context* ctx, int W, const uint64_t* d, int64_t n);
struct IntsView;
IntsView view(const Integrals& I);
// Product-space structure of a determinant list: n = nruns * L, det
// r*L + k = (outer string O[r], inner string I[k]). Device tables: outer
// neighbours of every run (r' << 2 | excitation level, r' ascending), their
// per-level counts, and the inner neighbour lists of every inner string at
// levels <= 1 and <= 2 (ascending k').
struct ProductSpace {
  int64_t L = 0, nruns = 0, max_seg = 0;  // max_seg: most outer neighbours of a run
  bool outer_alpha = true;
  DevBuf arena, lists;  // owners of the tables below
  const uint64_t *ostr = nullptr, *istr = nullptr;  // outer / inner strings
  const int32_t *ocnt = nullptr;                    // [3*nruns] neighbours per level
  const int32_t *onb_ptr = nullptr, *onb = nullptr; // outer lists, r' << 2 | level
  const int32_t *ptr1 = nullptr, *idx1 = nullptr;   // inner lists, levels <= 1
  const int32_t *ptr2 = nullptr, *idx2 = nullptr;   // inner lists, levels <= 2
  // element tables (filled per build from the integrals): per outer entry
  // the double element or the single's partial sum and packed excitation;
  // per inner entry likewise (level <= 2 list) and the packed single
  // (level <= 1 list)
  double *oval = nullptr, *ival2 = nullptr;
  int32_t *oexc = nullptr, *iexc1 = nullptr, *iexc2 = nullptr;
};
// Recognises a product-space list (host + device checks); false otherwise.
bool detect_product_space(mhip_context* ctx, int W, const uint64_t* d, int64_t n,
                          ProductSpace& ps);
// Structural row counts of a product-space list.
void product_count(mhip_context* ctx, const ProductSpace& ps, int64_t n, uint32_t* counts);
// Fill + Slater-Condon elements in one pass (kept counts and dropped flag
// as the element pass of the tiled path).
void product_fill_elements(mhip_context* ctx, int W, const ProductSpace& ps, const IntsView& I,
                           int64_t n, const double* diag, const int64_t* rowptr, int32_t* colind,
                           double* nzval, double thresh, int64_t* kept, int* dropped);
// Product spaces of the instances of a batched build: instances with the
// same determinant list share one ProductSpace; every instance has its own
// element tables and device view (first row, structure, integrals).
struct ProductBatch {
  std::vector<std::unique_ptr<ProductSpace>> spaces;
  std::vector<int> space_of;  // instance -> spaces index
  DevBuf tables, views, run_off, inner_off;
  int ninst = 0;
  int64_t max_seg = 0;  // most outer neighbours of a run over the instances
  int64_t total_runs = 0, total_inner = 0;
};
// list_id[i]: the first instance whose determinant list equals instance i's
// (i itself when none). False when an instance is not a product space.
bool prepare_batch_product(mhip_context* ctx, int W, const uint64_t* d_dets,
                           const std::vector<int64_t>& off, const std::vector<int>& list_id,
                           const IntsView* d_views, ProductBatch& pb);
void batch_product_count(mhip_context* ctx, const ProductBatch& pb, const int32_t* row_inst,
                         int64_t n, uint32_t* counts);
void batch_product_fill_elements(mhip_context* ctx, const ProductBatch& pb,
                                 const int32_t* row_inst, int64_t n, const double* diag,
                                 const int64_t* rowptr, int32_t* colind, double* nzval,
                                 double thresh, int64_t* kept, int* dropped);
// Optional work-item table / per-instance integrals of a batched build.
// Spin-string groups of a square list with fixed alpha and beta electron
// counts (mhip_spingroup.hip): per spin, the groups of equal string (start
// of each group's members in the sorted order), the total membership of
// each group's single-excitation neighbour groups, and each group's flat
// list (own members, then the neighbour groups' members) as other-spin
// string + determinant index.
struct GroupSide {
  int64_t ng = 0, total = 0;
  DevBuf start, nsize, foff, gof, fw, fi;
  DevBuf perm;  // determinants in group order (ascending index within a group)
};
struct SpinGroups {
  int64_t n = 0;
  GroupSide A, B;
  DevBuf cnt_b;  // per row: hits of its beta-side scan
};
// False when the counts differ between determinants or the flat lists
// would exceed max_bytes.
bool prepare_spin_groups(mhip_context* ctx, int W, const uint64_t* d, int64_t n,
                         double max_bytes, SpinGroups& g);
void spin_group_count(mhip_context* ctx, int W, const SpinGroups& g, const uint64_t* d,
                      int64_t n, uint32_t* counts);
void spin_group_fill(mhip_context* ctx, int W, const SpinGroups& g, const uint64_t* d, int64_t n,
                     const int64_t* rowptr, int32_t* col);
// Ascending columns within every row (segmented radix sort over the bits
// of ncols); col may be swapped for the sorted buffer.
void sort_row_columns(mhip_context* ctx, int64_t n, const int64_t* rowptr, int64_t nnz,
                      int64_t ncols, DevBuf& col);

struct BuildExtra {
  const int64_t* items = nullptr;  // DEVICE, 4 int64 per item
  int64_t nitems = 0;
  const IntsView* views = nullptr;  // DEVICE, one per instance
  const int32_t* row_inst = nullptr;  // DEVICE, instance of every row
  const ProductSpace* product = nullptr;  // structure passes of a product list
  const ProductBatch* bproduct = nullptr;  // per-instance product spaces of a batch
  const SpinGroups* groups = nullptr;      // spin-string groups of a selected-CI list
  double detect_ms = 0.;                  // host time spent recognising it
};
template <int W>
void build_impl(mhip_context* ctx, const uint64_t* d_bra, int64_t nbra, const uint64_t* d_ket,
                int64_t nket, double thresh, mhip_csr* csr, const BuildExtra& ex);
// Host symmetric eigensolver on the lower triangle (cyclic Jacobi).
void syev_lower(int n, double* A, int lda, double* W);
}  // namespace mhip
