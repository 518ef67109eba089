// Internal definitions shared by the HIP translation units of libmatrixgraph_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "matrixgraph_amd.h"

namespace mgamd {

// ---------------------------------------------------------------------------
// Errors: every C entry point runs its body under guarded(), which maps the
// exceptions below to status codes and records the message for mg_last_error().
// ---------------------------------------------------------------------------
struct Failure : std::runtime_error {
  int code;
  Failure(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Failure(code, msg); }
[[noreturn]] inline void invalid(const std::string& msg) { throw Failure(MG_ERR_INVALID, msg); }

inline void hip_check(hipError_t e, const char* what, const char* file, int line) {
  if (e == hipSuccess) return;
  (void)hipGetLastError();  // clear the sticky launch error so later calls report their own
  const int code = (e == hipErrorOutOfMemory) ? MG_ERR_OOM : MG_ERR_HIP;
  throw Failure(code, std::string(what) + ": " + hipGetErrorString(e) + " (" + file + ":" +
                          std::to_string(line) + ")");
}
#define MG_HIP(call) ::mgamd::hip_check((call), #call, __FILE__, __LINE__)
#define MG_LAUNCH_CHECK() ::mgamd::hip_check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)

template <class F>
int guarded(F&& f) {
  try {
    f();
    return MG_OK;
  } catch (const Failure& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return MG_ERR_OOM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return MG_ERR_INTERNAL;
  } catch (...) {
    set_last_error("unknown failure");
    return MG_ERR_INTERNAL;
  }
}

#define MG_REQUIRE(cond, msg) \
  do {                        \
    if (!(cond)) ::mgamd::invalid(msg); \
  } while (0)

// ---------------------------------------------------------------------------
// Device buffer (RAII, move-only). Allocation is explicit; freeing happens in
// the destructor. Only setup paths allocate; timed hop loops reuse buffers.
// ---------------------------------------------------------------------------
template <class T>
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  void alloc(size_t n) {
    release();
    if (n == 0) return;
    MG_HIP(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)));
    n_ = n;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  size_t bytes() const { return n_ * sizeof(T); }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

// Stream-ordered temporary from the device's default memory pool: hipMallocAsync and
// hipFreeAsync on one stream (the caller's context stream). Kernels queued on that stream
// before the buffer goes out of scope finish before its memory is reused, so a call can
// return without synchronising just to free its temporaries. mg_context_create keeps freed
// memory in the pool (release threshold: never), so repeated calls reuse it instead of
// going back to the driver. Counterpart of DeviceOwnedBuffer::Init (device_buffer.cuh:92-106)
// for temporaries; long-lived results keep DevBuf.
template <class T>
class StreamBuf {
 public:
  StreamBuf() = default;
  StreamBuf(size_t n, hipStream_t s) { alloc(n, s); }
  ~StreamBuf() { release(); }
  // At least one element, so get() is never null.
  void alloc(size_t n, hipStream_t s) {
    release();
    s_ = s;
    if (n == 0) n = 1;
    MG_HIP(hipMallocAsync(reinterpret_cast<void**>(&p_), n * sizeof(T), s_));
    n_ = n;
  }
  void release() {
    if (p_) (void)hipFreeAsync(p_, s_);
    p_ = nullptr;
    n_ = 0;
  }
  StreamBuf(const StreamBuf&) = delete;
  StreamBuf& operator=(const StreamBuf&) = delete;
  StreamBuf(StreamBuf&& o) noexcept : p_(o.p_), n_(o.n_), s_(o.s_) { o.p_ = nullptr; o.n_ = 0; }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  size_t bytes() const { return n_ * sizeof(T); }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
  hipStream_t s_ = nullptr;
};

// Pinned host staging buffer (hipHostMalloc; never hipHostRegister).
template <class T>
class PinnedBuf {
 public:
  PinnedBuf() = default;
  explicit PinnedBuf(size_t n) { alloc(n); }
  ~PinnedBuf() { if (p_) (void)hipHostFree(p_); }
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  void alloc(size_t n) {
    if (p_) (void)hipHostFree(p_);
    p_ = nullptr;
    if (n == 0) return;
    MG_HIP(hipHostMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T), hipHostMallocDefault));
  }
  T* get() const { return p_; }

 private:
  T* p_ = nullptr;
};

// Scoped device switch: restores the caller's current device.
struct DeviceGuard {
  int prev = 0;
  explicit DeviceGuard(int dev) {
    MG_HIP(hipGetDevice(&prev));
    if (prev != dev) MG_HIP(hipSetDevice(dev));
  }
  ~DeviceGuard() { (void)hipSetDevice(prev); }
};

}  // namespace mgamd

// ---------------------------------------------------------------------------
// Handle definitions (opaque to C callers).
// ---------------------------------------------------------------------------
struct mg_context {
  int device = 0;
  int n_cu = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  mgamd::DevBuf<unsigned long long> counters;  // small scratch for reductions (host-synchronised calls)
  hipEvent_t timer[2] = {nullptr, nullptr};    // mg_context_timer_start / _stop
  // Per-call device temporaries of asynchronous calls on `stream` (calls on one context run
  // in stream order, so the next call may reuse them); grown by ctx_scratch, stream-ordered.
  mgamd::StreamBuf<uint64_t> scratch;
};

struct mg_tiled_matrix {
  int device = 0;
  uint32_t n_strips = 0;
  uint32_t n_nz_tile = 0;
  uint32_t tile_size = 64;
  mgamd::DevBuf<uint32_t> offset_row;  // n_strips + 1
  mgamd::DevBuf<uint32_t> row_idx;     // n_nz_tile
  mgamd::DevBuf<uint32_t> col_idx;     // n_nz_tile
  mgamd::DevBuf<uint64_t> data;        // 64 * n_nz_tile
  // Bumped by every library call that writes this matrix's tiles (the numeric passes into C).
  uint64_t version = 0;
  // mg_tiled_matrix_prepare_operand: the product path's per-tile payloads (16 B) and set-bit
  // counts, valid while prep_version == version.
  mgamd::DevBuf<uint64_t> prep_payload;
  mgamd::DevBuf<uint32_t> prep_bits;
  uint64_t prep_version = ~0ull;
};

struct mg_edge_list {
  int device = 0;
  uint32_t n_vertices = 0;
  uint64_t n_edges = 0;
  mgamd::DevBuf<uint64_t> keys;  // (src << 32) | dst, ascending, unique
};

namespace mgamd {

// At least n words of ctx->scratch; growing it first waits for the work queued on the
// context's stream, which may still read the old buffer.
inline uint64_t* ctx_scratch(mg_context* ctx, size_t n) {
  // The old buffer is freed in stream order: work already queued that reads it finishes first.
  if (ctx->scratch.size() < n) ctx->scratch.alloc(n, ctx->stream);
  return ctx->scratch.get();
}

// Static payload scratch of mg_multiply_numeric_noalloc (mg_spgemm.hip): 32 MiB per device,
// 16 bytes per tile of both operands plus 16 (kNoAllocMaxTiles tiles in A and Bt together).
constexpr size_t kNoAllocScratchBytes = 32ull << 20;
constexpr uint64_t kNoAllocMaxTiles = kNoAllocScratchBytes / 16 - 1;
constexpr int kMaxDevices = 64;
// Looks up this device's address of the static scratch (device current); once per device.
void noalloc_scratch_resolve(int device);
// The address looked up by noalloc_scratch_resolve for `device`.
void* noalloc_scratch(int device);

// Input and output widths of a cost-model MLP (mg_mlp.hip).
void mlp_dims(const mg_mlp* p, uint32_t* n_in, uint32_t* n_out);

// Helpers implemented in mg_matrix.hip / mg_graph.hip, shared across units.
void alloc_tiled_matrix(mg_tiled_matrix* m, uint32_t n_strips, uint32_t n_nz_tile);
uint64_t device_popcount(mg_context* ctx, const uint64_t* data, uint64_t n_words);
// Sum of n u32 values, in 64 bits (for totals that a u32 scan would wrap).
uint64_t device_sum_u32(mg_context* ctx, const uint32_t* data, uint64_t n);
// Exclusive scan of n u32 counts into out[0..n] (out[n] = total), on ctx->stream.
void exclusive_scan_u32(mg_context* ctx, const uint32_t* in, uint32_t* out, uint32_t n);
void exclusive_scan_u32_on(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t n);
// Sort u64 keys ascending and drop duplicates; returns the unique count.
uint64_t sort_unique_u64(mg_context* ctx, DevBuf<uint64_t>& keys, uint64_t n, int end_bit);

inline unsigned grid_for(uint64_t work, unsigned per_block, unsigned cap = 1u << 20) {
  uint64_t g = (work + per_block - 1) / per_block;
  if (g == 0) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

}  // namespace mgamd
