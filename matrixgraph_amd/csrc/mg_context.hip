// Context lifecycle, error reporting and host-only helpers of the C-ABI.
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_set>

#include "mg_device.hpp"
#include "mg_internal.hpp"

namespace mgamd {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace mgamd

using namespace mgamd;

extern "C" const char* mg_version(void) { return "matrixgraph_amd 0.1.0 (gfx950)"; }

extern "C" const char* mg_last_error(void) { return g_last_error.c_str(); }

extern "C" int mg_context_create(int device, mg_context** out) {
  return guarded([&] {
    MG_REQUIRE(out != nullptr, "mg_context_create: out is NULL");
    *out = nullptr;
    if (device < 0) {
      // Same role as MATRIXGRAPH_CUDA_DEVICE in core/go_api/matrixgraph_go_api.cu:23-36.
      const char* env = std::getenv("MATRIXGRAPH_HIP_DEVICE");
      device = env ? std::atoi(env) : 0;
    }
    int n_dev = 0;
    MG_HIP(hipGetDeviceCount(&n_dev));
    if (device >= n_dev)
      invalid("mg_context_create: device " + std::to_string(device) + " not present (" +
              std::to_string(n_dev) + " devices)");
    DeviceGuard g(device);
    noalloc_scratch_resolve(device);
    {
      // Temporaries come from the device's default pool (StreamBuf); keep freed memory there.
      hipMemPool_t pool = nullptr;
      MG_HIP(hipDeviceGetDefaultMemPool(&pool, device));
      uint64_t keep = ~0ull;
      MG_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
    }
    hipDeviceProp_t prop;
    MG_HIP(hipGetDeviceProperties(&prop, device));
    auto* ctx = new mg_context();
    ctx->device = device;
    ctx->n_cu = prop.multiProcessorCount;
    hipError_t e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete ctx;
      MG_HIP(e);
    }
    ctx->stream = ctx->own_stream;
    try {
      ctx->counters.alloc(64);
    } catch (...) {
      (void)hipStreamDestroy(ctx->own_stream);
      delete ctx;
      throw;
    }
    *out = ctx;
  });
}

extern "C" int mg_context_destroy(mg_context* ctx) {
  return guarded([&] {
    if (!ctx) return;
    DeviceGuard g(ctx->device);
    ctx->scratch.release();  // stream-ordered: before the stream goes away
    (void)hipStreamSynchronize(ctx->stream);
    ctx->counters.release();
    for (hipEvent_t& e : ctx->timer)
      if (e) (void)hipEventDestroy(e);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
  });
}

extern "C" int mg_context_set_stream(mg_context* ctx, void* stream) {
  return guarded([&] {
    MG_REQUIRE(ctx != nullptr, "mg_context_set_stream: ctx is NULL");
    const hipStream_t next = stream ? static_cast<hipStream_t>(stream) : ctx->own_stream;
    if (next != ctx->stream) {
      // Work queued on the previous stream may still read the context's scratch, which the
      // next call (ordered on the new stream only) may overwrite or regrow.
      DeviceGuard g(ctx->device);
      ctx->scratch.release();  // freed on the stream that used it; the next call allocates on the new one
      MG_HIP(hipStreamSynchronize(ctx->stream));
    }
    ctx->stream = next;
  });
}

extern "C" int mg_context_synchronize(mg_context* ctx) {
  return guarded([&] {
    MG_REQUIRE(ctx != nullptr, "mg_context_synchronize: ctx is NULL");
    DeviceGuard g(ctx->device);
    MG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

extern "C" int mg_context_device(const mg_context* ctx, int* device, int* n_compute_units) {
  return guarded([&] {
    MG_REQUIRE(ctx != nullptr, "mg_context_device: ctx is NULL");
    if (device) *device = ctx->device;
    if (n_compute_units) *n_compute_units = ctx->n_cu;
  });
}

// Floyd's algorithm (Bentley & Floyd, "Programming pearls: a sample of
// brilliance", CACM 1987): for j = n-count .. n-1 draw r in [0, j]; take r unless
// already taken, else take j. Draw i uses SplitMix64 counter i of the
// "sources" stream; r = (draw >> 11) % (j+1).
extern "C" int mg_sample_sources(uint32_t n_vertices, uint32_t count, uint64_t seed, uint32_t* out) {
  return guarded([&] {
    MG_REQUIRE(out != nullptr || count == 0, "mg_sample_sources: out is NULL");
    MG_REQUIRE(count <= n_vertices, "mg_sample_sources: count exceeds n_vertices");
    const uint64_t key = stream_key(seed, kTagSources);
    std::unordered_set<uint32_t> taken;
    taken.reserve(count * 2 + 1);
    uint32_t k = 0;
    for (uint64_t j = static_cast<uint64_t>(n_vertices) - count; j < n_vertices; ++j, ++k) {
      const uint64_t r = (sm64_at(key, k) >> 11) % (j + 1);
      const uint32_t pick = taken.count(static_cast<uint32_t>(r)) ? static_cast<uint32_t>(j)
                                                                   : static_cast<uint32_t>(r);
      taken.insert(pick);
      out[k] = pick;
    }
  });
}

extern "C" int mg_context_stream(const mg_context* ctx, void** stream) {
  return guarded([&] {
    MG_REQUIRE(ctx && stream, "mg_context_stream: NULL argument");
    *stream = ctx->stream;
  });
}

extern "C" int mg_buffer_alloc(mg_context* ctx, size_t bytes, void** dptr) {
  return guarded([&] {
    MG_REQUIRE(ctx && dptr, "mg_buffer_alloc: NULL argument");
    *dptr = nullptr;
    DeviceGuard g(ctx->device);
    void* p = nullptr;
    MG_HIP(hipMalloc(&p, bytes ? bytes : 1));
    const hipError_t e = hipMemsetAsync(p, 0, bytes ? bytes : 1, ctx->stream);
    if (e == hipSuccess) {
      MG_HIP(hipStreamSynchronize(ctx->stream));
    } else {
      (void)hipFree(p);
      MG_HIP(e);
    }
    *dptr = p;
  });
}

extern "C" int mg_buffer_free(mg_context* ctx, void* dptr) {
  return guarded([&] {
    MG_REQUIRE(ctx != nullptr, "mg_buffer_free: NULL ctx");
    if (!dptr) return;
    DeviceGuard g(ctx->device);
    MG_HIP(hipStreamSynchronize(ctx->stream));
    MG_HIP(hipFree(dptr));
  });
}

extern "C" int mg_buffer_upload(mg_context* ctx, void* dptr, const void* host, size_t bytes) {
  return guarded([&] {
    MG_REQUIRE(ctx && dptr && (host || bytes == 0), "mg_buffer_upload: NULL argument");
    DeviceGuard g(ctx->device);
    if (bytes) MG_HIP(hipMemcpyAsync(dptr, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    MG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

extern "C" int mg_buffer_download(mg_context* ctx, void* host, const void* dptr, size_t bytes) {
  return guarded([&] {
    MG_REQUIRE(ctx && dptr && (host || bytes == 0), "mg_buffer_download: NULL argument");
    DeviceGuard g(ctx->device);
    if (bytes) MG_HIP(hipMemcpyAsync(host, dptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
    MG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

extern "C" int mg_buffer_copy(mg_context* ctx, void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    MG_REQUIRE(ctx && dst && src, "mg_buffer_copy: NULL argument");
    DeviceGuard g(ctx->device);
    if (bytes) MG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  });
}

extern "C" int mg_context_timer_start(mg_context* ctx) {
  return guarded([&] {
    MG_REQUIRE(ctx != nullptr, "mg_context_timer_start: ctx is NULL");
    DeviceGuard g(ctx->device);
    for (hipEvent_t& e : ctx->timer)
      if (!e) MG_HIP(hipEventCreate(&e));
    MG_HIP(hipEventRecord(ctx->timer[0], ctx->stream));
  });
}

extern "C" int mg_context_timer_stop(mg_context* ctx, double* ms) {
  return guarded([&] {
    MG_REQUIRE(ctx && ms, "mg_context_timer_stop: NULL argument");
    MG_REQUIRE(ctx->timer[0] && ctx->timer[1], "mg_context_timer_stop: timer not started");
    DeviceGuard g(ctx->device);
    MG_HIP(hipEventRecord(ctx->timer[1], ctx->stream));
    MG_HIP(hipEventSynchronize(ctx->timer[1]));
    float t = 0;
    MG_HIP(hipEventElapsedTime(&t, ctx->timer[0], ctx->timer[1]));
    *ms = t;
  });
}
