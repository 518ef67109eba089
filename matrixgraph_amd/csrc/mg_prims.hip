// Setup-path primitives: scans and sorts (hipCUB), popcount reduction.
// None of these run inside a walk hop; they build indices and inputs.
#include <hipcub/hipcub.hpp>

#include "mg_device.hpp"
#include "mg_internal.hpp"

namespace mgamd {

namespace {

__global__ void __launch_bounds__(256) popcount_kernel(const uint64_t* __restrict__ data, uint64_t n,
                                                       unsigned long long* __restrict__ out) {
  uint64_t local = 0;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    local += __popcll(data[i]);
  local = wave_sum(local);
  if (lane_id() == 0 && local) atomicAdd(out, static_cast<unsigned long long>(local));
}

__global__ void __launch_bounds__(256) sum_u32_kernel(const uint32_t* __restrict__ data, uint64_t n,
                                                      unsigned long long* __restrict__ out) {
  uint64_t local = 0;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) local += data[i];
  local = wave_sum(local);
  if (lane_id() == 0 && local) atomicAdd(out, static_cast<unsigned long long>(local));
}

__global__ void mark_dups_kernel(const uint64_t* __restrict__ keys, uint64_t n,
                                 uint32_t* __restrict__ keep) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x) keep[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

}  // namespace

uint64_t device_popcount(mg_context* ctx, const uint64_t* data, uint64_t n_words) {
  unsigned long long* d = ctx->counters.get();
  MG_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), ctx->stream));
  if (n_words) {
    popcount_kernel<<<grid_for(n_words, 256 * 16, 8192), 256, 0, ctx->stream>>>(data, n_words, d);
    MG_LAUNCH_CHECK();
  }
  unsigned long long h = 0;
  MG_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  MG_HIP(hipStreamSynchronize(ctx->stream));
  return h;
}

uint64_t device_sum_u32(mg_context* ctx, const uint32_t* data, uint64_t n) {
  unsigned long long* d = ctx->counters.get();
  MG_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), ctx->stream));
  if (n) {
    sum_u32_kernel<<<grid_for(n, 256 * 16, 8192), 256, 0, ctx->stream>>>(data, n, d);
    MG_LAUNCH_CHECK();
  }
  unsigned long long h = 0;
  MG_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  MG_HIP(hipStreamSynchronize(ctx->stream));
  return h;
}

void exclusive_scan_u32_on(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t n) {
  // out[0..n-1] = exclusive prefix, out[n] = total: scan n+1 items with a zero appended
  // is avoided by scanning n items then adding the last count on device.
  size_t temp_bytes = 0;
  MG_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, temp_bytes, in, out + 1, n, st));
  StreamBuf<uint8_t> temp(temp_bytes, st);
  MG_HIP(hipMemsetAsync(out, 0, sizeof(uint32_t), st));
  if (n) MG_HIP(hipcub::DeviceScan::InclusiveSum(temp.get(), temp_bytes, in, out + 1, n, st));
}

void exclusive_scan_u32(mg_context* ctx, const uint32_t* in, uint32_t* out, uint32_t n) {
  exclusive_scan_u32_on(ctx->stream, in, out, n);
}

uint64_t sort_unique_u64(mg_context* ctx, DevBuf<uint64_t>& keys, uint64_t n, int end_bit) {
  if (n == 0) return 0;
  MG_REQUIRE(n < (1ull << 31), "sort_unique_u64: more than 2^31 keys");
  const int ni = static_cast<int>(n);
  DevBuf<uint64_t> alt(n);
  hipcub::DoubleBuffer<uint64_t> db(keys.get(), alt.get());
  size_t temp_bytes = 0;
  MG_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, temp_bytes, db, ni, 0, end_bit, ctx->stream));
  StreamBuf<uint8_t> temp(temp_bytes, ctx->stream);
  MG_HIP(hipcub::DeviceRadixSort::SortKeys(temp.get(), temp_bytes, db, ni, 0, end_bit, ctx->stream));
  uint64_t* sorted = db.Current();
  uint64_t* other = (sorted == keys.get()) ? alt.get() : keys.get();
  // Unique: flag first occurrences, then select.
  StreamBuf<uint32_t> keep(n, ctx->stream);
  mark_dups_kernel<<<grid_for(n, 256), 256, 0, ctx->stream>>>(sorted, n, keep.get());
  MG_LAUNCH_CHECK();
  StreamBuf<int> n_sel(1, ctx->stream);
  size_t temp2 = 0;
  MG_HIP(hipcub::DeviceSelect::Flagged(nullptr, temp2, sorted, keep.get(), other, n_sel.get(), ni,
                                       ctx->stream));
  StreamBuf<uint8_t> temp_sel(temp2, ctx->stream);
  MG_HIP(hipcub::DeviceSelect::Flagged(temp_sel.get(), temp2, sorted, keep.get(), other, n_sel.get(),
                                       ni, ctx->stream));
  int h = 0;
  MG_HIP(hipMemcpyAsync(&h, n_sel.get(), sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  MG_HIP(hipStreamSynchronize(ctx->stream));
  // Result lives in `other`; make it the caller's buffer.
  if (other == alt.get()) keys = std::move(alt);
  return static_cast<uint64_t>(h);
}

}  // namespace mgamd
