// Measurement helper: the gather floor of the walk's hop kernel.
//
// Every set bit of A^T costs the hop kernel one gather of a 512-byte frontier row (64 words,
// 4096 sources). This kernel does only those gathers, in a given order, with nothing else:
// no tile decode, no LDS accumulation, no write-out. Its time over the hop's row list (the
// source vertex of every A^T bit, in strip order) is the floor that any kernel built on one
// row gather per bit can reach on the device. Rows are read the way the hop kernel reads
// them: lane c of each half-wave loads words c and c + 32 of one row, kProbeUnroll rows per
// half-wave in flight. The loaded words are folded into one word per lane so the loads stay.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "matrixgraph_amd.h"
#include "mg_device.hpp"
#include "mg_internal.hpp"

namespace mgamd {
namespace {

constexpr int kProbeUnroll = 8;

__global__ void __launch_bounds__(256) gather_rows_kernel(const uint32_t* __restrict__ rows_p, uint64_t n,
                                                          uint64_t per_wave, const uint64_t* __restrict__ x_p,
                                                          uint64_t* __restrict__ out) {
  const gptr<const uint32_t> rows = as_global(rows_p);
  const gptr<const uint64_t> x = as_global(x_p);
  const uint32_t lane = lane_id();
  const uint64_t wave = static_cast<uint64_t>(blockIdx.x) * 4 + threadIdx.x / 64;
  const uint64_t b = wave * per_wave;
  const uint64_t e = b + per_wave < n ? b + per_wave : n;
  const uint32_t c = lane & 31, h = lane >> 5;
  uint64_t a0 = 0, a1 = 0;
  for (uint64_t j = b + h; j < e; j += 2 * kProbeUnroll) {
    uint32_t r[kProbeUnroll];
#pragma unroll
    for (int k = 0; k < kProbeUnroll; ++k) {
      const uint64_t jj = j + 2 * k;
      r[k] = jj < e ? rows[jj] : 0u;
    }
    uint64_t v0[kProbeUnroll], v1[kProbeUnroll];
#pragma unroll
    for (int k = 0; k < kProbeUnroll; ++k) {
      const gptr<const uint64_t> row = x + static_cast<uint64_t>(r[k]) * 64 + c;
      v0[k] = row[0];
      v1[k] = row[32];
    }
#pragma unroll
    for (int k = 0; k < kProbeUnroll; ++k) {
      a0 |= v0[k];
      a1 ^= v1[k];
    }
  }
  out[wave * 64 + lane] = a0 ^ a1;
}

}  // namespace
}  // namespace mgamd

using namespace mgamd;

extern "C" int mg_probe_gather_rows(mg_context* ctx, const uint64_t* x, uint64_t n_x_rows, const uint32_t* rows,
                                    uint64_t n_rows, uint32_t grid, uint32_t reps, double* ms_per_pass) {
  return guarded([&] {
    MG_REQUIRE(ctx && x && ms_per_pass && (rows || n_rows == 0), "mg_probe_gather_rows: NULL argument");
    MG_REQUIRE(grid >= 1 && grid <= (1u << 20), "mg_probe_gather_rows: grid must be in [1, 2^20]");
    MG_REQUIRE(reps >= 1, "mg_probe_gather_rows: reps must be >= 1");
    DeviceGuard g(ctx->device);
    // Every row index must address a row of x: checked on the host before any launch.
    std::vector<uint32_t> h(n_rows);
    if (n_rows) MG_HIP(hipMemcpy(h.data(), rows, sizeof(uint32_t) * n_rows, hipMemcpyDeviceToHost));
    for (uint32_t r : h) MG_REQUIRE(r < n_x_rows, "mg_probe_gather_rows: row index out of range");
    DevBuf<uint64_t> out(static_cast<uint64_t>(grid) * 256);
    const uint64_t waves = static_cast<uint64_t>(grid) * 4;
    const uint64_t per_wave = (n_rows + waves - 1) / waves;
    hipStream_t st = ctx->stream;
    hipEvent_t e0, e1;
    MG_HIP(hipEventCreate(&e0));
    MG_HIP(hipEventCreate(&e1));
    gather_rows_kernel<<<grid, 256, 0, st>>>(rows, n_rows, per_wave, x, out.get());
    MG_LAUNCH_CHECK();
    MG_HIP(hipEventRecord(e0, st));
    for (uint32_t i = 0; i < reps; ++i) {
      gather_rows_kernel<<<grid, 256, 0, st>>>(rows, n_rows, per_wave, x, out.get());
      MG_LAUNCH_CHECK();
    }
    MG_HIP(hipEventRecord(e1, st));
    MG_HIP(hipEventSynchronize(e1));
    float ms = 0;
    MG_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_pass = static_cast<double>(ms) / reps;
  });
}
