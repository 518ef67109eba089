// Device-resident bit-tiled matrices: upload, download, digest, popcount and
// the reference's Count / MatrixBitCount entry points.
#include <algorithm>
#include <cstring>
#include <string>

#include "mg_device.hpp"
#include "mg_internal.hpp"

namespace mgamd {

void alloc_tiled_matrix(mg_tiled_matrix* m, uint32_t n_strips, uint32_t n_nz_tile) {
  m->n_strips = n_strips;
  m->n_nz_tile = n_nz_tile;
  m->tile_size = kTile;
  m->offset_row.alloc(static_cast<size_t>(n_strips) + 1);
  m->row_idx.alloc(n_nz_tile);
  m->col_idx.alloc(n_nz_tile);
  m->data.alloc(static_cast<size_t>(n_nz_tile) * kTile);
}

namespace {

// *count += popcount(bits [0, size_bits) of data). Bits past size_bits in the
// last word are masked (the layout's own padding), unlike the reference kernel
// that counts the whole last word (matrix_operations.cu:282-296).
__global__ void __launch_bounds__(256) bit_count_kernel(const uint64_t* __restrict__ data,
                                                        uint64_t size_bits,
                                                        unsigned long long* __restrict__ count) {
  const uint64_t n_words = (size_bits + 63) / 64;
  uint64_t local = 0;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n_words;
       i += stride) {
    uint64_t w = data[i];
    const uint64_t rem = size_bits - i * 64;
    if (rem < 64) w &= (1ull << rem) - 1;
    local += __popcll(w);
  }
  local = wave_sum(local);
  if (lane_id() == 0 && local) atomicAdd(count, static_cast<unsigned long long>(local));
}

// Digest of a tiled matrix: per tile one index term, per non-zero word one data term.
__global__ void __launch_bounds__(256) matrix_digest_kernel(const uint32_t* __restrict__ row_idx,
                                                            const uint32_t* __restrict__ col_idx,
                                                            const uint64_t* __restrict__ data,
                                                            uint32_t n_nz_tile, uint32_t n_strips,
                                                            unsigned long long* __restrict__ out) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t n_waves = gridDim.x * blockDim.x / kWave;
  const uint32_t lane = lane_id();
  uint64_t acc = 0;
  for (uint32_t t = wave; t < n_nz_tile; t += n_waves) {
    const uint64_t key = static_cast<uint64_t>(row_idx[t]) * n_strips + col_idx[t];
    const uint64_t w = data[static_cast<uint64_t>(t) * kTile + lane];
    if (w) acc += digest_term(key * kTile + lane, w);
    if (lane == 0) acc += digest_term(~key, 0x5EEDull);
  }
  acc = wave_sum(acc);
  if (lane == 0) atomicAdd(out, static_cast<unsigned long long>(acc));
}

void validate_host_index(uint32_t n_strips, uint32_t n_nz_tile, const uint32_t* off, const uint32_t* row,
                         const uint32_t* col) {
  MG_REQUIRE(off != nullptr && (n_nz_tile == 0 || (row != nullptr && col != nullptr)),
             "tiled matrix: index arrays must not be NULL");
  MG_REQUIRE(off[0] == 0, "tiled matrix: tile_offset_row[0] != 0");
  for (uint32_t x = 0; x < n_strips; ++x) {
    MG_REQUIRE(off[x + 1] >= off[x], "tiled matrix: tile_offset_row not monotone at row " + std::to_string(x));
    for (uint32_t i = off[x]; i < off[x + 1]; ++i) {
      MG_REQUIRE(i < n_nz_tile, "tiled matrix: tile_offset_row exceeds n_nz_tile");
      MG_REQUIRE(row[i] == x, "tiled matrix: tile_row_idx disagrees with tile_offset_row at tile " +
                                  std::to_string(i));
      MG_REQUIRE(col[i] < n_strips, "tiled matrix: tile_col_idx out of range at tile " + std::to_string(i));
      MG_REQUIRE(i == off[x] || col[i] > col[i - 1],
                 "tiled matrix: tiles of row " + std::to_string(x) + " not in ascending column order");
    }
  }
  MG_REQUIRE(off[n_strips] == n_nz_tile, "tiled matrix: tile_offset_row[n_strips] != n_nz_tile");
}

}  // namespace
}  // namespace mgamd

using namespace mgamd;

extern "C" int mg_tiled_matrix_upload(mg_context* ctx, uint32_t n_strips, uint32_t n_nz_tile,
                                      uint32_t tile_size, const uint32_t* tile_offset_row,
                                      const uint32_t* tile_row_idx, const uint32_t* tile_col_idx,
                                      const uint64_t* data, mg_tiled_matrix** out) {
  return guarded([&] {
    MG_REQUIRE(ctx && out, "mg_tiled_matrix_upload: NULL ctx/out");
    *out = nullptr;
    if (tile_size != kTile)
      fail(MG_ERR_UNSUPPORTED, "mg_tiled_matrix_upload: tile_size " + std::to_string(tile_size) +
                                   " unsupported (only 64)");
    MG_REQUIRE(n_strips > 0, "mg_tiled_matrix_upload: n_strips == 0");
    validate_host_index(n_strips, n_nz_tile, tile_offset_row, tile_row_idx, tile_col_idx);
    DeviceGuard g(ctx->device);
    auto m = std::make_unique<mg_tiled_matrix>();
    m->device = ctx->device;
    alloc_tiled_matrix(m.get(), n_strips, n_nz_tile);
    MG_HIP(hipMemcpyAsync(m->offset_row.get(), tile_offset_row, sizeof(uint32_t) * (n_strips + 1ull),
                          hipMemcpyHostToDevice, ctx->stream));
    if (n_nz_tile) {
      MG_HIP(hipMemcpyAsync(m->row_idx.get(), tile_row_idx, sizeof(uint32_t) * n_nz_tile,
                            hipMemcpyHostToDevice, ctx->stream));
      MG_HIP(hipMemcpyAsync(m->col_idx.get(), tile_col_idx, sizeof(uint32_t) * n_nz_tile,
                            hipMemcpyHostToDevice, ctx->stream));
      if (data)
        MG_HIP(hipMemcpyAsync(m->data.get(), data, sizeof(uint64_t) * kTile * n_nz_tile,
                              hipMemcpyHostToDevice, ctx->stream));
      else
        MG_HIP(hipMemsetAsync(m->data.get(), 0, sizeof(uint64_t) * kTile * n_nz_tile, ctx->stream));
    }
    MG_HIP(hipStreamSynchronize(ctx->stream));
    *out = m.release();
  });
}

extern "C" int mg_tiled_matrix_metadata(const mg_tiled_matrix* m, uint32_t* n_strips, uint32_t* n_nz_tile,
                                        uint32_t* tile_size) {
  return guarded([&] {
    MG_REQUIRE(m != nullptr, "mg_tiled_matrix_metadata: m is NULL");
    if (n_strips) *n_strips = m->n_strips;
    if (n_nz_tile) *n_nz_tile = m->n_nz_tile;
    if (tile_size) *tile_size = m->tile_size;
  });
}

extern "C" int mg_tiled_matrix_download(mg_context* ctx, const mg_tiled_matrix* m, uint32_t* tile_offset_row,
                                        uint32_t* tile_row_idx, uint32_t* tile_col_idx, uint64_t* data,
                                        uint64_t* nz_tile_bm) {
  return guarded([&] {
    MG_REQUIRE(ctx && m, "mg_tiled_matrix_download: NULL ctx/m");
    DeviceGuard g(ctx->device);
    const uint32_t n = m->n_nz_tile;
    if (tile_offset_row)
      MG_HIP(hipMemcpyAsync(tile_offset_row, m->offset_row.get(), sizeof(uint32_t) * (m->n_strips + 1ull),
                            hipMemcpyDeviceToHost, ctx->stream));
    std::vector<uint32_t> rtmp, ctmp;
    uint32_t* rows = tile_row_idx;
    uint32_t* cols = tile_col_idx;
    if (nz_tile_bm && n) {
      if (!rows) { rtmp.resize(n); rows = rtmp.data(); }
      if (!cols) { ctmp.resize(n); cols = ctmp.data(); }
    }
    if (rows && n)
      MG_HIP(hipMemcpyAsync(rows, m->row_idx.get(), sizeof(uint32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
    if (cols && n)
      MG_HIP(hipMemcpyAsync(cols, m->col_idx.get(), sizeof(uint32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
    if (data && n)
      MG_HIP(hipMemcpyAsync(data, m->data.get(), sizeof(uint64_t) * kTile * n, hipMemcpyDeviceToHost,
                            ctx->stream));
    MG_HIP(hipStreamSynchronize(ctx->stream));
    if (nz_tile_bm) {
      const uint64_t bits = static_cast<uint64_t>(m->n_strips) * m->n_strips;
      std::memset(nz_tile_bm, 0, sizeof(uint64_t) * (bits / 64 + 1));
      for (uint32_t i = 0; i < n; ++i) {
        const uint64_t b = static_cast<uint64_t>(rows[i]) * m->n_strips + cols[i];
        nz_tile_bm[b >> 6] |= 1ull << (b & 63);
      }
    }
  });
}

extern "C" int mg_tiled_matrix_device_arrays(const mg_tiled_matrix* m, const uint32_t** tile_offset_row,
                                             const uint32_t** tile_row_idx, const uint32_t** tile_col_idx,
                                             const uint64_t** data) {
  return guarded([&] {
    MG_REQUIRE(m != nullptr, "mg_tiled_matrix_device_arrays: m is NULL");
    if (tile_offset_row) *tile_offset_row = m->offset_row.get();
    if (tile_row_idx) *tile_row_idx = m->row_idx.get();
    if (tile_col_idx) *tile_col_idx = m->col_idx.get();
    if (data) *data = m->data.get();
  });
}

extern "C" int mg_tiled_matrix_digest(mg_context* ctx, const mg_tiled_matrix* m, uint64_t* digest) {
  return guarded([&] {
    MG_REQUIRE(ctx && m && digest, "mg_tiled_matrix_digest: NULL argument");
    DeviceGuard g(ctx->device);
    unsigned long long* d = ctx->counters.get();
    MG_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), ctx->stream));
    if (m->n_nz_tile) {
      matrix_digest_kernel<<<grid_for(m->n_nz_tile, 4, 16384), 256, 0, ctx->stream>>>(
          m->row_idx.get(), m->col_idx.get(), m->data.get(), m->n_nz_tile, m->n_strips, d);
      MG_LAUNCH_CHECK();
    }
    unsigned long long h = 0;
    MG_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    MG_HIP(hipStreamSynchronize(ctx->stream));
    *digest = h;
  });
}

extern "C" int mg_tiled_matrix_popcount(mg_context* ctx, const mg_tiled_matrix* m, uint64_t* count) {
  return guarded([&] {
    MG_REQUIRE(ctx && m && count, "mg_tiled_matrix_popcount: NULL argument");
    DeviceGuard g(ctx->device);
    *count = device_popcount(ctx, m->data.get(), static_cast<uint64_t>(m->n_nz_tile) * kTile);
  });
}

extern "C" int mg_tiled_matrix_destroy(mg_tiled_matrix* m) {
  return guarded([&] {
    if (!m) return;
    DeviceGuard g(m->device);
    delete m;
  });
}

extern "C" int mg_matrix_bit_count(void* stream, const uint64_t* matrix, uint64_t* count, uint64_t size) {
  return guarded([&] {
    MG_REQUIRE(count != nullptr && (matrix != nullptr || size == 0), "mg_matrix_bit_count: NULL buffer");
    if (size == 0) return;
    const uint64_t words = (size + 63) / 64;
    bit_count_kernel<<<grid_for(words, 256 * 8, 8192), 256, 0, static_cast<hipStream_t>(stream)>>>(
        matrix, size, reinterpret_cast<unsigned long long*>(count));
    MG_LAUNCH_CHECK();
  });
}

extern "C" int mg_count(void* stream, size_t tile_size, size_t n_strips, size_t n_nz_tile, const uint64_t* data,
                        uint64_t* count) {
  (void)n_strips;
  return guarded([&] {
    if (tile_size != kTile) fail(MG_ERR_UNSUPPORTED, "mg_count: tile_size must be 64");
    MG_REQUIRE(count != nullptr && (data != nullptr || n_nz_tile == 0), "mg_count: NULL buffer");
    if (n_nz_tile == 0) return;
    const uint64_t bits = static_cast<uint64_t>(n_nz_tile) * kTile * 64;
    bit_count_kernel<<<grid_for(bits / 64, 256 * 8, 8192), 256, 0, static_cast<hipStream_t>(stream)>>>(
        data, bits, reinterpret_cast<unsigned long long*>(count));
    MG_LAUNCH_CHECK();
  });
}
