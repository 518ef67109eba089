// General bit-tiled Boolean product C = A * B with B supplied transposed (Bt):
//   - layout-level entry points of the reference wrapper (MatrixBitAnd,
//     InitBitTiledMatrixMetadataByLayoutMatrix, FillTiles);
//   - the handle-based symbolic pass (structural tile set of C, tile-product count)
//     and numeric pass.
// Tile product on a wave64: lane r holds L[r] (row r of the left tile), lane c holds
// Rt[c] (row c of the transposed right tile); output row r bit c = (L[r] & Rt[c]) != 0.
// The loop runs only over non-zero Rt rows (wave ballot), so sparse tiles cost a few
// instructions; dense tiles (64 non-zero rows) take 64 iterations. The handle-based numeric
// pass multiplies pairs of tiles with at most 10 bits from their entry lists instead, one
// pair per lane (fill_tiles_payload_kernel), and keeps the wave product for denser tiles.
#include <mutex>
#include <string>

#include <hipcub/hipcub.hpp>

#include "mg_device.hpp"
#include "mg_internal.hpp"

namespace mgamd {
namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
#ifndef MG_PAIR_BATCH
#define MG_PAIR_BATCH 4
#endif
constexpr int kPairBatch = MG_PAIR_BATCH;  // tile pairs whose loads fill_tiles_kernel issues together
constexpr uint32_t kRowFeatures = MG_ROW_FEATURES;  // features per C row (mg_multiply_row_features)

__device__ inline uint64_t tile_product(uint64_t l_row, uint64_t rt_row) {
  uint64_t nz = ballot(rt_row != 0);
  if (ballot(l_row != 0) == 0) return 0;
  uint64_t out = 0;
  while (nz) {
    const int c = __ffsll(static_cast<unsigned long long>(nz)) - 1;
    nz &= nz - 1;
    const uint64_t w = read_lane(rt_row, c);
    out |= static_cast<uint64_t>((l_row & w) != 0) << c;
  }
  return out;
}

// Dense tile pairs on the matrix cores, as one contraction per C tile: for every pair
// (A(x,k), Bt(y,k)) whose two tiles both hold at least min_bits set bits, the vector kernel
// records the pair (its two tile indices) instead of multiplying it; then
// dense_contract_kernel sums, per C tile, the int8 0/1 products of all its recorded pairs in
// int32 (v_mfma_i32_32x32x32_i8, four 32 x 32 output blocks, two K-steps of 32 per pair) and
// turns counts > 0 into bit rows once, ORed into C. Operands are expanded from the bit tiles
// in registers: a lane loads one 64-bit row of each tile (A row r, Bt row c = B column c) and
// turns 4 bits at a time into 4 bytes of 0/1 with a 24-bit multiply.
// Lane layout (32x32x32 i8): lane l holds 16 bytes of A row l % 32 and of B column l % 32,
// at k = 16 * (l / 32) .. +15 of the K-step (any k order works if both operands use it);
// result register v of lane l is C[8 * (v / 4) + 4 * (l / 32) + v % 4][l % 32].
typedef int mg_v16i __attribute__((ext_vector_type(16)));
typedef int mg_v4i __attribute__((ext_vector_type(4)));

// Bits 0..15 of x as 16 bytes of value 0 or 1 (byte j = bit j): nibble * 0x204081 puts bit
// i of the nibble at bit 8i (the four shifted copies do not overlap), the mask keeps those.
__device__ inline mg_v4i bits16_to_bytes(uint32_t x) {
  mg_v4i out;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    out[q] = static_cast<int>(__umul24((x >> (4 * q)) & 0xFu, 0x204081u) & 0x01010101u);
  return out;
}

// Recorded dense pairs of one C tile (from one 64-tile step of its index intersection).
struct DenseTask {
  uint32_t tile, base, n;
};
struct DenseRecord {
  uint2* pairs;  // (A tile, Bt tile)
  DenseTask* tasks;
  unsigned long long* counter;  // pairs recorded (low 32 bits) and tasks (high 32 bits)
  uint32_t cap;            // capacity of pairs and of tasks (the exact dense pair count)
  const uint32_t* bits_a;  // set bits per tile
  const uint32_t* bits_bt;
  uint32_t min_bits;
};

// One wave per recorded task: int32 counts of the 64 x 64 output over all the task's pairs
// on the matrix cores, then one threshold; the rows are ORed into C atomically (several
// tasks may share a C tile).
__global__ void __launch_bounds__(kBlock) dense_contract_kernel(const DenseTask* __restrict__ tasks,
                                                                const unsigned long long* __restrict__ counter,
                                                                uint32_t cap, const uint2* __restrict__ pairs,
                                                                const uint64_t* __restrict__ data_a,
                                                                const uint64_t* __restrict__ data_bt,
                                                                uint64_t* __restrict__ data_c) {
  const uint32_t lane = lane_id();
  const uint32_t lo = lane & 31, hi = lane >> 5;
  const uint32_t n_tasks = min(static_cast<uint32_t>(*counter >> 32), cap);
  const uint32_t n_waves = gridDim.x * kWavesPerBlock;
  for (uint32_t t = read_first_lane(blockIdx.x * kWavesPerBlock + threadIdx.x / kWave); t < n_tasks; t += n_waves) {
    const DenseTask T = tasks[t];
    mg_v16i acc00 = {}, acc01 = {}, acc10 = {}, acc11 = {};
    const uint32_t p_end = min(T.base + T.n, cap);
    for (uint32_t p = T.base; p < p_end; ++p) {
      const uint2 pr = pairs[p];
      const uint64_t* A = data_a + static_cast<uint64_t>(pr.x) * kTile;
      const uint64_t* B = data_bt + static_cast<uint64_t>(pr.y) * kTile;
      const uint64_t ra0 = A[lo], ra1 = A[32 + lo], rb0 = B[lo], rb1 = B[32 + lo];
#pragma unroll
      for (uint32_t st = 0; st < 2; ++st) {
        const uint32_t k = 32 * st + 16 * hi;
        const mg_v4i a0 = bits16_to_bytes(static_cast<uint32_t>(ra0 >> k));
        const mg_v4i a1 = bits16_to_bytes(static_cast<uint32_t>(ra1 >> k));
        const mg_v4i b0 = bits16_to_bytes(static_cast<uint32_t>(rb0 >> k));
        const mg_v4i b1 = bits16_to_bytes(static_cast<uint32_t>(rb1 >> k));
        acc00 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, b0, acc00, 0, 0, 0);
        acc01 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, b1, acc01, 0, 0, 0);
        acc10 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, b0, acc10, 0, 0, 0);
        acc11 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, b1, acc11, 0, 0, 0);
      }
    }
    uint64_t out = 0;  // output row `lane`
#pragma unroll
    for (uint32_t v = 0; v < 16; ++v) {
      const uint32_t i0 = 8 * (v / 4) + (v % 4);
      const uint64_t b00 = ballot(acc00[v] > 0), b01 = ballot(acc01[v] > 0);
      const uint64_t b10 = ballot(acc10[v] > 0), b11 = ballot(acc11[v] > 0);
      if (lane == i0) out |= (b00 & 0xFFFFFFFFull) | ((b01 & 0xFFFFFFFFull) << 32);
      if (lane == i0 + 4) out |= (b00 >> 32) | ((b01 >> 32) << 32);
      if (lane == 32 + i0) out |= (b10 & 0xFFFFFFFFull) | ((b11 & 0xFFFFFFFFull) << 32);
      if (lane == 32 + i0 + 4) out |= (b10 >> 32) | ((b11 >> 32) << 32);
    }
    if (out) atomicOr(reinterpret_cast<unsigned long long*>(data_c) + static_cast<uint64_t>(T.tile) * kTile + lane,
                      static_cast<unsigned long long>(out));
  }
}

// Per Bt column k: the dense Bt tiles in it; then the dense pairs of the product: sum over
// dense A tiles (x, k) of that count (the structural pair count restricted to dense tiles).
__global__ void __launch_bounds__(kBlock) dense_col_count_kernel(const uint32_t* __restrict__ col, uint32_t n,
                                                                 const uint32_t* __restrict__ bits,
                                                                 uint32_t min_bits, uint32_t* __restrict__ count) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (bits[i] >= min_bits) atomicAdd(count + col[i], 1u);
}
__global__ void __launch_bounds__(kBlock) dense_pair_count_kernel(const uint32_t* __restrict__ col, uint32_t n,
                                                                  const uint32_t* __restrict__ bits,
                                                                  uint32_t min_bits,
                                                                  const uint32_t* __restrict__ count,
                                                                  unsigned long long* __restrict__ out) {
  uint64_t local = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (bits[i] >= min_bits) local += count[col[i]];
  local = wave_sum(local);
  if (lane_id() == 0 && local) atomicAdd(out, static_cast<unsigned long long>(local));
}

// Set bits of every tile (for the dense-pair dispatch).
__global__ void __launch_bounds__(kBlock) tile_bits_kernel(const uint64_t* __restrict__ data, uint32_t n_tiles,
                                                           uint32_t* __restrict__ bits) {
  const uint32_t n_waves = gridDim.x * kWavesPerBlock;
  for (uint32_t i = read_first_lane(blockIdx.x * kWavesPerBlock + threadIdx.x / kWave); i < n_tiles; i += n_waves) {
    const uint64_t c = wave_sum(__popcll(data[static_cast<uint64_t>(i) * kTile + lane_id()]));
    if (lane_id() == 0) bits[i] = static_cast<uint32_t>(c);
  }
}

// Index of key in sorted a[lo, hi), or -1.
__device__ inline int64_t find_sorted(const uint32_t* __restrict__ a, uint32_t lo, uint32_t hi, uint32_t key) {
  uint32_t l = lo, h = hi;
  while (l < h) {
    const uint32_t mid = l + ((h - l) >> 1);
    if (a[mid] < key) l = mid + 1; else h = mid;
  }
  return (l < hi && a[l] == key) ? static_cast<int64_t>(l) : -1;
}

// One wave per tile of C: intersect A.row(x) with Bt.row(y) on the tile column,
// OR the tile products, OR the result into data_c (FillTiles semantics,
// core/task/gpu_task/kernel/matrix_operations.cu:318-366 with the intended
// offsets of B at :356 and the intended 64x64 product of :212-233). Whole-tile products on
// the vector ALUs (the layer-1 entry point, whose operands are raw tile arrays).
__global__ void __launch_bounds__(kBlock) fill_tiles_kernel(
    uint32_t n_c, const uint32_t* __restrict__ off_a, const uint32_t* __restrict__ off_bt,
    const uint32_t* __restrict__ row_c, const uint32_t* __restrict__ col_c, const uint32_t* __restrict__ col_a,
    const uint32_t* __restrict__ col_bt, const uint64_t* __restrict__ data_a, const uint64_t* __restrict__ data_bt,
    uint64_t* __restrict__ data_c) {
  const uint32_t lane = lane_id();
  const uint32_t n_waves = gridDim.x * kWavesPerBlock;
  for (uint32_t i = read_first_lane(blockIdx.x * kWavesPerBlock + threadIdx.x / kWave); i < n_c; i += n_waves) {
    const uint32_t x = row_c[i];
    const uint32_t y = col_c[i];
    const uint32_t a0 = off_a[x], a1 = off_a[x + 1];
    const uint32_t b0 = off_bt[y], b1 = off_bt[y + 1];
    const bool a_short = (a1 - a0) <= (b1 - b0);
    const uint32_t s0 = a_short ? a0 : b0, s1 = a_short ? a1 : b1;
    const uint32_t l0 = a_short ? b0 : a0, l1 = a_short ? b1 : a1;
    const uint32_t* s_col = a_short ? col_a : col_bt;
    const uint32_t* l_col = a_short ? col_bt : col_a;
    uint64_t acc = 0;
    for (uint32_t base = s0; base < s1; base += kWave) {
      const uint32_t j = base + lane;
      int64_t pos = -1;
      if (j < s1) pos = find_sorted(l_col, l0, l1, s_col[j]);
      uint64_t hit = ballot(pos >= 0);
      // kPairBatch matching pairs at a time, all their tile loads issued before the first
      // product, so a wave waits once per batch instead of once per pair.
      while (hit) {
        uint64_t l_row[kPairBatch], rt_row[kPairBatch];
        uint32_t n_pairs = 0;  // wave-uniform
#pragma unroll
        for (int q = 0; q < kPairBatch; ++q) {
          l_row[q] = 0;
          rt_row[q] = 0;
          if (hit) {
            const int k = __ffsll(static_cast<unsigned long long>(hit)) - 1;
            hit &= hit - 1;
            const uint32_t js = base + k;
            const uint32_t jl = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(pos), k));
            const uint64_t ia = a_short ? js : jl;
            const uint64_t ib = a_short ? jl : js;
            l_row[q] = data_a[ia * kTile + lane];
            rt_row[q] = data_bt[ib * kTile + lane];
            ++n_pairs;
          }
        }
#pragma unroll
        for (int q = 0; q < kPairBatch; ++q)
          if (static_cast<uint32_t>(q) < n_pairs) acc |= tile_product(l_row[q], rt_row[q]);
      }
    }
    uint64_t* dst = data_c + static_cast<uint64_t>(i) * kTile + lane;
    if (acc) *dst |= acc;
  }
}

// Product of two dense tiles through their common inner indices (`common`: the columns in use
// in both, wave-uniform): for each such j, the rows r with L(r, j) set receive the columns c
// with Rt(c, j) set (a ballot). The loop runs over the j the two tiles share, where
// tile_product runs over every non-zero row of Rt; on R-MAT tiles above kWideMax bits that
// is about half the steps.
__device__ inline uint64_t tile_product_common(uint64_t l_row, uint64_t rt_row, uint64_t common) {
  uint64_t out = 0;
  for (; common; common &= common - 1) {
    const int j = __ffsll(static_cast<unsigned long long>(common)) - 1;
    const uint64_t cols = ballot((rt_row >> j) & 1);
    if ((l_row >> j) & 1) out |= cols;
  }
  return out;
}

// 16-byte compact payload of a tile with at most kWideMax set bits: entry i = (row << 6) |
// column in bits 12(i % 5).. of word i / 5, the entry count in bits 60..63 of word 1; count
// kWideDense marks a denser tile, whose word 0 then holds the OR of its rows (the columns in
// use: the inner indices j through which it meets the other operand's tiles).
typedef uint64_t WidePayload __attribute__((ext_vector_type(2)));
constexpr uint32_t kWideMax = 10;
constexpr uint32_t kWideDense = 15;

__device__ inline uint32_t wide_count(WidePayload p) { return static_cast<uint32_t>(p.y >> 60); }
__device__ inline uint32_t wide_entry(WidePayload p, uint32_t i) {
  return static_cast<uint32_t>((i < 5 ? p.x : p.y) >> (12 * (i < 5 ? i : i - 5))) & 0xFFFu;
}

// Wide payload of every tile of both operands, one wave per tile (lane r holds row r): tiles
// [0, n_a) of data_a, then the n_bt tiles of data_bt, at the same positions of `payload`; with
// `bits`, also each tile's set-bit count at the same position.
__global__ void __launch_bounds__(kBlock) payload_kernel(const uint64_t* __restrict__ data_a, uint32_t n_a,
                                                         const uint64_t* __restrict__ data_bt, uint32_t n_bt,
                                                         WidePayload* __restrict__ payload,
                                                         uint32_t* __restrict__ bits = nullptr) {
  const uint32_t n_waves = gridDim.x * kWavesPerBlock;
  const uint64_t n_tiles = static_cast<uint64_t>(n_a) + n_bt;
  for (uint64_t i = read_first_lane(blockIdx.x * kWavesPerBlock + threadIdx.x / kWave); i < n_tiles; i += n_waves) {
    const uint64_t word = i < n_a ? data_a[i * kTile + lane_id()] : data_bt[(i - n_a) * kTile + lane_id()];
    const uint64_t total = wave_sum(__popcll(word));
    WidePayload pl = {0, static_cast<uint64_t>(kWideDense) << 60};
    const uint64_t cols = wave_or(word);
    if (total > kWideMax) pl.x = cols;
    if (total <= kWideMax) {
      uint64_t w[2] = {0, 0};
      uint32_t n = 0;
      for (uint64_t m = ballot(word != 0); m; m &= m - 1) {
        const int r = __ffsll(static_cast<unsigned long long>(m)) - 1;
        for (uint64_t wr = read_lane(word, r); wr; wr &= wr - 1) {
          const uint64_t c = static_cast<uint64_t>(__ffsll(static_cast<unsigned long long>(wr)) - 1);
          w[n / 5] |= ((static_cast<uint64_t>(r) << 6) | c) << (12 * (n % 5));
          ++n;
        }
      }
      pl.x = w[0];
      pl.y = w[1] | (static_cast<uint64_t>(n) << 60);
    }
    if (lane_id() == 0) {
      payload[i] = pl;
      if (bits) bits[i] = static_cast<uint32_t>(total);
    }
  }
}

bool prepared(const mg_tiled_matrix* m) { return m->prep_version == m->version && m->prep_payload.get(); }

// Payloads of A's tiles at pl[0, n_a) and of Bt's at pl[n_a, n_a + n_bt), one launch (with
// `bits`, the tiles' set-bit counts at the same positions). An operand prepared by
// mg_tiled_matrix_prepare_operand is copied from its kept payloads instead (16 + 4 bytes per
// tile rather than a read of its 512-byte tiles).
void launch_payloads(hipStream_t st, const mg_tiled_matrix* a, const mg_tiled_matrix* bt, WidePayload* pl,
                     uint32_t* bits = nullptr) {
  const bool pa = prepared(a) && a->n_nz_tile, pb = prepared(bt) && bt->n_nz_tile;
  const uint32_t na = a->n_nz_tile, nb = bt->n_nz_tile;
  if (pa) {
    MG_HIP(hipMemcpyAsync(pl, a->prep_payload.get(), sizeof(WidePayload) * na, hipMemcpyDeviceToDevice, st));
    if (bits) MG_HIP(hipMemcpyAsync(bits, a->prep_bits.get(), sizeof(uint32_t) * na, hipMemcpyDeviceToDevice, st));
  }
  if (pb) {
    MG_HIP(hipMemcpyAsync(pl + na, bt->prep_payload.get(), sizeof(WidePayload) * nb, hipMemcpyDeviceToDevice, st));
    if (bits)
      MG_HIP(hipMemcpyAsync(bits + na, bt->prep_bits.get(), sizeof(uint32_t) * nb, hipMemcpyDeviceToDevice, st));
  }
  // The rest in one launch: both operands, or the one that is not prepared.
  const uint64_t n = (pa ? 0ull : na) + (pb ? 0ull : nb);
  if (!n) return;
  if (!pa && !pb)
    payload_kernel<<<grid_for(n, kWavesPerBlock, 65536), kBlock, 0, st>>>(a->data.get(), na, bt->data.get(), nb, pl,
                                                                          bits);
  else if (!pa)
    payload_kernel<<<grid_for(n, kWavesPerBlock, 65536), kBlock, 0, st>>>(a->data.get(), na, nullptr, 0, pl, bits);
  else
    payload_kernel<<<grid_for(n, kWavesPerBlock, 65536), kBlock, 0, st>>>(bt->data.get(), nb, nullptr, 0, pl + na,
                                                                          bits ? bits + na : nullptr);
  MG_LAUNCH_CHECK();
}

// fill_tiles_kernel for operands whose tiles are mostly sparse, from their wide payloads
// (pl_a, pl_bt). The intersection gives each lane its own tile pair; a pair of two compact
// tiles is multiplied by that lane alone: every entry (r, j) of the A tile meets every entry
// (c, j) of the Bt tile with the same j and sets bit c of row r in the wave's LDS
// accumulator. Pairs with a denser tile take the full-tile wave product, kPairBatch at a time.
// kMfma: a pair whose two tiles both hold at least rec.min_bits set bits is recorded for
// dense_contract_kernel instead.
// kOut: where a finished C tile goes (PrunedOut): ORed into data_c at its structural index
// (kOutStruct), or, for a pruned product, only if it is non-zero: appended at the next slot
// of out.data with its structural index in out.key (kOutAppend, slots past out.capacity are
// counted but not written), flagged in out.flag (kOutFlag, no data written), or written at
// out.pos[i] (kOutPos).
enum : int { kOutStruct = 0, kOutAppend = 1, kOutFlag = 2, kOutPos = 3 };
struct PrunedOut {
  uint64_t* data;
  uint32_t* key;
  uint32_t* flag;
  const uint32_t* pos;
  unsigned long long* counter;
  uint64_t capacity;
};
template <bool kMfma, int kOut = kOutStruct>
__global__ void __launch_bounds__(kBlock) fill_tiles_payload_kernel(
    uint32_t n_c, const uint32_t* __restrict__ off_a, const uint32_t* __restrict__ off_bt,
    const uint32_t* __restrict__ row_c, const uint32_t* __restrict__ col_c, const uint32_t* __restrict__ col_a,
    const uint32_t* __restrict__ col_bt, const uint64_t* __restrict__ data_a, const uint64_t* __restrict__ data_bt,
    const WidePayload* __restrict__ pl_a, const WidePayload* __restrict__ pl_bt, uint64_t* __restrict__ data_c,
    DenseRecord rec, PrunedOut out = PrunedOut{}) {
  __shared__ unsigned long long s_acc[kWavesPerBlock][kTile];
  const uint32_t lane = lane_id();
  unsigned long long* const acc_l = s_acc[threadIdx.x / kWave];
  acc_l[lane] = 0;
  wave_lds_sync();
  const uint32_t n_waves = gridDim.x * kWavesPerBlock;
  for (uint32_t i = read_first_lane(blockIdx.x * kWavesPerBlock + threadIdx.x / kWave); i < n_c; i += n_waves) {
    const uint32_t x = row_c[i];
    const uint32_t y = col_c[i];
    const uint32_t a0 = off_a[x], a1 = off_a[x + 1];
    const uint32_t b0 = off_bt[y], b1 = off_bt[y + 1];
    const bool a_short = (a1 - a0) <= (b1 - b0);
    const uint32_t s0 = a_short ? a0 : b0, s1 = a_short ? a1 : b1;
    const uint32_t l0 = a_short ? b0 : a0, l1 = a_short ? b1 : a1;
    const uint32_t* s_col = a_short ? col_a : col_bt;
    const uint32_t* l_col = a_short ? col_bt : col_a;
    uint64_t acc = 0;
    for (uint32_t base = s0; base < s1; base += kWave) {
      const uint32_t j = base + lane;
      int64_t pos = -1;
      if (j < s1) pos = find_sorted(l_col, l0, l1, s_col[j]);
      const bool has = pos >= 0;
      const uint32_t ia = has ? (a_short ? j : static_cast<uint32_t>(pos)) : 0u;
      const uint32_t ib = has ? (a_short ? static_cast<uint32_t>(pos) : j) : 0u;
      bool mfma = false;
      if constexpr (kMfma) mfma = has && rec.bits_a[ia] >= rec.min_bits && rec.bits_bt[ib] >= rec.min_bits;
      // Payloads only for the pairs the vector ALUs multiply.
      WidePayload pa = {0, 0}, pb = {0, 0};
      if (has && !mfma) {
        pa = pl_a[ia];
        pb = pl_bt[ib];
      }
      const uint32_t na = wide_count(pa), nb = wide_count(pb);
      const bool dense = has && !mfma && (na == kWideDense || nb == kWideDense);
      if (has && !mfma && !dense) {
        // Only the inner indices j present in both tiles give bits: per common j, the Bt
        // entries' columns are collected into one mask and ORed into the row of every A
        // entry with that j (na + nb steps per common j instead of na * nb comparisons).
        uint64_t jm_a = 0, jm_b = 0;
        for (uint32_t u = 0; u < na; ++u) jm_a |= 1ull << (wide_entry(pa, u) & 63u);
        for (uint32_t v = 0; v < nb; ++v) jm_b |= 1ull << (wide_entry(pb, v) & 63u);
        for (uint64_t common = jm_a & jm_b; common; common &= common - 1) {
          const uint32_t jj = static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(common)) - 1);
          uint64_t cols = 0;
          for (uint32_t v = 0; v < nb; ++v) {
            const uint32_t eb = wide_entry(pb, v);
            if ((eb & 63u) == jj) cols |= 1ull << (eb >> 6);
          }
          for (uint32_t u = 0; u < na; ++u) {
            const uint32_t ea = wide_entry(pa, u);
            if ((ea & 63u) == jj) atomicOr(acc_l + (ea >> 6), cols);
          }
        }
      }
      if constexpr (kMfma) {
        // Record this step's dense pairs as one task of C tile i (dense_contract_kernel).
        const uint64_t mm = ballot(mfma);
        if (mm) {
          const uint32_t n = static_cast<uint32_t>(__popcll(mm));
          // One atomic reserves both the pair positions and the task slot.
          unsigned long long both = 0;
          if (lane == 0) both = atomicAdd(rec.counter, (1ull << 32) | n);
          both = read_first_lane64(both);
          const uint32_t base = static_cast<uint32_t>(both);
          const uint32_t task = static_cast<uint32_t>(both >> 32);
          if (mfma) {
            const uint32_t at = base + static_cast<uint32_t>(__popcll(mm & ((1ull << lane) - 1)));
            if (at < rec.cap) rec.pairs[at] = make_uint2(ia, ib);
          }
          if (lane == 0 && task < rec.cap) rec.tasks[task] = DenseTask{i, base, n};
        }
      }
      // Pairs with one dense tile: the dense tile's rows (lane = row) against the other tile's
      // entries, kPairBatch pairs' loads at a time. An A entry (r, j) ORs column j of the Bt
      // tile (a ballot over its rows) into row r; a Bt entry (c, j) sets bit c of every row r
      // whose A row has bit j.
      const bool a_dense = na == kWideDense, b_dense = nb == kWideDense;
      for (uint64_t dm = ballot(dense && !a_dense); dm;) {
        uint64_t rt_row[kPairBatch], used[kPairBatch];
        WidePayload ent[kPairBatch];
        uint32_t n_pairs = 0;  // wave-uniform
#pragma unroll
        for (int q = 0; q < kPairBatch; ++q) {
          rt_row[q] = 0;
          used[q] = 0;
          ent[q] = WidePayload{0, 0};
          if (dm) {
            const int k = __ffsll(static_cast<unsigned long long>(dm)) - 1;
            dm &= dm - 1;
            const uint64_t kb = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(ib), k));
            rt_row[q] = data_bt[kb * kTile + lane];
            ent[q].x = read_lane(pa.x, k);
            ent[q].y = read_lane(pa.y, k);
            used[q] = read_lane(pb.x, k);  // columns in use of the dense Bt tile
            ++n_pairs;
          }
        }
#pragma unroll
        for (int q = 0; q < kPairBatch; ++q) {
          if (static_cast<uint32_t>(q) >= n_pairs) continue;
          const uint32_t n = wide_count(ent[q]);
          for (uint32_t u = 0; u < n; ++u) {
            const uint32_t e = wide_entry(ent[q], u);
            if (!((used[q] >> (e & 63u)) & 1)) continue;  // no Bt row has bit j
            const uint64_t col = ballot((rt_row[q] >> (e & 63u)) & 1);
            if (lane == (e >> 6)) acc |= col;
          }
        }
      }
      for (uint64_t dm = ballot(dense && a_dense && !b_dense); dm;) {
        uint64_t l_row[kPairBatch], used[kPairBatch];
        WidePayload ent[kPairBatch];
        uint32_t n_pairs = 0;  // wave-uniform
#pragma unroll
        for (int q = 0; q < kPairBatch; ++q) {
          l_row[q] = 0;
          used[q] = 0;
          ent[q] = WidePayload{0, 0};
          if (dm) {
            const int k = __ffsll(static_cast<unsigned long long>(dm)) - 1;
            dm &= dm - 1;
            const uint64_t ka = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(ia), k));
            l_row[q] = data_a[ka * kTile + lane];
            ent[q].x = read_lane(pb.x, k);
            ent[q].y = read_lane(pb.y, k);
            used[q] = read_lane(pa.x, k);  // columns in use of the dense A tile
            ++n_pairs;
          }
        }
#pragma unroll
        for (int q = 0; q < kPairBatch; ++q) {
          if (static_cast<uint32_t>(q) >= n_pairs) continue;
          const uint32_t n = wide_count(ent[q]);
          for (uint32_t v = 0; v < n; ++v) {
            const uint32_t e = wide_entry(ent[q], v);
            if (!((used[q] >> (e & 63u)) & 1)) continue;  // no A row has bit j
            acc |= ((l_row[q] >> (e & 63u)) & 1ull) << (e >> 6);
          }
        }
      }
      // Both tiles dense: the wave product through the inner indices both tiles use.
      for (uint64_t dm = ballot(dense && a_dense && b_dense); dm;) {
        uint64_t l_row[kPairBatch], rt_row[kPairBatch], common[kPairBatch];
        uint32_t n_pairs = 0;  // wave-uniform
#pragma unroll
        for (int q = 0; q < kPairBatch; ++q) {
          l_row[q] = 0;
          rt_row[q] = 0;
          common[q] = 0;
          if (dm) {
            const int k = __ffsll(static_cast<unsigned long long>(dm)) - 1;
            dm &= dm - 1;
            const uint64_t ka = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(ia), k));
            const uint64_t kb = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(ib), k));
            l_row[q] = data_a[ka * kTile + lane];
            rt_row[q] = data_bt[kb * kTile + lane];
            common[q] = read_lane(pa.x, k) & read_lane(pb.x, k);
            ++n_pairs;
          }
        }
#pragma unroll
        for (int q = 0; q < kPairBatch; ++q)
          if (static_cast<uint32_t>(q) < n_pairs) acc |= tile_product_common(l_row[q], rt_row[q], common[q]);
      }
    }
    wave_lds_sync();
    acc |= acc_l[lane];
    acc_l[lane] = 0;
    wave_lds_sync();
    if constexpr (kOut == kOutStruct) {
      uint64_t* dst = data_c + static_cast<uint64_t>(i) * kTile + lane;
      if (acc) *dst |= acc;
    } else {
      const bool nz = ballot(acc != 0) != 0;  // wave-uniform
      if constexpr (kOut == kOutFlag) {
        if (lane == 0) out.flag[i] = nz ? 1u : 0u;
      } else if constexpr (kOut == kOutPos) {
        if (nz) out.data[static_cast<uint64_t>(out.pos[i]) * kTile + lane] = acc;
      } else {
        if (nz) {
          unsigned long long slot = 0;
          if (lane == 0) slot = atomicAdd(out.counter, 1ull);
          slot = read_first_lane64(slot);
          if (slot < out.capacity) {
            out.data[slot * kTile + lane] = acc;
            if (lane == 0) out.key[slot] = i;
          }
        }
      }
    }
  }
}

// len (<= 64) bits of a bit-packed array starting at bit `start`.
__device__ inline uint64_t bit_window(const uint64_t* __restrict__ p, uint64_t start, uint32_t len) {
  const uint64_t w = start >> 6;
  const uint32_t s = static_cast<uint32_t>(start & 63);
  uint64_t v = p[w] >> s;
  if (s != 0 && s + len > 64) v |= p[w + 1] << (64 - s);
  if (len < 64) v &= (1ull << len) - 1;
  return v;
}

// MatrixBitAnd (matrix_operations.cu:256-280), intended semantics: wave per
// (row i of a, 64 rows of b); c bit i*n+j |= any(a row i & b row j).
__global__ void __launch_bounds__(kBlock) bit_and_kernel(const uint64_t* __restrict__ a,
                                                         const uint64_t* __restrict__ b,
                                                         unsigned long long* __restrict__ c, uint32_t m,
                                                         uint32_t k, uint32_t n) {
  const uint32_t lane = lane_id();
  const uint32_t nj = (n + 63) / 64;
  const uint64_t n_tasks = static_cast<uint64_t>(m) * nj;
  const uint64_t n_waves = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
  for (uint64_t task = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; task < n_tasks; task += n_waves) {
    const uint32_t i = static_cast<uint32_t>(task / nj);
    const uint32_t jc = static_cast<uint32_t>(task % nj);
    const uint32_t j = jc * 64 + lane;
    bool hit = false;
    if (j < n) {
      for (uint32_t t0 = 0; t0 < k && !hit; t0 += 64) {
        const uint32_t len = (k - t0) < 64 ? (k - t0) : 64;
        const uint64_t wa = bit_window(a, static_cast<uint64_t>(i) * k + t0, len);
        const uint64_t wb = bit_window(b, static_cast<uint64_t>(j) * k + t0, len);
        hit = (wa & wb) != 0;
      }
    }
    const uint64_t mask = ballot(hit);
    if (lane == 0 && mask) {
      const uint64_t start = static_cast<uint64_t>(i) * n + static_cast<uint64_t>(jc) * 64;
      const uint32_t s = static_cast<uint32_t>(start & 63);
      atomicOr(c + (start >> 6), static_cast<unsigned long long>(mask << s));
      if (s != 0 && (mask >> (64 - s)) != 0)
        atomicOr(c + (start >> 6) + 1, static_cast<unsigned long long>(mask >> (64 - s)));
    }
  }
}

// Layout row counts: cnt[x] = set bits in row x of an n x n layout.
__global__ void __launch_bounds__(kBlock) layout_row_count_kernel(const uint64_t* __restrict__ layout,
                                                                  uint32_t n, uint32_t* __restrict__ cnt) {
  const uint32_t lane = lane_id();
  const uint32_t n_waves = gridDim.x * kWavesPerBlock;
  for (uint32_t x = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; x < n; x += n_waves) {
    uint64_t c = 0;
    for (uint32_t y0 = lane * 64; y0 < n; y0 += 64 * kWave) {
      const uint32_t len = (n - y0) < 64 ? (n - y0) : 64;
      c += __popcll(bit_window(layout, static_cast<uint64_t>(x) * n + y0, len));
    }
    c = wave_sum(c);
    if (lane == 0) cnt[x] = static_cast<uint32_t>(c);
  }
}

// Rank emission: (row_idx, col_idx)[rank] for every set bit, rank = row-major order.
__global__ void __launch_bounds__(kBlock) layout_emit_kernel(const uint64_t* __restrict__ layout, uint32_t n,
                                                             const uint32_t* __restrict__ row_start,
                                                             uint32_t* __restrict__ row_idx,
                                                             uint32_t* __restrict__ col_idx) {
  const uint32_t lane = lane_id();
  const uint32_t n_waves = gridDim.x * kWavesPerBlock;
  for (uint32_t x = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; x < n; x += n_waves) {
    uint32_t rank = row_start[x];
    for (uint32_t y0 = 0; y0 < n; y0 += 64) {
      const uint32_t len = (n - y0) < 64 ? (n - y0) : 64;
      const uint64_t w = bit_window(layout, static_cast<uint64_t>(x) * n + y0, len);  // uniform
      if (w == 0) continue;
      if ((w >> lane) & 1) {
        const uint32_t r = rank + __popcll(w & ((1ull << lane) - 1));
        row_idx[r] = x;
        col_idx[r] = y0 + lane;
      }
      rank += __popcll(w);
    }
  }
}

__global__ void add_counts_kernel(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src, uint32_t n) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x) dst[i] += src[i];
}

// --- symbolic pass ---------------------------------------------------------

// Row form of B from Bt: histogram of Bt's column index (entries [0, n_bt)); in the same
// launch, the rows of B that A references (its column index, entries [n_bt, n_bt + n_a)).
__global__ void col_hist_mark_kernel(const uint32_t* __restrict__ col_bt, uint32_t n_bt,
                                     const uint32_t* __restrict__ col_a, uint32_t n_a, uint32_t* __restrict__ hist,
                                     uint32_t* __restrict__ ref) {
  const uint64_t n = static_cast<uint64_t>(n_bt) + n_a;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    if (i < n_bt)
      atomicAdd(hist + col_bt[i], 1u);
    else
      ref[col_a[i - n_bt]] = 1u;
  }
}
__global__ void col_scatter_kernel(const uint32_t* __restrict__ row, const uint32_t* __restrict__ col, uint32_t n,
                                   const uint32_t* __restrict__ start, uint32_t* __restrict__ cursor,
                                   uint32_t* __restrict__ out) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t t = col[i];
    out[start[t] + atomicAdd(cursor + t, 1u)] = row[i];
  }
}

// Block-wide exclusive scan of one u32 per thread (kWaves waves).
template <int kWaves>
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_tmp, uint32_t* total) {
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / kWave;
  uint32_t incl = v;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t o = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), d, kWave));
    if (lane >= static_cast<uint32_t>(d)) incl += o;
  }
  if (lane == kWave - 1) s_tmp[wave] = incl;
  __syncthreads();
  uint32_t base = 0, tot = 0;
  for (int w = 0; w < kWaves; ++w) {
    const uint32_t t = s_tmp[w];
    if (w < static_cast<int>(wave)) base += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return base + incl - v;
}

// Per output row x: union over t in A.row(x) of B.row(t) in an LDS bitmap.
// EMIT = false: count[x] = |union|, tp[x] = sum |B.row(t)|; with `sums`, the block's totals are
// added to sums[0] (tile products), sums[1] (tiles of the B rows that A references, ref[x]:
// row x of B, as x runs over the strips) and sums[2] (C tiles).
// EMIT = true : write the union as ascending col_idx at off_c[x], row_idx = x.
// BLOCK threads per row (kRowBlock = 1024: a heavy row, such as a dense frontier's tile row
// against thousands of B rows, is the whole pass when few rows carry the work; 1024 threads
// took the 3-hop density-mixed product walk's symbolic passes from 39 to 16 ms and C = A*A of
// R-MAT 14 from 0.64 to 0.62 ms, with the two-hop of R-MAT 18 unchanged).
constexpr int kRowBlock = 1024;
template <bool EMIT, int BLOCK = kRowBlock>
__global__ void __launch_bounds__(BLOCK) symbolic_row_kernel(
    uint32_t n_strips, const uint32_t* __restrict__ off_a, const uint32_t* __restrict__ col_a,
    const uint32_t* __restrict__ off_b, const uint32_t* __restrict__ col_b, uint32_t* __restrict__ count,
    unsigned long long* __restrict__ tp, const uint32_t* __restrict__ off_c, uint32_t* __restrict__ row_c,
    uint32_t* __restrict__ col_c, const uint32_t* __restrict__ ref = nullptr,
    unsigned long long* __restrict__ sums = nullptr) {
  constexpr int kWaves = BLOCK / kWave;
  extern __shared__ uint64_t s_bm[];
  __shared__ uint32_t s_tmp[kWaves];
  __shared__ unsigned long long s_tp;
  const uint32_t n_words = (n_strips + 63) / 64;
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / kWave;
  uint64_t blk_tp = 0, blk_used = 0, blk_c = 0;  // thread 0's totals over the block's rows
  for (uint32_t x = blockIdx.x; x < n_strips; x += gridDim.x) {
    for (uint32_t w = threadIdx.x; w < n_words; w += BLOCK) s_bm[w] = 0;
    if (threadIdx.x == 0) s_tp = 0;
    __syncthreads();
    const uint32_t a0 = off_a[x], a1 = off_a[x + 1];
    uint64_t my_tp = 0;
    for (uint32_t k = a0 + wave; k < a1; k += kWaves) {
      const uint32_t t = col_a[k];
      const uint32_t r0 = off_b[t], r1 = off_b[t + 1];
      my_tp += r1 - r0;
      for (uint32_t j = r0 + lane; j < r1; j += kWave) {
        const uint32_t y = col_b[j];
        atomicOr(reinterpret_cast<unsigned long long*>(s_bm + (y >> 6)), 1ull << (y & 63));
      }
    }
    if (!EMIT && lane == 0 && my_tp) atomicAdd(&s_tp, static_cast<unsigned long long>(my_tp));
    __syncthreads();
    // Contiguous word range per thread so emitted columns come out ascending.
    const uint32_t per = (n_words + BLOCK - 1) / BLOCK;
    const uint32_t w0 = threadIdx.x * per;
    const uint32_t w1 = (w0 + per < n_words) ? w0 + per : n_words;
    uint32_t c = 0;
    for (uint32_t w = w0; w < w1; ++w) c += __popcll(s_bm[w]);
    uint32_t total = 0;
    const uint32_t before = block_exclusive_scan<kWaves>(c, s_tmp, &total);
    if (!EMIT) {
      if (threadIdx.x == 0) {
        count[x] = total;
        tp[x] = s_tp;
        blk_tp += s_tp;
        blk_c += total;
        if (ref && ref[x]) blk_used += off_b[x + 1] - off_b[x];
      }
    } else {
      uint32_t pos = off_c[x] + before;
      for (uint32_t w = w0; w < w1; ++w) {
        uint64_t bits = s_bm[w];
        while (bits) {
          const int b = __ffsll(static_cast<unsigned long long>(bits)) - 1;
          bits &= bits - 1;
          row_c[pos] = x;
          col_c[pos] = w * 64 + b;
          ++pos;
        }
      }
    }
    __syncthreads();
  }
  if (!EMIT && sums && threadIdx.x == 0) {
    if (blk_tp) atomicAdd(sums, static_cast<unsigned long long>(blk_tp));
    if (blk_used) atomicAdd(sums + 1, static_cast<unsigned long long>(blk_used));
    if (blk_c) atomicAdd(sums + 2, static_cast<unsigned long long>(blk_c));
  }
}

constexpr uint32_t kMaxLdsBitmapWords = 8192;  // 64 KiB: n_strips <= 524288

void launch_fill_tiles(hipStream_t s, uint32_t n_c, const uint32_t* off_a, const uint32_t* off_bt,
                       const uint32_t* row_c, const uint32_t* col_c, const uint32_t* col_a,
                       const uint32_t* col_bt, const uint64_t* data_a, const uint64_t* data_bt,
                       uint64_t* data_c) {
  if (n_c == 0) return;
  fill_tiles_kernel<<<grid_for(n_c, kWavesPerBlock, 65536), kBlock, 0, s>>>(n_c, off_a, off_bt, row_c, col_c,
                                                                            col_a, col_bt, data_a, data_bt, data_c);
  MG_LAUNCH_CHECK();
}

// Numeric pass of the handle-level product: payloads of both operands (stream-ordered in the
// context's scratch), then fill_tiles_payload_kernel; with a dense record (rec.pairs set), the
// pairs of dense tiles are recorded there and dense_contract_kernel multiplies them after.
// `pl` holds payload_words(a, bt) words of device scratch (the context's, or the static array
// of the allocation-free entry point).
size_t payload_words(const mg_tiled_matrix* a, const mg_tiled_matrix* bt) {
  return 2 * (static_cast<size_t>(a->n_nz_tile) + bt->n_nz_tile + 1);
}

void launch_payload_product(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt,
                            mg_tiled_matrix* c, WidePayload* pl, const DenseRecord& rec) {
  if (c->n_nz_tile == 0) return;
  hipStream_t st = ctx->stream;
  WidePayload* pl_a = pl;
  WidePayload* pl_bt = pl + a->n_nz_tile;
  launch_payloads(st, a, bt, pl);
  const uint32_t grid = grid_for(c->n_nz_tile, kWavesPerBlock, 65536);
  if (rec.pairs)
    fill_tiles_payload_kernel<true><<<grid, kBlock, 0, st>>>(
        c->n_nz_tile, a->offset_row.get(), bt->offset_row.get(), c->row_idx.get(), c->col_idx.get(),
        a->col_idx.get(), bt->col_idx.get(), a->data.get(), bt->data.get(), pl_a, pl_bt, c->data.get(), rec);
  else
    fill_tiles_payload_kernel<false><<<grid, kBlock, 0, st>>>(
        c->n_nz_tile, a->offset_row.get(), bt->offset_row.get(), c->row_idx.get(), c->col_idx.get(),
        a->col_idx.get(), bt->col_idx.get(), a->data.get(), bt->data.get(), pl_a, pl_bt, c->data.get(), rec);
  MG_LAUNCH_CHECK();
}


// --- structural pass shared by the handle-level products ----------------------
// Row form of B (B.row(t) = { y : Bt(y,t) present }), per-row structural counts and tile
// products of C = A*B, C's row pointer, and the totals.
struct Structural {
  uint32_t n = 0;  // n_strips
  StreamBuf<uint32_t> off_b, col_b, cnt, off_c;
  StreamBuf<unsigned long long> tp;
  uint64_t n_c = 0, tile_products = 0, right_used = 0;
};

void structural_pass(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt, Structural& S,
                     const char* who) {
  MG_REQUIRE(a->n_strips == bt->n_strips, std::string(who) + ": operands differ in n_strips");
  const uint32_t n = a->n_strips;
  const uint32_t n_words = (n + 63) / 64;
  if (n_words > kMaxLdsBitmapWords) fail(MG_ERR_UNSUPPORTED, std::string(who) + ": n_strips > 524288 not supported");
  hipStream_t s = ctx->stream;
  S.n = n;
  // One zeroed block: the totals (tile products, right tiles used, C tiles; u64), then the
  // histogram, the scatter cursors and the referenced-row flags of the strips (u32 each).
  StreamBuf<unsigned long long> zero(3 + (3ull * n + 1) / 2, s);
  unsigned long long* const sums = zero.get();
  uint32_t* const hist = reinterpret_cast<uint32_t*>(sums + 3);
  uint32_t* const cursor = hist + n;
  uint32_t* const ref = cursor + n;
  MG_HIP(hipMemsetAsync(zero.get(), 0, zero.bytes(), s));
  S.off_b.alloc(n + 1ull, s);
  S.col_b.alloc(bt->n_nz_tile, s);
  if (bt->n_nz_tile || a->n_nz_tile) {
    col_hist_mark_kernel<<<grid_for(static_cast<uint64_t>(bt->n_nz_tile) + a->n_nz_tile, 256), 256, 0, s>>>(
        bt->col_idx.get(), bt->n_nz_tile, a->col_idx.get(), a->n_nz_tile, hist, ref);
    MG_LAUNCH_CHECK();
  }
  exclusive_scan_u32(ctx, hist, S.off_b.get(), n);
  if (bt->n_nz_tile) {
    col_scatter_kernel<<<grid_for(bt->n_nz_tile, 256), 256, 0, s>>>(bt->row_idx.get(), bt->col_idx.get(),
                                                                    bt->n_nz_tile, S.off_b.get(), cursor,
                                                                    S.col_b.get());
    MG_LAUNCH_CHECK();
  }
  S.cnt.alloc(n, s);
  S.off_c.alloc(n + 1ull, s);
  S.tp.alloc(n, s);
  const size_t lds = sizeof(uint64_t) * n_words;
  const unsigned grid = grid_for(n, 1, 16384);
  // The row kernel also adds up the totals. C's tile index is 32-bit (the reference's
  // n_nz_tile is u32): the 64-bit C tile total is read with the other sums and checked before
  // anything uses the 32-bit scan.
  symbolic_row_kernel<false><<<grid, kRowBlock, lds, s>>>(n, a->offset_row.get(), a->col_idx.get(),
                                                          S.off_b.get(), S.col_b.get(), S.cnt.get(), S.tp.get(),
                                                          nullptr, nullptr, nullptr, ref, sums);
  MG_LAUNCH_CHECK();
  exclusive_scan_u32(ctx, S.cnt.get(), S.off_c.get(), n);
  unsigned long long h_sums[3] = {0, 0, 0};
  MG_HIP(hipMemcpyAsync(h_sums, sums, sizeof(h_sums), hipMemcpyDeviceToHost, s));
  MG_HIP(hipStreamSynchronize(s));  // the one synchronisation of the pass: the totals
  const uint64_t n_c64 = h_sums[2];
  MG_REQUIRE(n_c64 <= 0xFFFFFFFFull, std::string(who) + ": " + std::to_string(n_c64) +
                                         " output tiles, more than a u32 tile index holds");
  S.n_c = n_c64;
  S.tile_products = h_sums[0];
  S.right_used = h_sums[1];
}

// The structural (row, col) index of C, row-major, into row_c / col_c (S.n_c entries).
void structural_emit(mg_context* ctx, const mg_tiled_matrix* a, const Structural& S, uint32_t* row_c,
                     uint32_t* col_c) {
  if (!S.n_c) return;
  const size_t lds = sizeof(uint64_t) * ((S.n + 63) / 64);
  symbolic_row_kernel<true><<<grid_for(S.n, 1, 16384), kRowBlock, lds, ctx->stream>>>(
      S.n, a->offset_row.get(), a->col_idx.get(), S.off_b.get(), S.col_b.get(), nullptr, nullptr, S.off_c.get(),
      row_c, col_c);
  MG_LAUNCH_CHECK();
}

// --- pruned product: only the numerically non-empty tiles of C ------------------
// Final index of the appended tiles: sorted keys (structural indices) give each output
// tile its row and column; data moves from its slot. One wave per tile.
__global__ void __launch_bounds__(kBlock) gather_pruned_kernel(uint32_t n, const uint32_t* __restrict__ keys,
                                                               const uint32_t* __restrict__ slots,
                                                               const uint32_t* __restrict__ row_s,
                                                               const uint32_t* __restrict__ col_s,
                                                               const uint64_t* __restrict__ scratch,
                                                               uint32_t* __restrict__ row_c,
                                                               uint32_t* __restrict__ col_c,
                                                               uint64_t* __restrict__ data_c) {
  const uint32_t lane = lane_id();
  const uint32_t n_waves = gridDim.x * kWavesPerBlock;
  for (uint32_t j = read_first_lane(blockIdx.x * kWavesPerBlock + threadIdx.x / kWave); j < n; j += n_waves) {
    const uint32_t k = keys[j];
    data_c[static_cast<uint64_t>(j) * kTile + lane] = scratch[static_cast<uint64_t>(slots[j]) * kTile + lane];
    if (lane == 0) {
      row_c[j] = row_s[k];
      col_c[j] = col_s[k];
    }
  }
}

// Index of the flagged tiles at their exact positions (kOutPos pass): pos = exclusive scan
// of the flags.
__global__ void __launch_bounds__(kBlock) place_pruned_index_kernel(uint32_t n_s, const uint32_t* __restrict__ flag,
                                                                    const uint32_t* __restrict__ pos,
                                                                    const uint32_t* __restrict__ row_s,
                                                                    const uint32_t* __restrict__ col_s,
                                                                    uint32_t* __restrict__ row_c,
                                                                    uint32_t* __restrict__ col_c) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_s; i += gridDim.x * blockDim.x)
    if (flag[i]) {
      row_c[pos[i]] = row_s[i];
      col_c[pos[i]] = col_s[i];
    }
}

__global__ void __launch_bounds__(kBlock) iota_u32_kernel(uint32_t* __restrict__ out, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = i;
}

__global__ void __launch_bounds__(kBlock) row_hist_kernel(const uint32_t* __restrict__ row, uint32_t n,
                                                          uint32_t* __restrict__ hist) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    atomicAdd(hist + row[i], 1u);
}

template <int kOut>
void launch_pruned(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt, uint32_t n_s,
                   const uint32_t* row_s, const uint32_t* col_s, WidePayload* pl, const PrunedOut& out) {
  if (!n_s) return;
  fill_tiles_payload_kernel<false, kOut><<<grid_for(n_s, kWavesPerBlock, 65536), kBlock, 0, ctx->stream>>>(
      n_s, a->offset_row.get(), bt->offset_row.get(), row_s, col_s, a->col_idx.get(), bt->col_idx.get(),
      a->data.get(), bt->data.get(), pl, pl + a->n_nz_tile, nullptr, DenseRecord{}, out);
  MG_LAUNCH_CHECK();
}

// Per C row: features of the cost model (see mg_multiply_row_features).
__global__ void __launch_bounds__(kBlock) brow_bits_kernel(const uint32_t* __restrict__ col, uint32_t n,
                                                           const uint32_t* __restrict__ bits,
                                                           unsigned long long* __restrict__ out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    atomicAdd(out + col[i], static_cast<unsigned long long>(bits[i]));
}

__global__ void __launch_bounds__(kBlock) row_features_kernel(uint32_t n, const uint32_t* __restrict__ off_a,
                                                              const uint32_t* __restrict__ col_a,
                                                              const uint32_t* __restrict__ bits_a,
                                                              const unsigned long long* __restrict__ brow_bits,
                                                              const uint32_t* __restrict__ cnt,
                                                              const unsigned long long* __restrict__ tp,
                                                              float* __restrict__ feat) {
  const uint32_t lane = lane_id();
  const uint32_t n_waves = gridDim.x * kWavesPerBlock;
  for (uint32_t x = read_first_lane(blockIdx.x * kWavesPerBlock + threadIdx.x / kWave); x < n; x += n_waves) {
    const uint32_t a0 = off_a[x], a1 = off_a[x + 1];
    uint64_t abits = 0, bbits = 0, amax = 0;
    for (uint32_t k = a0 + lane; k < a1; k += kWave) {
      const uint64_t ab = bits_a[k];
      abits += ab;
      amax = ab > amax ? ab : amax;
      bbits += brow_bits[col_a[k]];
    }
    abits = wave_sum(abits);
    bbits = wave_sum(bbits);
    for (int o = 32; o >= 1; o >>= 1) {
      const uint64_t v = static_cast<uint64_t>(__shfl_xor(static_cast<long long>(amax), o, kWave));
      amax = v > amax ? v : amax;
    }
    if (lane == 0) {
      const float nt = static_cast<float>(a1 - a0);
      const float c = static_cast<float>(cnt[x]);
      const float t = static_cast<float>(tp[x]);
      float* f = feat + static_cast<uint64_t>(x) * kRowFeatures;
      f[0] = log1pf(c) / 16.0f;
      f[1] = log1pf(t) / 20.0f;
      f[2] = log1pf(nt) / 16.0f;
      f[3] = log1pf(static_cast<float>(abits)) / 20.0f;
      f[4] = log1pf(static_cast<float>(bbits)) / 24.0f;
      f[5] = log1pf(nt > 0 ? static_cast<float>(abits) / nt : 0.0f) / 8.0f;
      f[6] = log1pf(t > 0 ? static_cast<float>(bbits) / t : 0.0f) / 8.0f;
      f[7] = c / static_cast<float>(n);
    }
  }
}
}  // namespace

// Payload scratch of mg_multiply_numeric_noalloc: one copy per device, held from the moment
// the code object is loaded until the process ends. Host code never passes the name to a
// kernel: it stands for a host-side shadow, not a device address. noalloc_scratch_resolve
// (called by mg_context_create with the device current) looks up this device's address once;
// every later use goes through that pointer. All calls on one device share the array, so
// calls that overlap on one device (different streams or contexts) race.
static __device__ __align__(256) uint64_t scratch[kNoAllocScratchBytes / sizeof(uint64_t)];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

void noalloc_scratch_resolve(int device) {
  MG_REQUIRE(device >= 0 && device < kMaxDevices, "device index beyond the static scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* ptr = nullptr;
  MG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  MG_REQUIRE(ptr != nullptr, "hipGetSymbolAddress returned NULL for the static scratch");
  g_scratch_ptr[device] = ptr;
}

void* noalloc_scratch(int device) {
  MG_REQUIRE(device >= 0 && device < kMaxDevices, "device index beyond the static scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  MG_REQUIRE(g_scratch_ptr[device] != nullptr, "static scratch not resolved for this device");
  return g_scratch_ptr[device];
}

}  // namespace mgamd

using namespace mgamd;

extern "C" int mg_matrix_bit_and(void* stream, const uint64_t* matrix_a, const uint64_t* matrix_b,
                                 uint64_t* matrix_c, uint32_t m, uint32_t k, uint32_t n) {
  return guarded([&] {
    MG_REQUIRE(matrix_a && matrix_b && matrix_c, "mg_matrix_bit_and: NULL buffer");
    if (m == 0 || n == 0 || k == 0) return;
    const uint64_t tasks = static_cast<uint64_t>(m) * ((n + 63) / 64);
    bit_and_kernel<<<grid_for(tasks, kWavesPerBlock, 65536), kBlock, 0, static_cast<hipStream_t>(stream)>>>(
        matrix_a, matrix_b, reinterpret_cast<unsigned long long*>(matrix_c), m, k, n);
    MG_LAUNCH_CHECK();
  });
}

extern "C" int mg_init_bit_tiled_matrix_metadata_by_layout_matrix(void* stream, const uint64_t* layout_matrix,
                                                                  uint32_t n_strips, uint32_t* tile_count_row,
                                                                  uint32_t* tile_row_idx, uint32_t* tile_col_idx) {
  return guarded([&] {
    MG_REQUIRE(layout_matrix && tile_count_row && tile_row_idx && tile_col_idx,
               "mg_init_bit_tiled_matrix_metadata_by_layout_matrix: NULL buffer");
    if (n_strips == 0) return;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // Temporary row counts and their exclusive scan (the reference computes ranks
    // with an O(words) prefix popcount per bit instead, matrix_operations.cu:149-174).
    StreamBuf<uint32_t> cnt(n_strips, s), start(n_strips + 1ull, s);
    layout_row_count_kernel<<<grid_for(n_strips, kWavesPerBlock, 65536), kBlock, 0, s>>>(layout_matrix, n_strips,
                                                                                          cnt.get());
    MG_LAUNCH_CHECK();
    mg_context tmp;  // scan helper needs only the stream
    tmp.stream = s;
    exclusive_scan_u32(&tmp, cnt.get(), start.get(), n_strips);
    layout_emit_kernel<<<grid_for(n_strips, kWavesPerBlock, 65536), kBlock, 0, s>>>(
        layout_matrix, n_strips, start.get(), tile_row_idx, tile_col_idx);
    MG_LAUNCH_CHECK();
    add_counts_kernel<<<grid_for(n_strips, 256), 256, 0, s>>>(tile_count_row, cnt.get(), n_strips);
    MG_LAUNCH_CHECK();
  });
}

extern "C" int mg_fill_tiles(void* stream, size_t tile_size, size_t n_strips, size_t n_nz_tile_a,
                             size_t n_nz_tile_b, size_t n_nz_tile_c, const uint64_t* layout_matrix_c,
                             const uint32_t* tile_offset_row_a, const uint32_t* tile_offset_row_b,
                             const uint32_t* tile_offset_row_c, const uint32_t* tile_row_idx_a,
                             const uint32_t* tile_row_idx_b, const uint32_t* tile_row_idx_c,
                             const uint32_t* tile_col_idx_a, const uint32_t* tile_col_idx_b,
                             const uint32_t* tile_col_idx_c, const uint64_t* data_a, const uint64_t* data_b,
                             uint64_t* data_c) {
  (void)layout_matrix_c; (void)tile_row_idx_a; (void)tile_row_idx_b; (void)tile_offset_row_c;
  (void)n_nz_tile_a; (void)n_nz_tile_b; (void)n_strips;
  return guarded([&] {
    if (tile_size != kTile) fail(MG_ERR_UNSUPPORTED, "mg_fill_tiles: tile_size must be 64");
    if (n_nz_tile_c == 0) return;
    MG_REQUIRE(tile_offset_row_a && tile_offset_row_b && tile_row_idx_c && tile_col_idx_a && tile_col_idx_b &&
                   tile_col_idx_c && data_a && data_b && data_c,
               "mg_fill_tiles: NULL buffer");
    MG_REQUIRE(n_nz_tile_c < (1ull << 32), "mg_fill_tiles: n_nz_tile_c exceeds u32");
    launch_fill_tiles(static_cast<hipStream_t>(stream), static_cast<uint32_t>(n_nz_tile_c), tile_offset_row_a,
                      tile_offset_row_b, tile_row_idx_c, tile_col_idx_c, tile_col_idx_a, tile_col_idx_b, data_a,
                      data_b, data_c);
  });
}

extern "C" int mg_multiply_symbolic(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt,
                                    mg_tiled_matrix** c_out, mg_product_stats* stats) {
  return guarded([&] {
    MG_REQUIRE(ctx && a && bt && c_out, "mg_multiply_symbolic: NULL argument");
    *c_out = nullptr;
    DeviceGuard g(ctx->device);
    hipStream_t s = ctx->stream;
    Structural S;
    structural_pass(ctx, a, bt, S, "mg_multiply_symbolic");
    auto c = std::make_unique<mg_tiled_matrix>();
    c->device = ctx->device;
    alloc_tiled_matrix(c.get(), S.n, static_cast<uint32_t>(S.n_c));
    MG_HIP(hipMemcpyAsync(c->offset_row.get(), S.off_c.get(), sizeof(uint32_t) * (S.n + 1ull),
                          hipMemcpyDeviceToDevice, s));
    if (S.n_c) {
      MG_HIP(hipMemsetAsync(c->data.get(), 0, c->data.bytes(), s));
      structural_emit(ctx, a, S, c->row_idx.get(), c->col_idx.get());
    }
    if (stats) {
      stats->tile_products = S.tile_products;
      stats->right_tiles_used = S.right_used;
      stats->out_tiles = S.n_c;
      stats->alg_bytes = 520ull * (a->n_nz_tile + S.right_used + S.n_c) + 4ull * (2ull * S.n + 2);
    }
    *c_out = c.release();
  });
}

extern "C" int mg_tiled_matrix_prepare_operand(mg_context* ctx, mg_tiled_matrix* m) {
  return guarded([&] {
    MG_REQUIRE(ctx && m, "mg_tiled_matrix_prepare_operand: NULL argument");
    MG_REQUIRE(m->tile_size == 64, "mg_tiled_matrix_prepare_operand: tile_size must be 64");
    if (prepared(m)) return;  // kept copy still current
    DeviceGuard g(ctx->device);
    hipStream_t st = ctx->stream;
    const uint32_t n = m->n_nz_tile;
    m->prep_payload.alloc(2ull * std::max<uint32_t>(1, n));
    m->prep_bits.alloc(std::max<uint32_t>(1, n));
    if (n) {
      payload_kernel<<<grid_for(n, kWavesPerBlock, 65536), kBlock, 0, st>>>(
          m->data.get(), n, nullptr, 0, reinterpret_cast<WidePayload*>(m->prep_payload.get()), m->prep_bits.get());
      MG_LAUNCH_CHECK();
    }
    MG_HIP(hipStreamSynchronize(st));
    m->prep_version = m->version;
  });
}

extern "C" int mg_multiply_numeric(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt,
                                   mg_tiled_matrix* c) {
  return guarded([&] {
    MG_REQUIRE(ctx && a && bt && c, "mg_multiply_numeric: NULL argument");
    MG_REQUIRE(a->n_strips == bt->n_strips && c->n_strips == a->n_strips,
               "mg_multiply_numeric: operands differ in n_strips");
    DeviceGuard g(ctx->device);
    ++c->version;
    if (c->n_nz_tile == 0) return;
    WidePayload* pl = reinterpret_cast<WidePayload*>(ctx_scratch(ctx, payload_words(a, bt)));
    launch_payload_product(ctx, a, bt, c, pl, DenseRecord{});
  });
}

namespace mgamd {
namespace {

// Structural index, operand payloads and features of a pruned product.
struct PrunedPlan {
  Structural S;
  StreamBuf<uint32_t> row_s, col_s;
  WidePayload* pl = nullptr;
  StreamBuf<uint32_t> bits;  // set bits per tile of A, then of Bt (from the payload pass)
};

void pruned_prepare(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt, PrunedPlan& P,
                    const char* who) {
  hipStream_t s = ctx->stream;
  structural_pass(ctx, a, bt, P.S, who);
  const uint32_t n_s = static_cast<uint32_t>(P.S.n_c);
  P.row_s.alloc(n_s, s);
  P.col_s.alloc(n_s, s);
  structural_emit(ctx, a, P.S, P.row_s.get(), P.col_s.get());
  P.pl = reinterpret_cast<WidePayload*>(ctx_scratch(ctx, payload_words(a, bt)));
  P.bits.alloc(static_cast<size_t>(a->n_nz_tile) + bt->n_nz_tile + 1, s);
  launch_payloads(s, a, bt, P.pl, P.bits.get());
}

// Features of every C row into feat (device, n_strips x kRowFeatures).
// known_bits: the tiles' set-bit counts of A then Bt when a payload pass already made them
// (else they are counted here).
void row_features_into(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt, const Structural& S,
                       float* feat, const uint32_t* known_bits = nullptr) {
  hipStream_t s = ctx->stream;
  const uint32_t n = S.n;
  StreamBuf<uint32_t> counted(known_bits ? 0 : static_cast<size_t>(a->n_nz_tile) + bt->n_nz_tile + 1, s);
  StreamBuf<unsigned long long> brow(n, s);
  MG_HIP(hipMemsetAsync(brow.get(), 0, sizeof(unsigned long long) * std::max<uint32_t>(1, n), s));
  if (!known_bits) {
    if (a->n_nz_tile) {
      tile_bits_kernel<<<grid_for(a->n_nz_tile, kWavesPerBlock, 65536), kBlock, 0, s>>>(a->data.get(),
                                                                                         a->n_nz_tile, counted.get());
      MG_LAUNCH_CHECK();
    }
    if (bt->n_nz_tile) {
      tile_bits_kernel<<<grid_for(bt->n_nz_tile, kWavesPerBlock, 65536), kBlock, 0, s>>>(
          bt->data.get(), bt->n_nz_tile, counted.get() + a->n_nz_tile);
      MG_LAUNCH_CHECK();
    }
  }
  const uint32_t* bits_a = known_bits ? known_bits : counted.get();
  const uint32_t* bits_b = bits_a + a->n_nz_tile;
  if (bt->n_nz_tile) {
    // Set bits of each row t of B = of the Bt tiles in column t.
    brow_bits_kernel<<<grid_for(bt->n_nz_tile, kBlock), kBlock, 0, s>>>(bt->col_idx.get(), bt->n_nz_tile, bits_b,
                                                                       brow.get());
    MG_LAUNCH_CHECK();
  }
  if (n) {
    row_features_kernel<<<grid_for(n, kWavesPerBlock, 65536), kBlock, 0, s>>>(
        n, a->offset_row.get(), a->col_idx.get(), bits_a, brow.get(), S.cnt.get(), S.tp.get(), feat);
    MG_LAUNCH_CHECK();
  }
}

// Output slots from a prediction: per row min(structural, ceil(structural * clamp(y) *
// (1 + margin))); the total into *out, the unrounded prediction per row into pred.
__global__ void __launch_bounds__(kBlock) capacity_kernel(uint32_t n, const uint32_t* __restrict__ cnt,
                                                          const float* __restrict__ y, uint32_t y_stride,
                                                          float margin, float* __restrict__ pred,
                                                          unsigned long long* __restrict__ out) {
  uint64_t local = 0;
  for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < n; x += gridDim.x * blockDim.x) {
    const float c = static_cast<float>(cnt[x]);
    const float f = fminf(fmaxf(y[static_cast<uint64_t>(x) * y_stride], 0.0f), 1.0f);
    const float pr = c * f;
    pred[x] = pr;
    const float cap = fminf(ceilf(pr * (1.0f + margin)), c);
    local += static_cast<uint64_t>(cap);
  }
  local = wave_sum(local);
  if (lane_id() == 0 && local) atomicAdd(out, static_cast<unsigned long long>(local));
}

// Numeric passes of a pruned product into *c (see mg_multiply_pruned for `capacity`).
void pruned_execute(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt, PrunedPlan& P,
                    uint64_t capacity, mg_tiled_matrix* c, mg_pruned_stats& st) {
  hipStream_t s = ctx->stream;
  const Structural& S = P.S;
  const uint32_t n_s = static_cast<uint32_t>(S.n_c);
  const uint32_t* row_s = P.row_s.get();
  const uint32_t* col_s = P.col_s.get();
  st.structural_tiles = n_s;
  st.tile_products = S.tile_products;
  uint64_t n_num = 0;
  uint64_t cap = capacity == MG_PRUNED_STRUCTURAL ? n_s : std::min<uint64_t>(capacity, n_s);
  if (capacity == MG_PRUNED_EXACT || n_s == 0) {
    // Exact: flag pass, scan, then each non-empty tile written at its final position.
    StreamBuf<uint32_t> flag(n_s, s), pos(n_s + 1ull, s);
    PrunedOut fo{};
    fo.flag = flag.get();
    launch_pruned<kOutFlag>(ctx, a, bt, n_s, row_s, col_s, P.pl, fo);
    exclusive_scan_u32(ctx, flag.get(), pos.get(), n_s);
    uint32_t n32 = 0;
    MG_HIP(hipMemcpyAsync(&n32, pos.get() + n_s, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    MG_HIP(hipStreamSynchronize(s));
    n_num = n32;
    alloc_tiled_matrix(c, S.n, n32);
    if (n32) {
      PrunedOut po{};
      po.data = c->data.get();
      po.pos = pos.get();
      launch_pruned<kOutPos>(ctx, a, bt, n_s, row_s, col_s, P.pl, po);
      place_pruned_index_kernel<<<grid_for(n_s, kBlock), kBlock, 0, s>>>(n_s, flag.get(), pos.get(), row_s, col_s,
                                                                        c->row_idx.get(), c->col_idx.get());
      MG_LAUNCH_CHECK();
    }
    st.passes = 2;
    st.capacity = n32;
    st.scratch_bytes = 0;
  } else {
    // One pass into `cap` slots; a capacity below the non-empty count is found by the
    // counter, which keeps counting past it, and the pass is run again with the exact count.
    StreamBuf<unsigned long long> counter(1, s);
    st.capacity = cap;
    for (int attempt = 0;; ++attempt) {
      const uint64_t slots = std::max<uint64_t>(1, cap);
      StreamBuf<uint64_t> scratch(slots * kTile, s);
      StreamBuf<uint32_t> keys(slots, s);
      MG_HIP(hipMemsetAsync(counter.get(), 0, sizeof(unsigned long long), s));
      PrunedOut ao{};
      ao.data = scratch.get();
      ao.key = keys.get();
      ao.counter = counter.get();
      ao.capacity = cap;
      launch_pruned<kOutAppend>(ctx, a, bt, n_s, row_s, col_s, P.pl, ao);
      unsigned long long got = 0;
      MG_HIP(hipMemcpyAsync(&got, counter.get(), sizeof(got), hipMemcpyDeviceToHost, s));
      MG_HIP(hipStreamSynchronize(s));
      st.passes += 1;
      st.scratch_bytes = std::max<uint64_t>(st.scratch_bytes, slots * kTile * sizeof(uint64_t));
      if (got > cap) {
        MG_REQUIRE(attempt == 0, "mg_multiply_pruned: internal capacity overflow after the exact retry");
        st.fallback = 1;
        cap = got;
        continue;
      }
      n_num = got;
      const uint32_t n32 = static_cast<uint32_t>(got);
      alloc_tiled_matrix(c, S.n, n32);
      if (n32) {
        // Row-major order: sort the slots by structural index, then move the tiles.
        StreamBuf<uint32_t> slot_in(n32, s), keys_out(n32, s), slot_out(n32, s);
        iota_u32_kernel<<<grid_for(n32, kBlock), kBlock, 0, s>>>(slot_in.get(), n32);
        MG_LAUNCH_CHECK();
        int bits = 1;
        while ((1ull << bits) < n_s) ++bits;
        size_t temp_bytes = 0;
        MG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys.get(), keys_out.get(), slot_in.get(),
                                                  slot_out.get(), static_cast<int>(n32), 0, bits, s));
        StreamBuf<uint8_t> temp(temp_bytes, s);
        MG_HIP(hipcub::DeviceRadixSort::SortPairs(temp.get(), temp_bytes, keys.get(), keys_out.get(), slot_in.get(),
                                                  slot_out.get(), static_cast<int>(n32), 0, bits, s));
        gather_pruned_kernel<<<grid_for(n32, kWavesPerBlock, 65536), kBlock, 0, s>>>(
            n32, keys_out.get(), slot_out.get(), row_s, col_s, scratch.get(), c->row_idx.get(), c->col_idx.get(),
            c->data.get());
        MG_LAUNCH_CHECK();
      }
      break;
    }
  }
  // Row pointer of the pruned C from its row index.
  StreamBuf<uint32_t> hist(S.n, s);
  MG_HIP(hipMemsetAsync(hist.get(), 0, sizeof(uint32_t) * S.n, s));
  if (n_num) {
    row_hist_kernel<<<grid_for(n_num, kBlock), kBlock, 0, s>>>(c->row_idx.get(), static_cast<uint32_t>(n_num),
                                                               hist.get());
    MG_LAUNCH_CHECK();
  }
  exclusive_scan_u32(ctx, hist.get(), c->offset_row.get(), S.n);
  MG_HIP(hipStreamSynchronize(s));
  st.numeric_tiles = n_num;
  st.output_bytes = n_num * kTile * sizeof(uint64_t);
}

}  // namespace
}  // namespace mgamd

extern "C" int mg_multiply_pruned(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt,
                                  uint64_t capacity, mg_tiled_matrix** c_out, mg_pruned_stats* stats) {
  return guarded([&] {
    MG_REQUIRE(ctx && a && bt && c_out, "mg_multiply_pruned: NULL argument");
    *c_out = nullptr;
    DeviceGuard g(ctx->device);
    PrunedPlan P;
    pruned_prepare(ctx, a, bt, P, "mg_multiply_pruned");
    auto c = std::make_unique<mg_tiled_matrix>();
    c->device = ctx->device;
    mg_pruned_stats st{};
    pruned_execute(ctx, a, bt, P, capacity, c.get(), st);
    if (stats) *stats = st;
    *c_out = c.release();
  });
}

extern "C" int mg_multiply_pruned_model(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt,
                                        const mg_mlp* model, float margin, float* row_predicted,
                                        mg_tiled_matrix** c_out, mg_pruned_stats* stats) {
  return guarded([&] {
    MG_REQUIRE(ctx && a && bt && model && c_out, "mg_multiply_pruned_model: NULL argument");
    *c_out = nullptr;
    MG_REQUIRE(margin >= 0.0f && margin <= 16.0f, "mg_multiply_pruned_model: margin must be in [0, 16]");
    uint32_t n_in = 0, n_out = 0;
    mlp_dims(model, &n_in, &n_out);
    MG_REQUIRE(n_in == kRowFeatures && n_out >= 1,
               "mg_multiply_pruned_model: the model must take MG_ROW_FEATURES inputs");
    DeviceGuard g(ctx->device);
    hipStream_t s = ctx->stream;
    PrunedPlan P;
    pruned_prepare(ctx, a, bt, P, "mg_multiply_pruned_model");
    const uint32_t n = P.S.n;
    uint64_t capacity = MG_PRUNED_EXACT;
    if (P.S.n_c) {
      StreamBuf<float> feat(static_cast<uint64_t>(n) * kRowFeatures, s), y(static_cast<uint64_t>(n) * n_out, s),
          pred(n, s);
      StreamBuf<unsigned long long> total(1, s);
      row_features_into(ctx, a, bt, P.S, feat.get(), P.bits.get());
      if (mg_mlp_forward(ctx, model, feat.get(), n, y.get()) != MG_OK)
        fail(MG_ERR_INTERNAL, std::string("mg_multiply_pruned_model: ") + mg_last_error());
      MG_HIP(hipMemsetAsync(total.get(), 0, sizeof(unsigned long long), s));
      capacity_kernel<<<grid_for(n, kBlock, 1024), kBlock, 0, s>>>(n, P.S.cnt.get(), y.get(), n_out, margin,
                                                                  pred.get(), total.get());
      MG_LAUNCH_CHECK();
      unsigned long long h = 0;
      MG_HIP(hipMemcpyAsync(&h, total.get(), sizeof(h), hipMemcpyDeviceToHost, s));
      if (row_predicted) MG_HIP(hipMemcpyAsync(row_predicted, pred.get(), sizeof(float) * n, hipMemcpyDeviceToHost, s));
      MG_HIP(hipStreamSynchronize(s));
      capacity = std::max<unsigned long long>(1, h);
    } else if (row_predicted && n) {
      std::fill(row_predicted, row_predicted + n, 0.0f);
    }
    auto c = std::make_unique<mg_tiled_matrix>();
    c->device = ctx->device;
    mg_pruned_stats st{};
    pruned_execute(ctx, a, bt, P, capacity, c.get(), st);
    if (stats) *stats = st;
    *c_out = c.release();
  });
}

extern "C" int mg_multiply_row_features(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt,
                                        float* features, uint32_t* structural, uint64_t* tile_products) {
  return guarded([&] {
    MG_REQUIRE(ctx && a && bt && features, "mg_multiply_row_features: NULL argument");
    DeviceGuard g(ctx->device);
    hipStream_t s = ctx->stream;
    Structural S;
    structural_pass(ctx, a, bt, S, "mg_multiply_row_features");
    const uint32_t n = S.n;
    StreamBuf<float> feat(static_cast<uint64_t>(n) * kRowFeatures, s);
    row_features_into(ctx, a, bt, S, feat.get());
    if (n) {
      MG_HIP(hipMemcpyAsync(features, feat.get(), sizeof(float) * n * kRowFeatures, hipMemcpyDeviceToHost, s));
      if (structural)
        MG_HIP(hipMemcpyAsync(structural, S.cnt.get(), sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
    }
    MG_HIP(hipStreamSynchronize(s));
    if (tile_products) *tile_products = S.tile_products;
  });
}

extern "C" int mg_multiply_numeric_noalloc(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt,
                                           mg_tiled_matrix* c) {
  return guarded([&] {
    MG_REQUIRE(ctx && a && bt && c, "mg_multiply_numeric_noalloc: NULL argument");
    MG_REQUIRE(a->n_strips == bt->n_strips && c->n_strips == a->n_strips,
               "mg_multiply_numeric_noalloc: operands differ in n_strips");
    ++c->version;
    const size_t words = payload_words(a, bt);
    if (words * sizeof(uint64_t) > kNoAllocScratchBytes)
      invalid("mg_multiply_numeric_noalloc: " + std::to_string(a->n_nz_tile) + " + " +
              std::to_string(bt->n_nz_tile) + " operand tiles need " + std::to_string(words * sizeof(uint64_t)) +
              " bytes of payload scratch, more than the static " + std::to_string(kNoAllocScratchBytes) +
              " (at most " + std::to_string(kNoAllocMaxTiles) + " tiles in A and Bt together)");
    DeviceGuard g(ctx->device);
    if (c->n_nz_tile == 0) return;
    // The device address of this device's copy of `scratch`, looked up when the context was
    // created; the kernels write every payload word they read, so nothing is cleared.
    WidePayload* pl = reinterpret_cast<WidePayload*>(noalloc_scratch(ctx->device));
    launch_payload_product(ctx, a, bt, c, pl, DenseRecord{});
  });
}

extern "C" int mg_multiply_numeric_dense(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt,
                                         mg_tiled_matrix* c, uint32_t min_bits, uint64_t* dense_products) {
  return guarded([&] {
    MG_REQUIRE(ctx && a && bt && c, "mg_multiply_numeric_dense: NULL argument");
    MG_REQUIRE(a->n_strips == bt->n_strips && c->n_strips == a->n_strips,
               "mg_multiply_numeric_dense: operands differ in n_strips");
    DeviceGuard g(ctx->device);
    ++c->version;
    hipStream_t st = ctx->stream;
    StreamBuf<uint32_t> ba(a->n_nz_tile, st), bb(bt->n_nz_tile, st);
    if (a->n_nz_tile) {
      tile_bits_kernel<<<grid_for(a->n_nz_tile, kWavesPerBlock, 65536), kBlock, 0, st>>>(a->data.get(), a->n_nz_tile,
                                                                                         ba.get());
      MG_LAUNCH_CHECK();
    }
    if (bt->n_nz_tile) {
      tile_bits_kernel<<<grid_for(bt->n_nz_tile, kWavesPerBlock, 65536), kBlock, 0, st>>>(bt->data.get(),
                                                                                          bt->n_nz_tile, bb.get());
      MG_LAUNCH_CHECK();
    }
    // Exact number of dense pairs: the capacity of the pair and task lists.
    uint64_t n_pairs = 0;
    if (a->n_nz_tile && bt->n_nz_tile) {
      StreamBuf<uint32_t> colcount(a->n_strips, st);
      StreamBuf<unsigned long long> total(1, st);
      MG_HIP(hipMemsetAsync(colcount.get(), 0, sizeof(uint32_t) * a->n_strips, st));
      MG_HIP(hipMemsetAsync(total.get(), 0, sizeof(unsigned long long), st));
      dense_col_count_kernel<<<grid_for(bt->n_nz_tile, kBlock), kBlock, 0, st>>>(bt->col_idx.get(), bt->n_nz_tile,
                                                                                bb.get(), min_bits, colcount.get());
      MG_LAUNCH_CHECK();
      dense_pair_count_kernel<<<grid_for(a->n_nz_tile, kBlock, 4096), kBlock, 0, st>>>(
          a->col_idx.get(), a->n_nz_tile, ba.get(), min_bits, colcount.get(), total.get());
      MG_LAUNCH_CHECK();
      unsigned long long h = 0;
      MG_HIP(hipMemcpyAsync(&h, total.get(), sizeof(h), hipMemcpyDeviceToHost, st));
      MG_HIP(hipStreamSynchronize(st));
      n_pairs = h;
    }
    MG_REQUIRE(n_pairs < (1ull << 31), "mg_multiply_numeric_dense: more than 2^31 dense tile pairs");
    const uint32_t cap = static_cast<uint32_t>(n_pairs);
    StreamBuf<uint2> pairs(cap, st);
    StreamBuf<DenseTask> tasks(cap, st);
    StreamBuf<unsigned long long> counter(1, st);
    MG_HIP(hipMemsetAsync(counter.get(), 0, sizeof(unsigned long long), st));
    DenseRecord rec{pairs.get(), tasks.get(), counter.get(), cap, ba.get(), bb.get(), min_bits};
    WidePayload* pl = reinterpret_cast<WidePayload*>(ctx_scratch(ctx, payload_words(a, bt)));
    launch_payload_product(ctx, a, bt, c, pl, rec);
    if (cap && c->n_nz_tile) {
      dense_contract_kernel<<<grid_for(cap, kWavesPerBlock, 16384), kBlock, 0, st>>>(
          tasks.get(), counter.get(), cap, pairs.get(), a->data.get(), bt->data.get(), c->data.get());
      MG_LAUNCH_CHECK();
    }
    unsigned long long h = 0;
    MG_HIP(hipMemcpyAsync(&h, counter.get(), sizeof(h), hipMemcpyDeviceToHost, st));
    MG_HIP(hipStreamSynchronize(st));
    const uint64_t recorded = h & 0xFFFFFFFFull;
    MG_REQUIRE(recorded <= n_pairs, "mg_multiply_numeric_dense: more dense pairs recorded than counted");
    if (dense_products) *dense_products = recorded;
  });
}
