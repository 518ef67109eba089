// Input production on the device: counter-based R-MAT / Erdos-Renyi / planted
// community generators, edge-list set operations, and the edge list -> bit-tile
// converter (same output layout as Edgelist2BitTiledMatrix,
// core/util/format_converter.cuh:292-415, built by sorting instead of a per-edge
// prefix popcount over the n_strips^2 occupancy bitmap).
//
// Generator definitions (restated independently by the CPU oracle):
//   R-MAT  edge e, level l (0 = most significant bit): r = sm64_at(K, e*scale + l) >> 32;
//          r < T(a) -> (0,0); r < T(a+b) -> (0,1); r < T(a+b+c) -> (1,0); else (1,1),
//          where (src bit, dst bit) are bit scale-1-l and T(p) = min(floor(p*2^32), 2^32-1).
//   Bernoulli word (ER, communities): 64 independent bits, bit i of draw j is bit
//          m-1-j (MSB first) of lane i's m-bit uniform U_i; the word is [U_i < p_num].
//          Word (u, column block j) uses draws sm64_at(K, (u*n_blocks + j)*m + step).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include <hipcub/hipcub.hpp>

#include "mg_device.hpp"
#include "mg_internal.hpp"

namespace mgamd {
namespace {

constexpr uint64_t kSentinel = ~0ull;

__global__ void __launch_bounds__(256) rmat_kernel(uint64_t n_draws, uint32_t scale, uint32_t ta, uint32_t tab,
                                                   uint32_t tabc, uint64_t key, uint64_t* __restrict__ out) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n_draws; e += stride) {
    uint32_t u = 0, v = 0;
    for (uint32_t l = 0; l < scale; ++l) {
      const uint32_t r = static_cast<uint32_t>(sm64_at(key, e * scale + l) >> 32);
      const uint32_t bit = 1u << (scale - 1 - l);
      if (r < ta) {
      } else if (r < tab) {
        v |= bit;
      } else if (r < tabc) {
        u |= bit;
      } else {
        u |= bit;
        v |= bit;
      }
    }
    out[e] = (u == v) ? kSentinel : ((static_cast<uint64_t>(u) << 32) | v);
  }
}

__device__ inline uint64_t bernoulli_word(uint64_t key, uint64_t word_ctr, uint64_t p_num, uint32_t m) {
  uint64_t lt = 0, eq = ~0ull;
  for (uint32_t j = 0; j < m; ++j) {
    const uint32_t bitpos = m - 1 - j;
    const uint64_t tb = ((p_num >> bitpos) & 1) ? ~0ull : 0ull;
    const uint64_t r = sm64_at(key, word_ctr * m + j);
    lt |= eq & ~r & tb;
    eq &= ~(r ^ tb);
    if (eq == 0) break;
  }
  return lt;
}

// Bernoulli adjacency rows. Row u covers column blocks [jb0(u), jb1(u)); the
// word mask also clears columns outside [c0, c1) and the diagonal.
// Mode 0 counts set bits per row into cnt[row]; mode 1 writes edges at off[row].
struct BernoulliSpec {
  uint32_t n_vertices;
  uint32_t n_blocks;   // ceil(n_vertices / 64): counter stride per row
  uint64_t p_num;
  uint32_t m;
  uint64_t key;
  // communities: row r -> vertex base + (r / size) * stride + r % size
  uint32_t n_rows;
  uint32_t comm_size;  // 0 = plain ER (row r = vertex r, all columns)
  uint32_t comm_stride;
};

__device__ inline void row_range(const BernoulliSpec& s, uint32_t row, uint32_t* u, uint32_t* c0, uint32_t* c1) {
  if (s.comm_size == 0) {
    *u = row;
    *c0 = 0;
    *c1 = s.n_vertices;
  } else {
    const uint32_t c = row / s.comm_size;
    *c0 = c * s.comm_stride;
    *u = *c0 + row % s.comm_size;
    *c1 = *c0 + s.comm_size;
  }
}

template <bool WRITE>
__global__ void __launch_bounds__(256) bernoulli_rows_kernel(BernoulliSpec s, uint32_t* __restrict__ cnt,
                                                             const uint64_t* __restrict__ off,
                                                             uint64_t* __restrict__ out) {
  const uint32_t lane = lane_id();
  const uint32_t n_waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t row = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; row < s.n_rows; row += n_waves) {
    uint32_t u, c0, c1;
    row_range(s, row, &u, &c0, &c1);
    const uint32_t jb0 = c0 / 64, jb1 = (c1 + 63) / 64;
    uint64_t pos = WRITE ? off[row] : 0;
    uint32_t total = 0;
    for (uint32_t jbase = jb0; jbase < jb1; jbase += kWave) {
      const uint32_t j = jbase + lane;
      uint64_t w = 0;
      if (j < jb1) {
        w = bernoulli_word(s.key, static_cast<uint64_t>(u) * s.n_blocks + j, s.p_num, s.m);
        const uint32_t lo = j * 64;
        // keep columns in [c0, c1)
        if (c0 > lo) w &= ~0ull << (c0 - lo);
        if (c1 < lo + 64) w &= (c1 - lo) >= 64 ? ~0ull : ((1ull << (c1 - lo)) - 1);
        if (u >= lo && u < lo + 64) w &= ~(1ull << (u - lo));
      }
      const uint32_t pc = __popcll(w);
      if (WRITE) {
        // exclusive prefix of pc over lanes
        uint32_t incl = pc;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
          const uint32_t o = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), d, kWave));
          if (lane >= static_cast<uint32_t>(d)) incl += o;
        }
        uint64_t p = pos + (incl - pc);
        while (w) {
          const int b = __ffsll(static_cast<unsigned long long>(w)) - 1;
          w &= w - 1;
          out[p++] = (static_cast<uint64_t>(u) << 32) | (j * 64 + b);
        }
        pos += static_cast<uint32_t>(__shfl(static_cast<int>(incl), kWave - 1, kWave));
      } else {
        total += pc;
      }
    }
    if (!WRITE) {
      const uint64_t t = wave_sum(total);
      if (lane == 0) cnt[row] = static_cast<uint32_t>(t);
    }
  }
}

__global__ void widen_kernel(const uint32_t* __restrict__ in, uint64_t* __restrict__ out, uint32_t n) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x) out[i] = in[i];
}

__global__ void pack_edges_kernel(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint64_t n,
                                  uint64_t* __restrict__ out) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x) out[i] = (static_cast<uint64_t>(src[i]) << 32) | dst[i];
}

// Tile sort key: (tile_row * n_strips + tile_col) << 12 | (row_in_tile << 6 | col_in_tile).
__global__ void tile_key_kernel(const uint64_t* __restrict__ edges, uint64_t n, uint32_t n_strips, int transpose,
                                uint64_t* __restrict__ out) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t a = static_cast<uint32_t>(edges[i] >> 32);
    uint32_t b = static_cast<uint32_t>(edges[i]);
    if (transpose) { const uint32_t t = a; a = b; b = t; }
    const uint64_t tile = static_cast<uint64_t>(a >> 6) * n_strips + (b >> 6);
    out[i] = (tile << 12) | ((a & 63u) << 6) | (b & 63u);
  }
}

__global__ void tile_flag_kernel(const uint64_t* __restrict__ keys, uint64_t n, uint32_t* __restrict__ flag) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x) flag[i] = (i == 0 || (keys[i] >> 12) != (keys[i - 1] >> 12)) ? 1u : 0u;
}

__global__ void tile_fill_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ flag,
                                 const uint32_t* __restrict__ incl, uint64_t n, uint32_t n_strips,
                                 uint32_t* __restrict__ row_idx, uint32_t* __restrict__ col_idx,
                                 uint32_t* __restrict__ row_cnt, unsigned long long* __restrict__ data) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t k = keys[i];
    const uint32_t t = incl[i] - 1;
    const uint64_t tile = k >> 12;
    if (flag[i]) {
      const uint32_t x = static_cast<uint32_t>(tile / n_strips);
      row_idx[t] = x;
      col_idx[t] = static_cast<uint32_t>(tile % n_strips);
      atomicAdd(row_cnt + x, 1u);
    }
    const uint32_t r = static_cast<uint32_t>((k >> 6) & 63), c = static_cast<uint32_t>(k & 63);
    atomicOr(data + static_cast<uint64_t>(t) * kTile + r, 1ull << c);
  }
}

__global__ void __launch_bounds__(256) edge_digest_kernel(const uint64_t* __restrict__ keys, uint64_t n,
                                                          unsigned long long* __restrict__ out) {
  uint64_t acc = 0;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    acc += digest_term(keys[i], 0xE0E0ull);
  acc = wave_sum(acc);
  if (lane_id() == 0) atomicAdd(out, static_cast<unsigned long long>(acc));
}

int bits_for(uint64_t v) {  // number of bits needed to represent v (v >= 1)
  int b = 0;
  while (b < 64 && (v >> b) != 0) ++b;
  return b;
}

uint32_t prob_threshold(double p) {
  const double t = std::floor(p * 4294967296.0);
  if (!(t > 0.0)) return 0;
  if (t >= 4294967295.0) return 4294967295u;
  return static_cast<uint32_t>(t);
}

std::unique_ptr<mg_edge_list> finish_edges(mg_context* ctx, DevBuf<uint64_t>&& keys, uint64_t n, uint32_t n_vertices,
                                           bool drop_sentinel) {
  auto el = std::make_unique<mg_edge_list>();
  el->device = ctx->device;
  el->n_vertices = n_vertices;
  uint64_t u = sort_unique_u64(ctx, keys, n, 64);
  if (drop_sentinel && u > 0) {
    uint64_t last = 0;
    MG_HIP(hipMemcpyAsync(&last, keys.get() + (u - 1), sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    MG_HIP(hipStreamSynchronize(ctx->stream));
    if (last == kSentinel) --u;
  }
  el->n_edges = u;
  el->keys = std::move(keys);
  return el;
}

std::unique_ptr<mg_edge_list> bernoulli_edges(mg_context* ctx, const BernoulliSpec& s) {
  hipStream_t st = ctx->stream;
  DevBuf<uint32_t> cnt(s.n_rows);
  const unsigned grid = grid_for(s.n_rows, 4, 65536);
  bernoulli_rows_kernel<false><<<grid, 256, 0, st>>>(s, cnt.get(), nullptr, nullptr);
  MG_LAUNCH_CHECK();
  // 64-bit offsets: per-row counts are small, totals may exceed 2^32.
  DevBuf<uint64_t> cnt64(s.n_rows), off(s.n_rows + 1ull);
  widen_kernel<<<grid_for(s.n_rows, 256), 256, 0, st>>>(cnt.get(), cnt64.get(), s.n_rows);
  MG_LAUNCH_CHECK();
  {
    size_t tb = 0;
    MG_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb, cnt64.get(), off.get() + 1, static_cast<int>(s.n_rows), st));
    DevBuf<uint8_t> tmp(tb ? tb : 1);
    MG_HIP(hipMemsetAsync(off.get(), 0, sizeof(uint64_t), st));
    MG_HIP(hipcub::DeviceScan::InclusiveSum(tmp.get(), tb, cnt64.get(), off.get() + 1, static_cast<int>(s.n_rows), st));
  }
  uint64_t total = 0;
  MG_HIP(hipMemcpyAsync(&total, off.get() + s.n_rows, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
  MG_HIP(hipStreamSynchronize(st));
  DevBuf<uint64_t> keys(total ? total : 1);
  if (total) {
    bernoulli_rows_kernel<true><<<grid, 256, 0, st>>>(s, nullptr, off.get(), keys.get());
    MG_LAUNCH_CHECK();
  }
  return finish_edges(ctx, std::move(keys), total, s.n_vertices, false);
}

}  // namespace
}  // namespace mgamd

using namespace mgamd;

extern "C" int mg_edges_rmat(mg_context* ctx, uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                             uint64_t seed, mg_edge_list** out) {
  return guarded([&] {
    MG_REQUIRE(ctx && out, "mg_edges_rmat: NULL argument");
    *out = nullptr;
    MG_REQUIRE(scale >= 1 && scale <= 31, "mg_edges_rmat: scale must be in [1, 31]");
    MG_REQUIRE(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, "mg_edges_rmat: invalid quadrant probabilities");
    const uint64_t n_draws = static_cast<uint64_t>(edge_factor) << scale;
    MG_REQUIRE(n_draws < (1ull << 31), "mg_edges_rmat: more than 2^31 draws");
    DeviceGuard g(ctx->device);
    DevBuf<uint64_t> keys(n_draws ? n_draws : 1);
    if (n_draws) {
      rmat_kernel<<<grid_for(n_draws, 256, 65536), 256, 0, ctx->stream>>>(
          n_draws, scale, prob_threshold(a), prob_threshold(a + b), prob_threshold(a + b + c),
          stream_key(seed, kTagRmat), keys.get());
      MG_LAUNCH_CHECK();
    }
    *out = finish_edges(ctx, std::move(keys), n_draws, 1u << scale, true).release();
  });
}

extern "C" int mg_edges_er(mg_context* ctx, uint32_t log2_n, uint64_t p_num, uint32_t p_log2_den, uint64_t seed,
                           mg_edge_list** out) {
  return guarded([&] {
    MG_REQUIRE(ctx && out, "mg_edges_er: NULL argument");
    *out = nullptr;
    MG_REQUIRE(log2_n >= 1 && log2_n <= 31, "mg_edges_er: log2_n must be in [1, 31]");
    MG_REQUIRE(p_log2_den >= 1 && p_log2_den <= 63 && p_num < (1ull << p_log2_den),
               "mg_edges_er: p must be p_num / 2^m with 1 <= m <= 63 and p_num < 2^m");
    DeviceGuard g(ctx->device);
    BernoulliSpec s{};
    s.n_vertices = 1u << log2_n;
    s.n_blocks = (s.n_vertices + 63) / 64;
    s.p_num = p_num;
    s.m = p_log2_den;
    s.key = stream_key(seed, kTagEr);
    s.n_rows = s.n_vertices;
    s.comm_size = 0;
    *out = bernoulli_edges(ctx, s).release();
  });
}

extern "C" int mg_edges_communities(mg_context* ctx, uint32_t n_vertices, uint32_t n_comm, uint32_t comm_size,
                                    uint32_t stride, uint64_t p_num, uint32_t p_log2_den, uint64_t seed,
                                    mg_edge_list** out) {
  return guarded([&] {
    MG_REQUIRE(ctx && out, "mg_edges_communities: NULL argument");
    *out = nullptr;
    MG_REQUIRE(n_comm >= 1 && comm_size >= 1, "mg_edges_communities: empty community spec");
    MG_REQUIRE(static_cast<uint64_t>(n_comm - 1) * stride + comm_size <= n_vertices,
               "mg_edges_communities: communities exceed n_vertices");
    MG_REQUIRE(stride >= comm_size || n_comm == 1, "mg_edges_communities: communities overlap");
    MG_REQUIRE(p_log2_den >= 1 && p_log2_den <= 63 && p_num < (1ull << p_log2_den),
               "mg_edges_communities: p must be p_num / 2^m with 1 <= m <= 63 and p_num < 2^m");
    MG_REQUIRE(static_cast<uint64_t>(n_comm) * comm_size < (1ull << 32), "mg_edges_communities: too many rows");
    DeviceGuard g(ctx->device);
    BernoulliSpec s{};
    s.n_vertices = n_vertices;
    s.n_blocks = (n_vertices + 63) / 64;
    s.p_num = p_num;
    s.m = p_log2_den;
    s.key = stream_key(seed, kTagCommunity);
    s.n_rows = n_comm * comm_size;
    s.comm_size = comm_size;
    s.comm_stride = stride;
    *out = bernoulli_edges(ctx, s).release();
  });
}

extern "C" int mg_edges_union(mg_context* ctx, const mg_edge_list* a, const mg_edge_list* b, mg_edge_list** out) {
  return guarded([&] {
    MG_REQUIRE(ctx && a && b && out, "mg_edges_union: NULL argument");
    *out = nullptr;
    MG_REQUIRE(a->n_vertices == b->n_vertices, "mg_edges_union: vertex counts differ");
    DeviceGuard g(ctx->device);
    const uint64_t n = a->n_edges + b->n_edges;
    DevBuf<uint64_t> keys(n ? n : 1);
    if (a->n_edges)
      MG_HIP(hipMemcpyAsync(keys.get(), a->keys.get(), a->n_edges * 8, hipMemcpyDeviceToDevice, ctx->stream));
    if (b->n_edges)
      MG_HIP(hipMemcpyAsync(keys.get() + a->n_edges, b->keys.get(), b->n_edges * 8, hipMemcpyDeviceToDevice,
                            ctx->stream));
    *out = finish_edges(ctx, std::move(keys), n, a->n_vertices, false).release();
  });
}

namespace mgamd {
namespace {

// Grid-cut block membership of each edge key (src << 32 | dst): src in block bx, dst in by.
__global__ void grid_block_flag_kernel(const uint64_t* __restrict__ keys, uint64_t n, uint32_t scope, uint32_t bx,
                                       uint32_t by, uint32_t* __restrict__ flag) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t k = keys[i];
    const uint32_t s = static_cast<uint32_t>(k >> 32), d = static_cast<uint32_t>(k);
    flag[i] = (s / scope == bx && d / scope == by) ? 1u : 0u;
  }
}

// Stable compaction of the block's edges into local ids (u - bx*S, v - by*S): the order of
// the input keys is kept, so the output is sorted and unique as well.
__global__ void grid_block_scatter_kernel(const uint64_t* __restrict__ keys, uint64_t n, uint32_t scope, uint32_t bx,
                                          uint32_t by, const uint32_t* __restrict__ flag,
                                          const uint32_t* __restrict__ pos, uint64_t* __restrict__ out) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    if (!flag[i]) continue;
    const uint64_t k = keys[i];
    const uint64_t s = (k >> 32) - static_cast<uint64_t>(bx) * scope;
    const uint64_t d = (k & 0xFFFFFFFFull) - static_cast<uint64_t>(by) * scope;
    out[pos[i]] = (s << 32) | d;
  }
}

// Tile -> edge-list expansion in (source, destination) order without a sort. The tiles of
// each strip x are cut into pieces of kPieceTiles consecutive tiles (ascending column y); one
// wave per piece. Counts (lane = row r, over the piece's tiles) are laid out in the output
// order (strip x, row r, piece) so that their exclusive scan gives every (row, piece) its
// first output slot. The write pass takes one wave per (piece, quarter of 16 rows) and lane =
// tile: each lane reads its tile's 16 rows (one 128-byte line), and for each row a prefix sum of
// the lanes' bit counts places each lane's edges, so the row's run of the piece is written in
// (tile, column) order, which is the destination order: by each lane for its own few bits, or,
// when the row holds more than four bits per tile that has any, tile by tile by the whole wave
// (lane = column). drop_diagonal leaves out the bits with x*64 + r == y*64 + c.
constexpr uint32_t kPieceTiles = 64;

struct PieceRef {
  uint32_t strip, first_tile, n_tiles, n_pieces_in_strip, piece, count_base;
};

__global__ void __launch_bounds__(256) expand_pieces_kernel(const uint32_t* __restrict__ offset_row, uint32_t n_strips,
                                                            const uint32_t* __restrict__ piece_off,
                                                            PieceRef* __restrict__ pieces) {
  for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < n_strips; x += gridDim.x * blockDim.x) {
    const uint32_t t0 = offset_row[x], t1 = offset_row[x + 1];
    const uint32_t np = piece_off[x + 1] - piece_off[x];
    for (uint32_t p = 0; p < np; ++p) {
      const uint32_t first = t0 + p * kPieceTiles;
      pieces[piece_off[x] + p] = {x, first, min(kPieceTiles, t1 - first), np, p, piece_off[x] * 64};
    }
  }
}

__global__ void __launch_bounds__(256) count_rows_kernel(const PieceRef* __restrict__ pieces, uint32_t n_pieces,
                                                         const uint32_t* __restrict__ col_idx,
                                                         const uint64_t* __restrict__ data, int drop_diagonal,
                                                         uint32_t* __restrict__ count) {
  const uint32_t lane = lane_id();
  const uint32_t n_waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t i = read_first_lane(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave); i < n_pieces;
       i += n_waves) {
    const PieceRef pc = pieces[i];
    const uint64_t slot = static_cast<uint64_t>(pc.count_base) + static_cast<uint64_t>(lane) * pc.n_pieces_in_strip +
                          pc.piece;
    uint32_t c_total = 0;
    for (uint32_t k = 0; k < pc.n_tiles; ++k) {
      const uint32_t t = pc.first_tile + k;
      uint64_t w = data[static_cast<uint64_t>(t) * 64 + lane];
      if (drop_diagonal && col_idx[t] == pc.strip) w &= ~(1ull << lane);
      c_total += __popcll(w);
    }
    count[slot] = c_total;
  }
}

__global__ void __launch_bounds__(256) write_rows_kernel(const PieceRef* __restrict__ pieces, uint32_t n_pieces,
                                                         const uint32_t* __restrict__ col_idx,
                                                         const uint64_t* __restrict__ data, int drop_diagonal,
                                                         const uint32_t* __restrict__ start,
                                                         uint64_t* __restrict__ out) {
  typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
  const uint32_t lane = lane_id();
  const uint32_t n_waves = gridDim.x * (blockDim.x / kWave);
  const uint64_t n_items = static_cast<uint64_t>(n_pieces) * 4;  // (piece, 16-row quarter)
  for (uint64_t i = read_first_lane(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave); i < n_items;
       i += n_waves) {
    const PieceRef pc = pieces[i / 4];
    const uint32_t r0 = 16 * static_cast<uint32_t>(i % 4);
    const bool has = lane < pc.n_tiles;  // lane = tile k of the piece
    const uint32_t t = pc.first_tile + (has ? lane : 0u);
    const uint64_t y = has ? col_idx[t] : 0u;
    const bool diag = drop_diagonal && has && y == pc.strip;
    const u64x2* const tile = reinterpret_cast<const u64x2*>(data + static_cast<uint64_t>(t) * 64);
    // First output position of the quarter's rows' runs (lane q: row r0 + q), read out per row.
    const uint32_t row_start = lane < 16 ? start[static_cast<uint64_t>(pc.count_base) +
                                                 static_cast<uint64_t>(r0 + lane) * pc.n_pieces_in_strip + pc.piece]
                                         : 0u;
    {
      // Rows r0 .. r0 + 15 of this lane's tile: one 128-byte line per lane.
      u64x2 v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = has ? tile[r0 / 2 + q] : u64x2{0ull, 0ull};
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const uint32_t r = r0 + q;
        uint64_t w = (q & 1) ? v[q / 2].y : v[q / 2].x;
        if (diag) w &= ~(1ull << r);
        const uint64_t act = ballot(w != 0);
        if (!act) continue;
        // Exclusive prefix of the lanes' bit counts: this lane's first position in the run.
        const uint32_t n = static_cast<uint32_t>(__popcll(w));
        uint32_t incl = n;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
          const uint32_t o = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), d, kWave));
          if (lane >= static_cast<uint32_t>(d)) incl += o;
        }
        uint64_t pos = static_cast<uint64_t>(static_cast<uint32_t>(
                           __builtin_amdgcn_readlane(static_cast<int>(row_start), q))) +
                       (incl - n);
        const uint64_t src = (static_cast<uint64_t>(pc.strip) * 64 + r) << 32;
        const uint32_t total = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), kWave - 1));
        if (total > 4u * static_cast<uint32_t>(__popcll(act))) {
          // Dense row: tile by tile, the whole wave writes the tile's edges (lane = column).
          for (uint64_t a = act; a; a &= a - 1) {
            const int k = __ffsll(static_cast<unsigned long long>(a)) - 1;
            const uint64_t m = read_lane(w, k);
            const uint64_t p = read_lane(pos, k);
            const uint64_t yk = read_lane(y, k);
            if ((m >> lane) & 1) {
              const uint32_t below = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                               __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
              out[p + below] = src | (yk * 64 + lane);
            }
          }
        } else {
          // Sparse row: each lane writes its few edges itself.
          while (w) {
            const uint64_t c = static_cast<uint64_t>(__ffsll(static_cast<unsigned long long>(w)) - 1);
            w &= w - 1;
            out[pos++] = src | (y * 64 + c);
          }
        }
      }
    }
  }
}

// Keep-flags of a stable filter: edge keys whose source and destination differ.
__global__ void off_diagonal_flag_kernel(const uint64_t* __restrict__ keys, uint64_t n, uint32_t* __restrict__ flag) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    flag[i] = (keys[i] >> 32) != (keys[i] & 0xFFFFFFFFull) ? 1u : 0u;
}

__global__ void compact_keys_kernel(const uint64_t* __restrict__ keys, uint64_t n, const uint32_t* __restrict__ flag,
                                    const uint32_t* __restrict__ pos, uint64_t* __restrict__ out) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    if (flag[i]) out[pos[i]] = keys[i];
}

std::unique_ptr<mg_edge_list> tiles_to_edges(mg_context* ctx, const mg_tiled_matrix* m, bool drop_diagonal);

// The edges of el whose source and destination differ, as a new list (order kept).
std::unique_ptr<mg_edge_list> without_self_loops(mg_context* ctx, const mg_edge_list* el) {
  hipStream_t st = ctx->stream;
  const uint64_t n = el->n_edges;
  MG_REQUIRE(n < (1ull << 31), "edge list filter: more than 2^31 edges");
  uint32_t count = 0;
  DevBuf<uint32_t> flag(n ? n : 1), pos(n + 1);
  if (n) {
    off_diagonal_flag_kernel<<<grid_for(n, 256), 256, 0, st>>>(el->keys.get(), n, flag.get());
    MG_LAUNCH_CHECK();
    exclusive_scan_u32(ctx, flag.get(), pos.get(), static_cast<uint32_t>(n));
    MG_HIP(hipMemcpyAsync(&count, pos.get() + n, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    MG_HIP(hipStreamSynchronize(st));
  }
  auto out = std::make_unique<mg_edge_list>();
  out->device = ctx->device;
  out->n_vertices = el->n_vertices;
  out->n_edges = count;
  out->keys.alloc(count ? count : 1);
  if (count) {
    compact_keys_kernel<<<grid_for(n, 256), 256, 0, st>>>(el->keys.get(), n, flag.get(), pos.get(), out->keys.get());
    MG_LAUNCH_CHECK();
  }
  MG_HIP(hipStreamSynchronize(st));
  return out;
}

}  // namespace
}  // namespace mgamd

extern "C" int mg_edges_two_hop(mg_context* ctx, const mg_edge_list* el, mg_edge_list** out) {
  return guarded([&] {
    MG_REQUIRE(ctx && el && out, "mg_edges_two_hop: NULL argument");
    *out = nullptr;
    DeviceGuard g(ctx->device);
    // Self-loops are neither first nor second steps, and a walk back to its source is no pair:
    // A' = A without its diagonal, C = A' * A' on the bit tiles, then C without its diagonal.
    auto a_el = without_self_loops(ctx, el);
    struct Owned {
      mg_tiled_matrix *a = nullptr, *at = nullptr, *c = nullptr;
      ~Owned() {
        if (a) mg_tiled_matrix_destroy(a);
        if (at) mg_tiled_matrix_destroy(at);
        if (c) mg_tiled_matrix_destroy(c);
      }
    } o;
    auto check = [](int rc) {
      if (rc != MG_OK) fail(rc, mg_last_error());
    };
    check(mg_edges_to_tiled_matrix(ctx, a_el.get(), 0, &o.a));
    check(mg_edges_to_tiled_matrix(ctx, a_el.get(), 1, &o.at));
    check(mg_multiply_symbolic(ctx, o.a, o.at, &o.c, nullptr));
    check(mg_multiply_numeric(ctx, o.a, o.at, o.c));
    auto r = tiles_to_edges(ctx, o.c, true);
    r->n_vertices = el->n_vertices;
    *out = r.release();
  });
}

namespace mgamd {
namespace {

std::unique_ptr<mg_edge_list> tiles_to_edges(mg_context* ctx, const mg_tiled_matrix* m, bool drop_diagonal) {
  MG_REQUIRE(m->tile_size == 64, "tile -> edge expansion: tile_size must be 64");
  MG_REQUIRE(static_cast<uint64_t>(m->n_strips) * 64 < (1ull << 32), "tile -> edge expansion: too many vertices");
  hipStream_t st = ctx->stream;
  const uint32_t ns = m->n_strips;
  // Pieces per strip, on the host from the row pointer.
  std::vector<uint32_t> off(ns + 1ull), piece_off(ns + 1ull, 0);
  if (ns) MG_HIP(hipMemcpyAsync(off.data(), m->offset_row.get(), sizeof(uint32_t) * (ns + 1ull), hipMemcpyDeviceToHost, st));
  MG_HIP(hipStreamSynchronize(st));
  for (uint32_t x = 0; x < ns; ++x)
    piece_off[x + 1] = piece_off[x] + (off[x + 1] - off[x] + kPieceTiles - 1) / kPieceTiles;
  const uint32_t n_pieces = ns ? piece_off[ns] : 0;
  MG_REQUIRE(static_cast<uint64_t>(n_pieces) * 64 < (1ull << 31), "tile -> edge expansion: too many tiles");
  auto el = std::make_unique<mg_edge_list>();
  el->device = ctx->device;
  el->n_vertices = ns * 64;
  el->n_edges = 0;
  if (n_pieces == 0) {
    el->keys.alloc(1);
    return el;
  }
  DevBuf<uint32_t> d_piece_off(ns + 1ull);
  MG_HIP(hipMemcpyAsync(d_piece_off.get(), piece_off.data(), sizeof(uint32_t) * (ns + 1ull), hipMemcpyHostToDevice, st));
  DevBuf<PieceRef> pieces(n_pieces);
  expand_pieces_kernel<<<grid_for(ns, 256), 256, 0, st>>>(m->offset_row.get(), ns, d_piece_off.get(), pieces.get());
  MG_LAUNCH_CHECK();
  const uint32_t n_slots = n_pieces * 64;
  DevBuf<uint32_t> count(n_slots), start(n_slots + 1ull);
  const unsigned grid = grid_for(n_pieces, 4, 65536);
  count_rows_kernel<<<grid, 256, 0, st>>>(pieces.get(), n_pieces, m->col_idx.get(), m->data.get(),
                                          drop_diagonal ? 1 : 0, count.get());
  MG_LAUNCH_CHECK();
  // The scan below is 32-bit: bound the 64-bit total first (the other edge-list paths stop at
  // 2^31 edges too).
  const uint64_t total64 = device_sum_u32(ctx, count.get(), n_slots);
  MG_REQUIRE(total64 < (1ull << 31), "tile -> edge expansion: " + std::to_string(total64) +
                                         " edges, more than 2^31 - 1");
  exclusive_scan_u32(ctx, count.get(), start.get(), n_slots);
  uint32_t total = 0;
  MG_HIP(hipMemcpyAsync(&total, start.get() + n_slots, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  MG_HIP(hipStreamSynchronize(st));
  el->n_edges = total;
  el->keys.alloc(total ? total : 1);
  if (total) {
    write_rows_kernel<<<grid_for(static_cast<uint64_t>(n_pieces) * 4, 4, 65536), 256, 0, st>>>(
        pieces.get(), n_pieces, m->col_idx.get(), m->data.get(), drop_diagonal ? 1 : 0, start.get(), el->keys.get());
    MG_LAUNCH_CHECK();
  }
  MG_HIP(hipStreamSynchronize(st));
  return el;
}

}  // namespace
}  // namespace mgamd

extern "C" int mg_tiled_matrix_to_edges(mg_context* ctx, const mg_tiled_matrix* m, mg_edge_list** out) {
  return guarded([&] {
    MG_REQUIRE(ctx && m && out, "mg_tiled_matrix_to_edges: NULL argument");
    *out = nullptr;
    DeviceGuard g(ctx->device);
    *out = tiles_to_edges(ctx, m, false).release();
  });
}

extern "C" int mg_edges_grid_block(mg_context* ctx, const mg_edge_list* el, uint32_t n_chunks, uint32_t bx,
                                   uint32_t by, mg_edge_list** out) {
  return guarded([&] {
    MG_REQUIRE(ctx && el && out, "mg_edges_grid_block: NULL argument");
    *out = nullptr;
    MG_REQUIRE(n_chunks >= 1 && n_chunks <= 256, "mg_edges_grid_block: n_chunks must be in [1, 256]");
    MG_REQUIRE(bx < n_chunks && by < n_chunks, "mg_edges_grid_block: block index out of range");
    DeviceGuard g(ctx->device);
    hipStream_t st = ctx->stream;
    // One block scope for the partition and the conversion (the reference's two differ).
    const uint32_t scope = static_cast<uint32_t>((static_cast<uint64_t>(el->n_vertices) + n_chunks - 1) / n_chunks);
    const uint64_t n = el->n_edges;
    MG_REQUIRE(n < (1ull << 31), "mg_edges_grid_block: more than 2^31 edges");
    uint32_t count = 0;
    DevBuf<uint32_t> flag(n ? n : 1), pos(n + 1);
    if (n) {
      grid_block_flag_kernel<<<grid_for(n, 256), 256, 0, st>>>(el->keys.get(), n, scope, bx, by, flag.get());
      MG_LAUNCH_CHECK();
      exclusive_scan_u32(ctx, flag.get(), pos.get(), static_cast<uint32_t>(n));
      MG_HIP(hipMemcpyAsync(&count, pos.get() + n, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
      MG_HIP(hipStreamSynchronize(st));
    }
    auto b = std::make_unique<mg_edge_list>();
    b->device = ctx->device;
    b->n_vertices = scope;
    b->n_edges = count;
    b->keys.alloc(count ? count : 1);
    if (count) {
      grid_block_scatter_kernel<<<grid_for(n, 256), 256, 0, st>>>(el->keys.get(), n, scope, bx, by, flag.get(),
                                                                  pos.get(), b->keys.get());
      MG_LAUNCH_CHECK();
    }
    MG_HIP(hipStreamSynchronize(st));
    *out = b.release();
  });
}

extern "C" int mg_edges_upload(mg_context* ctx, uint32_t n_vertices, uint64_t n_edges, const uint32_t* src,
                               const uint32_t* dst, mg_edge_list** out) {
  return guarded([&] {
    MG_REQUIRE(ctx && out, "mg_edges_upload: NULL argument");
    *out = nullptr;
    MG_REQUIRE(n_edges == 0 || (src && dst), "mg_edges_upload: NULL edge arrays");
    MG_REQUIRE(n_vertices >= 1, "mg_edges_upload: n_vertices == 0");
    for (uint64_t i = 0; i < n_edges; ++i)
      MG_REQUIRE(src[i] < n_vertices && dst[i] < n_vertices,
                 "mg_edges_upload: edge " + std::to_string(i) + " has a vertex id >= n_vertices");
    DeviceGuard g(ctx->device);
    DevBuf<uint32_t> ds(n_edges ? n_edges : 1), dd(n_edges ? n_edges : 1);
    DevBuf<uint64_t> keys(n_edges ? n_edges : 1);
    if (n_edges) {
      MG_HIP(hipMemcpyAsync(ds.get(), src, 4 * n_edges, hipMemcpyHostToDevice, ctx->stream));
      MG_HIP(hipMemcpyAsync(dd.get(), dst, 4 * n_edges, hipMemcpyHostToDevice, ctx->stream));
      pack_edges_kernel<<<grid_for(n_edges, 256), 256, 0, ctx->stream>>>(ds.get(), dd.get(), n_edges, keys.get());
      MG_LAUNCH_CHECK();
    }
    *out = finish_edges(ctx, std::move(keys), n_edges, n_vertices, false).release();
  });
}

extern "C" int mg_edges_info(const mg_edge_list* el, uint32_t* n_vertices, uint64_t* n_edges) {
  return guarded([&] {
    MG_REQUIRE(el != nullptr, "mg_edges_info: NULL edge list");
    if (n_vertices) *n_vertices = el->n_vertices;
    if (n_edges) *n_edges = el->n_edges;
  });
}

extern "C" int mg_edges_download(mg_context* ctx, const mg_edge_list* el, uint32_t* src, uint32_t* dst) {
  return guarded([&] {
    MG_REQUIRE(ctx && el, "mg_edges_download: NULL argument");
    if (el->n_edges == 0) return;
    MG_REQUIRE(src && dst, "mg_edges_download: NULL output arrays");
    DeviceGuard g(ctx->device);
    std::vector<uint64_t> h(el->n_edges);
    MG_HIP(hipMemcpyAsync(h.data(), el->keys.get(), 8 * el->n_edges, hipMemcpyDeviceToHost, ctx->stream));
    MG_HIP(hipStreamSynchronize(ctx->stream));
    for (uint64_t i = 0; i < el->n_edges; ++i) {
      src[i] = static_cast<uint32_t>(h[i] >> 32);
      dst[i] = static_cast<uint32_t>(h[i]);
    }
  });
}

extern "C" int mg_edges_digest(mg_context* ctx, const mg_edge_list* el, uint64_t* digest) {
  return guarded([&] {
    MG_REQUIRE(ctx && el && digest, "mg_edges_digest: NULL argument");
    DeviceGuard g(ctx->device);
    unsigned long long* d = ctx->counters.get();
    MG_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), ctx->stream));
    if (el->n_edges) {
      edge_digest_kernel<<<grid_for(el->n_edges, 256 * 16, 8192), 256, 0, ctx->stream>>>(el->keys.get(),
                                                                                          el->n_edges, d);
      MG_LAUNCH_CHECK();
    }
    unsigned long long h = 0;
    MG_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    MG_HIP(hipStreamSynchronize(ctx->stream));
    *digest = h;
  });
}

extern "C" int mg_edges_to_tiled_matrix(mg_context* ctx, const mg_edge_list* el, int transpose,
                                        mg_tiled_matrix** out) {
  return guarded([&] {
    MG_REQUIRE(ctx && el && out, "mg_edges_to_tiled_matrix: NULL argument");
    *out = nullptr;
    DeviceGuard g(ctx->device);
    hipStream_t st = ctx->stream;
    const uint32_t n_strips = (el->n_vertices + 63) / 64;
    MG_REQUIRE(n_strips < (1u << 26), "mg_edges_to_tiled_matrix: n_strips >= 2^26");
    const uint64_t n = el->n_edges;
    auto m = std::make_unique<mg_tiled_matrix>();
    m->device = ctx->device;
    if (n == 0) {
      alloc_tiled_matrix(m.get(), n_strips, 0);
      MG_HIP(hipMemsetAsync(m->offset_row.get(), 0, sizeof(uint32_t) * (n_strips + 1ull), st));
      MG_HIP(hipStreamSynchronize(st));
      *out = m.release();
      return;
    }
    MG_REQUIRE(n < (1ull << 31), "mg_edges_to_tiled_matrix: more than 2^31 edges");
    DevBuf<uint64_t> keys(n);
    tile_key_kernel<<<grid_for(n, 256), 256, 0, st>>>(el->keys.get(), n, n_strips, transpose, keys.get());
    MG_LAUNCH_CHECK();
    const uint64_t max_key = ((static_cast<uint64_t>(n_strips) * n_strips) << 12);
    sort_unique_u64(ctx, keys, n, bits_for(max_key));  // keys are already unique; this sorts them
    DevBuf<uint32_t> flag(n), incl(n);
    tile_flag_kernel<<<grid_for(n, 256), 256, 0, st>>>(keys.get(), n, flag.get());
    MG_LAUNCH_CHECK();
    {
      size_t tb = 0;
      MG_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb, flag.get(), incl.get(), static_cast<int>(n), st));
      DevBuf<uint8_t> tmp(tb ? tb : 1);
      MG_HIP(hipcub::DeviceScan::InclusiveSum(tmp.get(), tb, flag.get(), incl.get(), static_cast<int>(n), st));
    }
    uint32_t n_tiles = 0;
    MG_HIP(hipMemcpyAsync(&n_tiles, incl.get() + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    MG_HIP(hipStreamSynchronize(st));
    alloc_tiled_matrix(m.get(), n_strips, n_tiles);
    DevBuf<uint32_t> row_cnt(n_strips);
    MG_HIP(hipMemsetAsync(row_cnt.get(), 0, sizeof(uint32_t) * n_strips, st));
    MG_HIP(hipMemsetAsync(m->data.get(), 0, m->data.bytes(), st));
    tile_fill_kernel<<<grid_for(n, 256), 256, 0, st>>>(keys.get(), flag.get(), incl.get(), n, n_strips,
                                                       m->row_idx.get(), m->col_idx.get(), row_cnt.get(),
                                                       reinterpret_cast<unsigned long long*>(m->data.get()));
    MG_LAUNCH_CHECK();
    exclusive_scan_u32(ctx, row_cnt.get(), m->offset_row.get(), n_strips);
    MG_HIP(hipStreamSynchronize(st));
    *out = m.release();
  });
}

extern "C" int mg_edges_destroy(mg_edge_list* el) {
  return guarded([&] {
    if (!el) return;
    DeviceGuard g(el->device);
    delete el;
  });
}
