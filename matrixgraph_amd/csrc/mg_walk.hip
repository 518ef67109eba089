// Multi-hop walks X_{k+1} = X_k * A over bit tiles, right operand supplied as A^T.
//
// Formulation. With A^T in tile-CSR (strip y = 64 target vertices; tile (y,t) word r
// = in-neighbour set of target y*64+r among sources t*64..t*64+63), the hop is the
// tile-row-wise product X_{k+1}^T(y,:) = OR_{t in A^T.row(y)} A^T(y,t) (x) X_k^T(t,:),
// and at bit level row r of A^T(y,t) (x) X^T(t,:) is the OR of rows t' of X^T(t,:)
// for the set bits t' of A^T(y,t)[r]. X_k is kept vertex-major ("strip-interleaved"
// tiles): vertex v's source set is W consecutive u64, so one set bit of A^T costs one
// coalesced W*8-byte gather, OR-ed into the LDS accumulator of strip y.
//
// Panels. A panel is up to 64*W sources of one graph walked together (one walk, or many
// batched queries packed at word boundaries). Every launch covers one hop level of any
// number of panels: chunks and items carry a panel index into a table of per-level
// panel descriptors. Words of queries that have reached their hop count are masked out
// of later gathers (alive masks), so each X_k keeps the finished queries' results.
//
// Per level k (X_k -> X_{k+1}) one kernel does both passes of the hop, tile by tile:
//   structural S_{k+1}[y] = OR_{t in A^T.row(y)} (S_k[t] & alive) per strip, with the
//              tile-product / tile counts of B_alg;
//   numeric    the gather/OR of the set bits, streaming A^T tiles in scheduler chunks.
// Only the non-zero rows of an A^T tile are read: a per-tile row-occupancy mask, built
// once per graph, predicates the 64 row loads of the tile.
// A finalize kernel completes the masks of strips split across chunks.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "mg_device.hpp"
#include "mg_internal.hpp"

namespace mgamd {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;
// Numeric hop kernel: waves per workgroup (all share one strip accumulator in LDS) and the
// minimum waves per SIMD the register allocation must allow.
#ifndef MG_NUMERIC_WAVES
#define MG_NUMERIC_WAVES 4
#endif
#ifndef MG_NUMERIC_MIN_WAVES_PER_SIMD
#define MG_NUMERIC_MIN_WAVES_PER_SIMD 4
#endif
constexpr int kNumWaves = MG_NUMERIC_WAVES;
constexpr int kNumBlock = kNumWaves * kWave;
#ifndef MG_DEFAULT_CHUNK_TILES
#define MG_DEFAULT_CHUNK_TILES 2048
#endif
constexpr uint32_t kDefaultChunkTiles = MG_DEFAULT_CHUNK_TILES;  // scheduler target tiles per chunk
#ifndef MG_CHUNK_STRIPS
#define MG_CHUNK_STRIPS 32
#endif
constexpr uint32_t kChunkStrips = MG_CHUNK_STRIPS;  // at most this many strips per chunk (< 64)
static_assert(kChunkStrips < 64, "one lane per strip of a chunk, plus the end pointer");
#ifndef MG_LOADS_IN_FLIGHT
#define MG_LOADS_IN_FLIGHT 8
#endif
#ifndef MG_TILE_BATCH
#define MG_TILE_BATCH 8
#endif
constexpr int kLoadsInFlight = MG_LOADS_IN_FLIGHT;  // frontier gathers per wave flush
constexpr int kTileBatch = MG_TILE_BATCH;           // A^T tile loads in flight per wave

enum Counter { kTP = 0, kLeft, kRightUsed, kOut, kOutNumeric, kPop, kEdges, kNumCounters };
// Each (panel, level) keeps kCounterSlots copies of its counters; workgroup b adds into slot
// b % kCounterSlots, so the atomics of a launch do not all queue on the same addresses.
// Readers sum the slots.
constexpr uint32_t kCounterSlots = 64;
constexpr uint32_t kCounterStride = kCounterSlots * kNumCounters;  // per (panel, level)
// Block-level partial counters of the numeric kernel, reduced in LDS before one atomic each.
enum BlockCounter { bTP = 0, bUsed, bLeft, bOut, bEdges, bPop, bNumeric, kBlockCounters };

// One panel at one hop level: the launch's inputs (X_k side) and outputs (X_{k+1} side).
struct PanelLevel {
  const uint32_t* off;
  const uint32_t* col;
  const uint32_t* row;      // per A^T tile: its strip (row index)
  const uint32_t* orig;     // per tile of the engine's index: its position in `data` (null: same)
  const uint64_t* data;
  const uint64_t* rmask;    // per A^T tile: bit r = row r non-zero
  const uint64_t* payload;  // per A^T tile: compact bit list (see tile_payload_kernel)
  const uint64_t* dense;    // rows of the dense A^T tiles, 64 words each, at their payload slot
  const uint32_t* csc_tile;     // column index of A^T (first-hop push; null unless built)
  const uint32_t* csc_row;
  const uint64_t* csc_payload;
  const uint64_t* x_in;
  const uint64_t* v_in;
  const uint64_t* s_in;
  uint64_t* x_out;
  uint64_t* n_out;
  uint64_t* v_out;
  uint64_t* z_out;  // per strip: rows of X_{k+1} that may be non-zero in memory
  uint64_t* s_out;
  unsigned long long* counters;
  const uint32_t* jobs;     // per A^T set bit: (source vertex << 6) | target row, tile order
  const uint32_t* job_off;  // per A^T tile: first position in jobs (n_tiles + 1 entries)
  const uint32_t* col_count;  // per tile column of A^T: its tiles (with the job list)
  uint64_t* part;           // segmented hop kernel: per (segment, strip) partial N, V, Z masks
  uint64_t alive_in;    // words of X_k that propagate into X_{k+1}
  uint64_t alive_next;  // words of X_{k+1} that propagate further (they decide V_{k+1})
  uint32_t n_strips;
  uint32_t pad;
};

// Chunk of one panel's A^T tile stream: strips [strip_begin, strip_end) restricted to
// tiles [tile_begin, tile_end). A strip whose tiles exceed one chunk is split in pieces.
// [col_begin, col_end): the tile columns whose per-column tile counts the chunk adds up
// (one-segment job-list kernel); the chunks of a graph cover every column once.
struct Chunk {
  uint32_t panel, strip_begin, strip_end, tile_begin, tile_end;
  uint32_t col_begin, col_end;
};
struct StripItem {
  uint32_t panel, strip;
};
// A strip split across chunks at some level: its rows of that level's output frontier and its
// S word, zeroed at the start of every run (its pieces OR into them).
struct ZeroItem {
  uint64_t* x;
  uint64_t* s;
};
// Source s of a panel: bit `col` of the X_0 row of vertex `vertex`.
struct SourceItem {
  uint32_t panel, col, vertex;
};
struct PanelInit {
  uint64_t* x0;
  uint64_t* s0;
  uint64_t* v0;
  uint64_t* n0;
};
struct FillRange {
  uint64_t* ptr;
  uint64_t n;
};
// Outcome of one scheduler chunk at one hop level (cost-model training data).
struct ChunkLog {
  uint32_t jobs;       // A^T bits whose source vertex was non-empty (gathers)
  uint32_t out_tiles;  // numerically non-empty output tiles of its unsplit strips
};
static_assert(sizeof(ChunkLog) == 8, "written as one 8-byte word");

// --- fills and X_0 ------------------------------------------------------------
// Zero-fill of a list of ranges (a kernel, so it can sit inside a captured graph).
__global__ void __launch_bounds__(kBlock) fill_zero_kernel(const FillRange* __restrict__ ranges, uint32_t n_ranges) {
  for (uint32_t r = 0; r < n_ranges; ++r) {
    uint64_t* p = ranges[r].ptr;
    const uint64_t n = ranges[r].n;
    const uint64_t n2 = n / 2;
    ulonglong2* p2 = reinterpret_cast<ulonglong2*>(p);
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n2;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
      p2[i] = make_ulonglong2(0, 0);
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) p[n - 1] = 0;
  }
}

// Before a run: zero the rows of X_0 that the previous run set (the rows marked in V_0 of the
// strips whose numeric mask is non-zero: the source rows) and reset the X_0 masks. X_0
// buffers start zeroed at allocation. One wave per strip, over the strips of all panels in
// turn: every lane reads the strip's masks before lane 0 resets them, so no barrier is needed.
template <int W>
__global__ void __launch_bounds__(kBlock) clear_x0_kernel(const PanelInit* __restrict__ panels,
                                                          const uint32_t* __restrict__ n_strips, uint32_t n_panels) {
  constexpr int kWords = 64 * W;
  const uint32_t lane = lane_id();
  const uint32_t n_waves = gridDim.x * kWaves;
  const uint32_t first = read_first_lane(blockIdx.x * kWaves + threadIdx.x / kWave);
  for (uint32_t p = 0; p < n_panels; ++p) {
    const PanelInit P = panels[p];
    const uint32_t ns = n_strips[p];
    for (uint32_t y = first; y < ns; y += n_waves) {
      if (P.n0[y] != 0) {
        // Only the rows that hold a source (V_0 marks them) are non-zero.
        uint64_t* dst = P.x0 + static_cast<uint64_t>(y) * kWords;
        for (uint64_t rows = P.v0[y]; rows; rows &= rows - 1) {
          const uint32_t r = static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(rows)) - 1);
          for (int i = lane; i < W; i += kWave) dst[r * W + i] = 0;
        }
      }
      if (lane == 0) {
        P.n0[y] = 0;
        P.s0[y] = 0;
        P.v0[y] = 0;
      }
    }
  }
}

__global__ void init_sources_kernel(const SourceItem* __restrict__ items, uint32_t n_items,
                                    const PanelInit* __restrict__ panels, uint32_t W) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_items; i += gridDim.x * blockDim.x) {
    const SourceItem it = items[i];
    const PanelInit p = panels[it.panel];
    const uint32_t v = it.vertex;
    atomicOr(reinterpret_cast<unsigned long long*>(p.x0) + static_cast<uint64_t>(v) * W + it.col / 64,
             1ull << (it.col % 64));
    atomicOr(reinterpret_cast<unsigned long long*>(p.s0) + v / 64, 1ull << (it.col / 64));
    atomicOr(reinterpret_cast<unsigned long long*>(p.n0) + v / 64, 1ull << (it.col / 64));
    atomicOr(reinterpret_cast<unsigned long long*>(p.v0) + v / 64, 1ull << (v % 64));
  }
}

// Row-occupancy mask of every A^T tile: bit r set when word r (target row r) is non-zero.
// Tile i of the engine's index is tile orig[i] of the matrix (orig null: the same tile).
__device__ inline uint64_t data_tile(const uint32_t* __restrict__ orig, uint32_t i) {
  return static_cast<uint64_t>(orig ? orig[i] : i) * kTile;
}

__global__ void __launch_bounds__(kBlock) row_mask_kernel(const uint64_t* __restrict__ data,
                                                          const uint32_t* __restrict__ orig, uint32_t n_tiles,
                                                          uint64_t* __restrict__ mask) {
  const uint32_t lane = lane_id();
  const uint32_t n_waves = gridDim.x * kWaves;
  for (uint32_t i = read_first_lane(blockIdx.x * kWaves + threadIdx.x / kWave); i < n_tiles; i += n_waves) {
    const uint64_t m = ballot(__builtin_nontemporal_load(data + data_tile(orig, i) + lane) != 0);
    if (lane == 0) mask[i] = m;
  }
}

// Compact payload of every A^T tile (tile_payload). A dense tile's payload also receives,
// once dense_compact_kernel has run, the index of its copy in the dense-row array in bits
// 13..44; it is read row by row through its row-occupancy mask.

__global__ void __launch_bounds__(kBlock) tile_payload_kernel(const uint64_t* __restrict__ data,
                                                              const uint32_t* __restrict__ orig, uint32_t n_tiles,
                                                              uint64_t* __restrict__ payload) {
  const uint32_t lane = lane_id();
  const uint32_t n_waves = gridDim.x * kWaves;
  for (uint32_t i = read_first_lane(blockIdx.x * kWaves + threadIdx.x / kWave); i < n_tiles; i += n_waves) {
    const uint64_t word = __builtin_nontemporal_load(data + data_tile(orig, i) + lane);
    const uint64_t pl = tile_payload(word);
    if (lane == 0) payload[i] = pl;
  }
}

__global__ void __launch_bounds__(kBlock) dense_flag_kernel(const uint64_t* __restrict__ payload, uint32_t n,
                                                            uint32_t* __restrict__ flag) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    flag[i] = (payload[i] & kPayloadDense) ? 1u : 0u;
}

// Copies the rows of every dense tile to its slot pos[i] of a compact array (64 words per
// tile) and records the slot in the payload: the hop kernels then read dense rows from a
// few hundred MB instead of the whole tile store.
__global__ void __launch_bounds__(kBlock) dense_compact_kernel(const uint64_t* __restrict__ data,
                                                               const uint32_t* __restrict__ orig, uint32_t n,
                                                               const uint32_t* __restrict__ pos,
                                                               uint64_t* __restrict__ payload,
                                                               uint64_t* __restrict__ dense) {
  const uint32_t lane = lane_id();
  const uint32_t n_waves = gridDim.x * kWaves;
  for (uint32_t i = read_first_lane(blockIdx.x * kWaves + threadIdx.x / kWave); i < n; i += n_waves) {
    const uint64_t pl = payload[i];
    if (!(pl & kPayloadDense)) continue;
    const uint64_t slot = pos[i];
    dense[slot * kTile + lane] = __builtin_nontemporal_load(data + data_tile(orig, i) + lane);
    if (lane == 0) payload[i] = (pl & ~(0xFFFFFFFFull << 13)) | (slot << 13);
  }
}

// --- numeric ----------------------------------------------------------------
// Gather queue of one wave: lane j holds job j (source vertex u and the mask of the
// strip's target rows that have u as in-neighbour); len is wave-uniform.
struct GatherQueue {
  uint32_t u;
  uint64_t rows;
  uint32_t len;
};

// A job gathers the W words of one source vertex: 64/W jobs share one 64-lane load,
// and a flush keeps at most kLoadsInFlight loads in flight.
template <int W>
struct QueueShape {
  static constexpr int kJobsPerLoad = 64 / W;
  static constexpr int kCapacity = (kLoadsInFlight * kJobsPerLoad < 64) ? kLoadsInFlight * kJobsPerLoad : 64;
  static constexpr int kLoads = (kCapacity + kJobsPerLoad - 1) / kJobsPerLoad;
};

// keep = alive mask of this lane's word (all ones or zero).
template <int W>
__device__ inline void flush_queue(GatherQueue& q, gptr<const uint64_t> __restrict__ x_in,
                                   unsigned long long* __restrict__ acc, uint32_t lane, uint64_t keep) {
  using S = QueueShape<W>;
  uint64_t v[S::kLoads];
  if constexpr (W == 64) {
    // One job per load: its source vertex is wave-uniform.
#pragma unroll
    for (int j = 0; j < S::kLoads; ++j) {
      v[j] = 0;
      if (static_cast<uint32_t>(j) < q.len) {
        const uint32_t u = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(q.u), j));
        v[j] = x_in[static_cast<uint64_t>(u) * W + lane] & keep;
      }
    }
#pragma unroll
    for (int j = 0; j < S::kLoads; ++j) {
      if (static_cast<uint32_t>(j) < q.len) {
        uint64_t rows = read_lane(q.rows, j);
        const uint64_t val = v[j];
        while (rows) {
          const int r = __ffsll(static_cast<unsigned long long>(rows)) - 1;
          rows &= rows - 1;
          if (val) atomicOr(acc + r * W + lane, static_cast<unsigned long long>(val));
        }
      }
    }
  } else {
    // Lane = (job slot, word): several jobs per load, per-lane source index.
    const uint32_t w = lane % W;
    const uint32_t slot = lane / W;
#pragma unroll
    for (int j = 0; j < S::kLoads; ++j) {
      const uint32_t job = j * S::kJobsPerLoad + slot;
      const uint32_t u = static_cast<uint32_t>(__shfl(static_cast<int>(q.u), static_cast<int>(job & 63), kWave));
      v[j] = (job < q.len) ? (x_in[static_cast<uint64_t>(u) * W + w] & keep) : 0;
    }
#pragma unroll
    for (int j = 0; j < S::kLoads; ++j) {
      const uint32_t job = j * S::kJobsPerLoad + slot;
      const uint32_t lo = static_cast<uint32_t>(__shfl(static_cast<int>(static_cast<uint32_t>(q.rows)),
                                                       static_cast<int>(job & 63), kWave));
      const uint32_t hi = static_cast<uint32_t>(__shfl(static_cast<int>(static_cast<uint32_t>(q.rows >> 32)),
                                                       static_cast<int>(job & 63), kWave));
      uint64_t rows = (job < q.len && v[j] != 0) ? ((static_cast<uint64_t>(hi) << 32) | lo) : 0;
      const uint64_t val = v[j];
      while (rows) {
        const int r = __ffsll(static_cast<unsigned long long>(rows)) - 1;
        rows &= rows - 1;
        atomicOr(acc + r * W + w, static_cast<unsigned long long>(val));
      }
    }
  }
  q.len = 0;
}

template <int W>
__device__ inline void push_job(GatherQueue& q, uint32_t u, uint64_t rows, uint32_t lane,
                                gptr<const uint64_t> __restrict__ x_in, unsigned long long* __restrict__ acc,
                                uint64_t keep) {
  if (lane == q.len) {
    q.u = u;
    q.rows = rows;
  }
  if (++q.len == QueueShape<W>::kCapacity) flush_queue<W>(q, x_in, acc, lane, keep);
}

// Extract the gather jobs of one A^T tile: for each source column t' with a set bit
// (restricted to non-empty sources vm), job (t*64+t', rows with bit t').
template <int W>
__device__ inline void tile_jobs(uint32_t t, uint64_t b, uint64_t vm, GatherQueue& q, uint32_t lane,
                                 gptr<const uint64_t> __restrict__ x_in, unsigned long long* __restrict__ acc,
                                 uint64_t keep, uint32_t& edges) {
  b &= vm;
  const uint64_t nz = ballot(b != 0);
  if (nz == 0) return;
  edges += __popcll(b);
  uint64_t cols;
  if (__popcll(nz) <= 6) {
    cols = 0;
    uint64_t m = nz;
    while (m) {
      const int r = __ffsll(static_cast<unsigned long long>(m)) - 1;
      m &= m - 1;
      cols |= read_lane(b, r);
    }
  } else {
    cols = read_first_lane64(wave_or(b));
  }
  while (cols) {
    const int tp = __ffsll(static_cast<unsigned long long>(cols)) - 1;
    cols &= cols - 1;
    const uint64_t rows = ballot((b >> tp) & 1);
    push_job<W>(q, t * 64u + static_cast<uint32_t>(tp), rows, lane, x_in, acc, keep);
  }
}

// Block-wide sum of per-thread counters into LDS (one atomic per wave), then one global
// atomic per non-zero counter.
__device__ inline void flush_block_counters(const uint32_t (&c)[kBlockCounters], unsigned long long* s_cnt,
                                            gptr<unsigned long long> __restrict__ counters,
                                            ChunkLog* __restrict__ log) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int i = 0; i < kBlockCounters; ++i) {
    const uint64_t v = wave_sum(c[i]);
    if (lane == 0 && v) atomicAdd(&s_cnt[i], static_cast<unsigned long long>(v));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    constexpr int map[kBlockCounters] = {kTP, kRightUsed, kLeft, kOut, kEdges, kPop, kOutNumeric};
    const gptr<unsigned long long> slot = counters + (blockIdx.x % kCounterSlots) * kNumCounters;
#pragma unroll
    for (int i = 0; i < kBlockCounters; ++i)
      if (s_cnt[i]) atomicAdd((unsigned long long*)(slot + map[i]), s_cnt[i]);
    if (log) {
      // ChunkLog {jobs, out_tiles} as one 8-byte store.
      const uint64_t rec = (s_cnt[bEdges] & 0xFFFFFFFFull) | ((s_cnt[bNumeric] & 0xFFFFFFFFull) << 32);
      as_global(reinterpret_cast<uint64_t*>(log))[blockIdx.x] = rec;
    }
  }
}

// Strip write-out shared by the hop kernels. An unsplit strip reduces its masks (N: non-zero
// words, V: non-empty vertices that propagate, Z: non-zero rows) and popcount, records S and
// its tile counts, and is stored with non-temporal stores. Memory can only be non-zero at
// the rows x words of the previous Z and N masks, so a zero word outside them is not
// written. A piece of a split strip ORs into memory that was zeroed before the launch.
// z_prev, n_prev: the strip's Z and N masks before this launch; left: S_k[y] & alive.
template <int W, int BLOCK>
__device__ inline void write_strip(const unsigned long long* __restrict__ acc, unsigned long long& s_n,
                                   unsigned long long& s_v, unsigned long long& s_z, unsigned long long sym,
                                   uint32_t y, bool split, gptr<uint64_t> x_out, gptr<uint64_t> n_out,
                                   gptr<uint64_t> v_out, gptr<uint64_t> z_out, gptr<uint64_t> s_out,
                                   uint64_t z_prev, uint64_t n_prev, uint64_t left, uint64_t alive_next,
                                   uint32_t (&cnt)[kBlockCounters]) {
  constexpr int kWords = 64 * W;
  const uint32_t lane = lane_id();
  const gptr<uint64_t> dst = x_out + static_cast<uint64_t>(y) * kWords;
  if (!split) {
    // One pass: masks and popcount reduced while the words are stored.
    uint64_t nmask = 0, vmask = 0, zmask = 0, pc = 0;
    for (int i = threadIdx.x; i < kWords; i += BLOCK) {
      const uint64_t val = acc[i];
      if (val) {
        nmask |= 1ull << (i % W);
        zmask |= 1ull << (i / W);
        if ((alive_next >> (i % W)) & 1) vmask |= 1ull << (i / W);
        pc += __popcll(val);
      }
      if (val != 0 || (((z_prev >> (i / W)) & 1) && ((n_prev >> (i % W)) & 1)))
        __builtin_nontemporal_store(val, dst + i);
    }
    nmask = wave_or(nmask);
    vmask = wave_or(vmask);
    zmask = wave_or(zmask);
    cnt[bPop] += static_cast<uint32_t>(pc);
    if (lane == 0) {
      if (nmask) atomicOr(&s_n, static_cast<unsigned long long>(nmask));
      if (vmask) atomicOr(&s_v, static_cast<unsigned long long>(vmask));
      if (zmask) atomicOr(&s_z, static_cast<unsigned long long>(zmask));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      n_out[y] = s_n;
      v_out[y] = s_v;
      z_out[y] = s_z;
      s_out[y] = sym;
      cnt[bNumeric] += __popcll(s_n);
      cnt[bOut] += __popcll(sym);
      cnt[bLeft] += __popcll(left);
    }
  } else {
    for (int i = threadIdx.x; i < kWords; i += BLOCK) {
      const uint64_t val = acc[i];
      if (val) atomicOr((unsigned long long*)(dst + i), static_cast<unsigned long long>(val));
    }
    if (threadIdx.x == 0 && sym) atomicOr((unsigned long long*)(s_out + y), sym);
  }
}

// One workgroup per scheduler chunk; per strip a 64 x W-word LDS accumulator. The same
// sweep over the strip's tiles computes the structural output mask (OR of the source
// strips' structural masks) and the tile-product counts.
template <int W>
__global__ void __launch_bounds__(kNumBlock, MG_NUMERIC_MIN_WAVES_PER_SIMD) walk_numeric_kernel(const Chunk* __restrict__ chunks,
                                                              const PanelLevel* __restrict__ lv,
                                                              ChunkLog* __restrict__ log) {
  constexpr int kWords = 64 * W;  // accumulator words per strip
  __shared__ unsigned long long acc[kWords];
  __shared__ unsigned long long s_n, s_v, s_z, s_sym;
  __shared__ unsigned long long s_cnt[kBlockCounters];
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / kWave;
  const Chunk ch = chunks[blockIdx.x];
  const PanelLevel& L = lv[ch.panel];
  const gptr<const uint32_t> __restrict__ off = as_global(L.off);
  const gptr<const uint32_t> __restrict__ col = as_global(L.col);
  const gptr<const uint64_t> __restrict__ data = as_global(L.data);
  const uint32_t* const orig = L.orig;
  const gptr<const uint64_t> __restrict__ rmask = as_global(L.rmask);
  const gptr<const uint64_t> __restrict__ x_in = as_global(L.x_in);
  const gptr<const uint64_t> __restrict__ v_in = as_global(L.v_in);
  const gptr<const uint64_t> __restrict__ s_in = as_global(L.s_in);
  const gptr<uint64_t> __restrict__ x_out = as_global(L.x_out);
  const gptr<uint64_t> __restrict__ n_out = as_global(L.n_out);
  const gptr<uint64_t> __restrict__ v_out = as_global(L.v_out);
  const gptr<uint64_t> __restrict__ s_out = as_global(L.s_out);
  const uint64_t alive_in = L.alive_in;
  const uint64_t keep = ((alive_in >> ((W == 64) ? lane : (lane % W))) & 1) ? ~0ull : 0ull;
  const uint64_t alive_next = L.alive_next;
  uint32_t cnt[kBlockCounters];  // per-thread partial counts (each fits 32 bits per chunk)
#pragma unroll
  for (int i = 0; i < kBlockCounters; ++i) cnt[i] = 0;
  if (threadIdx.x < kBlockCounters) s_cnt[threadIdx.x] = 0;
  __syncthreads();

  for (uint32_t y = ch.strip_begin; y < ch.strip_end; ++y) {
    const uint32_t r0 = off[y], r1 = off[y + 1];
    const uint32_t seg0 = r0 > ch.tile_begin ? r0 : ch.tile_begin;
    const uint32_t seg1 = r1 < ch.tile_end ? r1 : ch.tile_end;
    const bool split = (seg0 != r0) || (seg1 != r1);
    for (int i = threadIdx.x; i < kWords; i += kNumBlock) acc[i] = 0;
    if (threadIdx.x == 0) { s_n = 0; s_v = 0; s_z = 0; s_sym = 0; }
    __syncthreads();

    GatherQueue q{0, 0, 0};
    uint64_t sym = 0;
    // Wave `wave` scans tiles p = seg0 + wave + kNumWaves*(64*g + lane): one coalesced load of
    // 64 tile columns and row masks and one gather of their source-strip masks per 64
    // tiles; the active tiles then go in groups of kTileBatch whose row loads are in
    // flight together.
    for (uint32_t g0 = seg0 + wave; g0 < seg1; g0 += kNumWaves * 64) {
      const uint32_t p = g0 + kNumWaves * lane;
      const bool valid = p < seg1;
      const uint32_t t_l = valid ? col[p] : 0;
      const uint32_t o_l = valid ? (orig ? as_global(orig)[p] : p) : 0;
      const uint64_t rm_l = valid ? rmask[p] : 0;
      const uint64_t vm_l = valid ? v_in[t_l] : 0;
      const uint64_t sm_l = valid ? (s_in[t_l] & alive_in) : 0;
      sym |= sm_l;
      cnt[bTP] += __popcll(sm_l);
      cnt[bUsed] += (sm_l != 0);
      uint64_t act = ballot(vm_l != 0 && rm_l != 0);
      while (act) {
        int ks[kTileBatch];
        int cnt_b = 0;
#pragma unroll
        for (int j = 0; j < kTileBatch; ++j) {
          ks[j] = 0;
          if (act) {
            ks[j] = __ffsll(static_cast<unsigned long long>(act)) - 1;
            act &= act - 1;
            cnt_b = j + 1;
          }
        }
        // A^T tiles are streamed once per hop: non-temporal loads keep them from
        // evicting the frontier words that the gathers re-read. Only rows set in the
        // tile's row mask are read.
        uint64_t b[kTileBatch];
#pragma unroll
        for (int j = 0; j < kTileBatch; ++j) {
          b[j] = 0;
          if (j < cnt_b) {
            const uint64_t rm = read_lane(rm_l, ks[j]);
            const uint32_t o = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(o_l), ks[j]));
            if ((rm >> lane) & 1) b[j] = __builtin_nontemporal_load(data + static_cast<uint64_t>(o) * kTile + lane);
          }
        }
#pragma unroll
        for (int j = 0; j < kTileBatch; ++j) {
          if (j < cnt_b) {
            const uint32_t t = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(t_l), ks[j]));
            const uint64_t vm = read_lane(vm_l, ks[j]);
            tile_jobs<W>(t, b[j], vm, q, lane, x_in, acc, keep, cnt[bEdges]);
          }
        }
      }
    }
    if (q.len) flush_queue<W>(q, x_in, acc, lane, keep);
    sym = wave_or(sym);
    if (lane == 0 && sym) atomicOr(&s_sym, static_cast<unsigned long long>(sym));
    __syncthreads();

    write_strip<W, kNumBlock>(acc, s_n, s_v, s_z, s_sym, y, split, x_out, n_out, v_out, as_global(L.z_out), s_out,
                              as_global(L.z_out)[y], n_out[y], s_in[y] & alive_in, alive_next, cnt);
    __syncthreads();
  }
  flush_block_counters(cnt, s_cnt, as_global(L.counters), log);
}

// --- numeric, 64 words per vertex ---------------------------------------------
// For the full 4096-source panel the hop kernel splits tile decoding from gathering. Lane i
// of a wave decodes tile i of a group of 64 from its compact payload (tiles with at most
// kPayloadMax bits) and appends one job (source vertex, target row) per set bit to the
// wave's job list in LDS; a denser tile is loaded by the whole wave (lane = row, only the
// rows in its row mask) and emits its jobs the same way. Draining the list, each half-wave
// gathers one 512-B source row (two 8-B loads per lane) and ORs it into the target row.
//
// A job is stored pre-scaled: source vertex * 512 (the byte offset of its frontier row) plus
// target row * 8, so the drain needs no index arithmetic beyond two masks. The list is
// padded to whole drain steps with jobs that read the frontier's extra all-zero row (vertex
// n_vertices, allocated past the last strip and never written), so no slot is tested.
#ifndef MG_DENSE_BATCH
#define MG_DENSE_BATCH 2
#endif
constexpr int kDenseBatch = MG_DENSE_BATCH;  // dense A^T tiles whose rows a wave loads together
constexpr int kJobCap = 228;     // jobs per wave list (a multiple of 2 * kDrainUnroll)
constexpr int kDrainUnroll = 3;  // job pairs per drain step (two steps in flight)
static_assert(kJobCap % (2 * kDrainUnroll) == 0, "job lists are drained in whole steps");

__device__ inline uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ inline unsigned long long make_job(uint64_t vertex, uint32_t row) {
  return (vertex << 9) | (static_cast<uint64_t>(row) << 3);
}

struct DrainStep {
  uint64_t v0[kDrainUnroll], v1[kDrainUnroll];  // words c and c + 32 of the source row
  uint32_t lds[kDrainUnroll];                    // LDS byte offset of the target words
};

// xl: frontier base + 8 * c (byte pointer); jobs[j .. j + 2 * kDrainUnroll) are valid slots.
__device__ inline void drain_issue(DrainStep& d, const unsigned long long* __restrict__ jobs_h, uint32_t j,
                                   gptr<const char> __restrict__ xl, uint32_t c8) {
#pragma unroll
  for (int k = 0; k < kDrainUnroll; ++k) {
    const unsigned long long e = jobs_h[j + 2 * k];
    const gptr<const char> src = xl + (e & ~0x1FFull);
    d.v0[k] = *reinterpret_cast<gptr<const uint64_t>>(src);
    d.v1[k] = *reinterpret_cast<gptr<const uint64_t>>(src + 256);
    d.lds[k] = ((static_cast<uint32_t>(e) & 0x1F8u) << 6) | c8;
  }
}

template <bool kMasked>
__device__ inline void drain_commit(const DrainStep& d, char* __restrict__ acc_b, uint64_t keep0, uint64_t keep1) {
#pragma unroll
  for (int k = 0; k < kDrainUnroll; ++k) {
    unsigned long long* a = reinterpret_cast<unsigned long long*>(acc_b + d.lds[k]);
    atomicOr(a, kMasked ? (d.v0[k] & keep0) : d.v0[k]);
    atomicOr(a + 32, kMasked ? (d.v1[k] & keep1) : d.v1[k]);
  }
}

template <bool kMasked>
__device__ inline void drain_loop(const unsigned long long* __restrict__ jobs_h, uint32_t n, gptr<const char> xl,
                                  char* __restrict__ acc_b, uint32_t c8, uint64_t keep0, uint64_t keep1) {
  constexpr uint32_t kStep = 2 * kDrainUnroll;
  // Ping-pong between two fixed register sets (no copies, so no wait before the next issue).
  DrainStep a, b;
  drain_issue(a, jobs_h, 0, xl, c8);
  for (uint32_t j = kStep;; j += 2 * kStep) {
    if (j >= n) {
      drain_commit<kMasked>(a, acc_b, keep0, keep1);
      break;
    }
    drain_issue(b, jobs_h, j, xl, c8);
    drain_commit<kMasked>(a, acc_b, keep0, keep1);
    if (j + kStep >= n) {
      drain_commit<kMasked>(b, acc_b, keep0, keep1);
      break;
    }
    drain_issue(a, jobs_h, j + kStep, xl, c8);
    drain_commit<kMasked>(b, acc_b, keep0, keep1);
  }
}

// Drains the first n jobs of this wave's list (n > 0). Lane c of a half-wave handles words
// c and c + 32 of a 64-word row, so each LDS OR covers 32 consecutive words of two rows.
// Words of finished queries (keep masks) are masked only in panels that have any.
__device__ inline void drain_jobs(unsigned long long* __restrict__ jobs, uint32_t n, gptr<const uint64_t> x_in,
                                  unsigned long long* __restrict__ acc, uint32_t lane, uint64_t keep0, uint64_t keep1,
                                  bool all_alive, unsigned long long pad_job) {
  constexpr uint32_t kStep = 2 * kDrainUnroll;
  const uint32_t padded = (n + kStep - 1) / kStep * kStep;
  if (lane < padded - n) jobs[n + lane] = pad_job;
  wave_lds_sync();
  const uint32_t c = lane & 31;
  const gptr<const char> xl = reinterpret_cast<gptr<const char>>(x_in) + 8 * c;
  const unsigned long long* jobs_h = jobs + (lane >> 5);
  char* const acc_b = reinterpret_cast<char*>(acc);
  if (all_alive)
    drain_loop<false>(jobs_h, padded, xl, acc_b, 8 * c, keep0, keep1);
  else
    drain_loop<true>(jobs_h, padded, xl, acc_b, 8 * c, keep0, keep1);
}

// Appends make_job(tag_vertex + lowest set bit of `bits`, row) for every lane whose `bits`
// is non-zero, clearing that bit, until no lane has bits left; drains the list whenever it
// could overflow.
__device__ inline void emit_jobs(uint64_t bits, uint64_t base_vertex, uint32_t row,
                                 unsigned long long* __restrict__ jobs, uint32_t& len,
                                 gptr<const uint64_t> __restrict__ x_in, unsigned long long* __restrict__ acc,
                                 uint32_t lane, uint64_t keep0, uint64_t keep1, bool all_alive,
                                 unsigned long long pad_job) {
  for (uint64_t has = ballot(bits != 0); has; has = ballot(bits != 0)) {
    if (len + 64 > kJobCap) {
      drain_jobs(jobs, len, x_in, acc, lane, keep0, keep1, all_alive, pad_job);
      wave_lds_sync();
      len = 0;
    }
    if (bits) {
      jobs[len + lanes_below(has)] =
          make_job(base_vertex + static_cast<uint64_t>(__ffsll(static_cast<unsigned long long>(bits)) - 1), row);
      bits &= bits - 1;
    }
    len += __popcll(has);
  }
}

// State a wave carries through the tiles of one strip.
struct WaveTileState {
  unsigned long long* jobs;
  uint32_t len;  // jobs in the list (wave-uniform)
  uint64_t sym;  // structural OR of the source strips (per lane)
};

// Decodes one group of 64 tiles (lane = tile) into jobs.
__device__ inline void decode_group(WaveTileState& st, uint32_t p, bool valid, uint32_t t_l, uint64_t pl,
                                    uint64_t rm_l, uint64_t vm_l, uint64_t sm_l, gptr<const uint64_t> __restrict__ dense_rows,
                                    gptr<const uint64_t> __restrict__ x_in,
                                    unsigned long long* __restrict__ acc, uint32_t lane, uint64_t keep0,
                                    uint64_t keep1, bool all_alive, unsigned long long pad_job,
                                    uint32_t (&cnt)[kBlockCounters]) {
  st.sym |= sm_l;
  cnt[bTP] += __popcll(sm_l);
  cnt[bUsed] += (sm_l != 0);
  const bool act = valid && vm_l != 0 && pl != 0;
  const bool dense = act && (pl & kPayloadDense) != 0;
  // Sparse tiles: entries of this lane's payload whose source vertex is non-empty.
  const uint64_t ent = (act && !dense) ? pl : 0ull;
  const uint32_t n_ent = static_cast<uint32_t>(ent >> 48) & 7u;
#pragma unroll
  for (uint32_t i = 0; i < kPayloadMax; ++i) {
    const uint32_t e = static_cast<uint32_t>(ent >> (12 * i)) & 0xFFFu;
    const bool ok = i < n_ent && ((vm_l >> (e & 63u)) & 1);
    cnt[bEdges] += ok;
    const uint64_t has = ballot(ok);
    if (has) {
      if (st.len + 64 > kJobCap) {
        drain_jobs(st.jobs, st.len, x_in, acc, lane, keep0, keep1, all_alive, pad_job);
        wave_lds_sync();
        st.len = 0;
      }
      if (ok) st.jobs[st.len + lanes_below(has)] = make_job(static_cast<uint64_t>(t_l) * 64 + (e & 63u), e >> 6);
      st.len += __popcll(has);
    }
  }
  // Dense tiles: the whole wave loads them (lane = row, only rows in the tile's row mask, which
  // rm_l holds per lane), kDenseBatch tiles at a time, and emits their jobs.
  for (uint64_t dm = ballot(dense); dm;) {
    uint64_t bits[kDenseBatch];
    uint32_t tb[kDenseBatch];
#pragma unroll
    for (int q = 0; q < kDenseBatch; ++q) {
      const bool any = dm != 0;
      const int k = any ? __ffsll(static_cast<unsigned long long>(dm)) - 1 : 0;
      dm &= dm - 1;
      const uint64_t slot = (read_lane(pl, k) >> 13) & 0xFFFFFFFFull;
      tb[q] = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(t_l), k));
      const uint64_t vm = read_lane(vm_l, k);
      const uint64_t rm = any ? read_lane(rm_l, k) : 0ull;
      bits[q] = ((rm >> lane) & 1) ? (__builtin_nontemporal_load(dense_rows + slot * kTile + lane) & vm) : 0ull;
      cnt[bEdges] += __popcll(bits[q]);
    }
#pragma unroll
    for (int q = 0; q < kDenseBatch; ++q)
      emit_jobs(bits[q], static_cast<uint64_t>(tb[q]) * 64, lane, st.jobs, st.len, x_in, acc, lane, keep0, keep1,
                all_alive, pad_job);
  }
}

__global__ void __launch_bounds__(kNumBlock, MG_NUMERIC_MIN_WAVES_PER_SIMD)
    walk_numeric64_kernel(const Chunk* __restrict__ chunks, const PanelLevel* __restrict__ lv,
                          ChunkLog* __restrict__ log) {
  constexpr int W = 64;
  constexpr int kWords = 64 * W;
  __shared__ unsigned long long acc[kWords];
  __shared__ unsigned long long jobs_all[kNumWaves][kJobCap];
  __shared__ unsigned long long s_n, s_v, s_z, s_sym;
  __shared__ unsigned long long s_cnt[kBlockCounters];
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / kWave;
  const Chunk ch = chunks[blockIdx.x];
  const PanelLevel& L = lv[ch.panel];
  const gptr<const uint32_t> __restrict__ off = as_global(L.off);
  const gptr<const uint32_t> __restrict__ col = as_global(L.col);
  const gptr<const uint64_t> __restrict__ dense_rows = as_global(L.dense);
  const gptr<const uint64_t> __restrict__ rmask = as_global(L.rmask);
  const gptr<const uint64_t> __restrict__ payload = as_global(L.payload);
  const gptr<const uint64_t> __restrict__ x_in = as_global(L.x_in);
  const gptr<const uint64_t> __restrict__ v_in = as_global(L.v_in);
  const gptr<const uint64_t> __restrict__ s_in = as_global(L.s_in);
  const uint64_t alive_in = L.alive_in;
  const uint64_t alive_next = L.alive_next;
  const uint64_t keep0 = ((alive_in >> (lane & 31)) & 1) ? ~0ull : 0ull;
  const uint64_t keep1 = ((alive_in >> ((lane & 31) + 32)) & 1) ? ~0ull : 0ull;
  const bool all_alive = alive_in == ~0ull;
  const unsigned long long pad_job = make_job(static_cast<uint64_t>(L.n_strips) * 64, 0);  // the zero row
  // Per-strip values of the chunk (at most kChunkStrips strips), one strip per lane, loaded
  // once: row pointers, previous Z and N masks, and S_k & alive for the tile counts.
  const uint32_t n_cs = ch.strip_end - ch.strip_begin;
  const uint32_t ys = ch.strip_begin + lane;
  const uint32_t off_l = lane <= n_cs ? off[ys] : 0u;
  const uint64_t zp_l = lane < n_cs ? as_global(L.z_out)[ys] : 0ull;
  const uint64_t np_l = lane < n_cs ? as_global(L.n_out)[ys] : 0ull;
  const uint64_t left_l = lane < n_cs ? (s_in[ys] & alive_in) : 0ull;
  uint32_t cnt[kBlockCounters];
#pragma unroll
  for (int i = 0; i < kBlockCounters; ++i) cnt[i] = 0;
  if (threadIdx.x < kBlockCounters) s_cnt[threadIdx.x] = 0;
  for (int i = threadIdx.x; i < kWords; i += kNumBlock) acc[i] = 0;
  if (threadIdx.x == 0) { s_n = 0; s_v = 0; s_z = 0; s_sym = 0; }
  __syncthreads();

  WaveTileState st{jobs_all[wave], 0u, 0ull};
  for (uint32_t i = 0; i < n_cs; ++i) {
    const uint32_t y = ch.strip_begin + i;
    const uint32_t r0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(off_l), static_cast<int>(i)));
    const uint32_t r1 =
        static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(off_l), static_cast<int>(i + 1)));
    const uint32_t seg0 = r0 > ch.tile_begin ? r0 : ch.tile_begin;
    const uint32_t seg1 = r1 < ch.tile_end ? r1 : ch.tile_end;
    const bool split = (seg0 != r0) || (seg1 != r1);
    st.len = 0;
    st.sym = 0;
    // Two groups of 64 tiles per step, their index and mask loads issued together. The
    // tile index and payload streams are read once per hop.
    for (uint32_t g0 = seg0 + 64 * wave; g0 < seg1; g0 += 128 * kNumWaves) {
      const uint32_t pa = g0 + lane, pb = g0 + 64 * kNumWaves + lane;
      const bool va = pa < seg1, vb = pb < seg1;
      const uint32_t ta = va ? __builtin_nontemporal_load(col + pa) : 0u;
      const uint32_t tb = vb ? __builtin_nontemporal_load(col + pb) : 0u;
      const uint64_t pla = va ? __builtin_nontemporal_load(payload + pa) : 0ull;
      const uint64_t plb = vb ? __builtin_nontemporal_load(payload + pb) : 0ull;
      const uint64_t vma = va ? v_in[ta] : 0ull, vmb = vb ? v_in[tb] : 0ull;
      const uint64_t sma = va ? (s_in[ta] & alive_in) : 0ull, smb = vb ? (s_in[tb] & alive_in) : 0ull;
      // Row masks of the dense tiles, loaded beside the strip masks.
      const uint64_t rma = (pla & kPayloadDense) ? __builtin_nontemporal_load(rmask + pa) : 0ull;
      const uint64_t rmb = (plb & kPayloadDense) ? __builtin_nontemporal_load(rmask + pb) : 0ull;
      decode_group(st, pa, va, ta, pla, rma, vma, sma, dense_rows, x_in, acc, lane, keep0, keep1, all_alive, pad_job,
                   cnt);
      decode_group(st, pb, vb, tb, plb, rmb, vmb, smb, dense_rows, x_in, acc, lane, keep0, keep1, all_alive, pad_job,
                   cnt);
    }
    if (st.len) drain_jobs(st.jobs, st.len, x_in, acc, lane, keep0, keep1, all_alive, pad_job);
    const uint64_t sym = wave_or(st.sym);
    if (lane == 0 && sym) atomicOr(&s_sym, static_cast<unsigned long long>(sym));
    __syncthreads();
    write_strip<W, kNumBlock>(acc, s_n, s_v, s_z, s_sym, y, split, as_global(L.x_out), as_global(L.n_out),
                              as_global(L.v_out), as_global(L.z_out), as_global(L.s_out), read_lane(zp_l, i),
                              read_lane(np_l, i), read_lane(left_l, i), alive_next, cnt);
    __syncthreads();
    // Ready the accumulator and reductions for the next strip.
    for (int k = threadIdx.x; k < kWords; k += kNumBlock) acc[k] = 0;
    if (threadIdx.x == 0) { s_n = 0; s_v = 0; s_z = 0; s_sym = 0; }
    __syncthreads();
  }
  flush_block_counters(cnt, s_cnt, as_global(L.counters), log);
}

// --- numeric, 64 words per vertex, from the job list, by row segments ------------
// The job list holds one u32 per set bit of A^T, (source vertex << 6) | target row, in tile
// order, built once per graph from the tile payloads. A workgroup of this kernel handles one
// scheduler chunk and one of G segments of the 4096-source row (64/G words): it gathers only
// that segment of every source row, so each workgroup moves 512/G bytes per set bit and keeps
// a 64 x (64/G)-word strip accumulator in LDS. Workgroup b takes chunk b / G and segment
// b % G; workgroups are placed on the eight XCDs round-robin by block index, so with G = 4
// an XCD only ever gathers one segment and its L2 holds four times as many hot source rows
// as when every workgroup gathers whole rows. No tile is decoded during the hop.
//
// Per strip: the gathers (lanes grouped by (64/G)/2, each lane 16 bytes of the segment,
// kSegUnroll jobs in flight per group, source rows whose V bit is clear skipped), the
// structural sweep of the strip's tiles (only in the chunk's owner segment, chunk % G, which
// also counts tile products and edges), one barrier, then the write-out of the segment with
// the N, V, Z bits it can see. The accumulator is double-buffered, so the write-out of one
// strip overlaps the next strip's gathers; the mask slots are triple-buffered and flushed one
// strip later, after the next barrier. Unsplit strips store their per-segment masks in
// `part`, which seg_combine_kernel ORs into N, V, Z after the launch; pieces of split strips
// OR into zeroed memory and finalize_strips_kernel completes them, as for the sweep kernel.
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
constexpr int kSegWaves = 4;
constexpr int kSegBlock = kSegWaves * kWave;
#ifndef MG_SEG_UNROLL
#define MG_SEG_UNROLL 4
#endif
constexpr int kSegUnroll = MG_SEG_UNROLL;  // jobs in flight per lane group
constexpr uint32_t kNoJob = 0xFFFFFFFFu;
// Largest vertex count a u32 job can address (26 bits of source vertex).
constexpr uint64_t kJobMaxVertices = 1ull << 26;

// Set bits per A^T tile, from its payload (the compact entry count, or a dense tile's count).
__global__ void __launch_bounds__(kBlock) tile_bits_kernel(const uint64_t* __restrict__ payload, uint32_t n,
                                                           uint32_t* __restrict__ bits) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint64_t pl = payload[i];
    bits[i] = (pl & kPayloadDense) ? static_cast<uint32_t>(pl & 0x1FFFu) : static_cast<uint32_t>((pl >> 48) & 7u);
  }
}

// Writes the jobs of every A^T tile at job_off[i], in column-major order within the tile
// (by source vertex, then target row), so the jobs that gather the same source row into one
// strip are adjacent: lane = tile for the compact payloads (entries sorted by column); a dense
// tile is transposed by the whole wave (64 ballots: lane = column) and expanded column by column.
__global__ void __launch_bounds__(kBlock) job_fill_kernel(const uint64_t* __restrict__ payload,
                                                          const uint32_t* __restrict__ col,
                                                          const uint64_t* __restrict__ dense,
                                                          const uint32_t* __restrict__ job_off, uint32_t n,
                                                          uint32_t* __restrict__ jobs) {
  const uint32_t lane = lane_id();
  const uint32_t n_waves = gridDim.x * kWaves;
  for (uint32_t g0 = read_first_lane((blockIdx.x * kWaves + threadIdx.x / kWave) * 64u); g0 < n;
       g0 += n_waves * 64u) {
    const uint32_t i = g0 + lane;
    const bool valid = i < n;
    const uint64_t pl = valid ? payload[i] : 0ull;
    const uint32_t t = valid ? col[i] : 0u;
    const uint32_t at = valid ? job_off[i] : 0u;
    const bool is_dense = valid && (pl & kPayloadDense) != 0;
    if (valid && !is_dense) {
      const uint32_t n_ent = static_cast<uint32_t>(pl >> 48) & 7u;
      // Entries as (column << 6) | row, insertion-sorted (at most kPayloadMax).
      uint32_t key[kPayloadMax];
#pragma unroll
      for (uint32_t k = 0; k < kPayloadMax; ++k) {
        const uint32_t e = static_cast<uint32_t>(pl >> (12 * k)) & 0xFFFu;  // (row << 6) | column
        key[k] = k < n_ent ? (((e & 63u) << 6) | (e >> 6)) : 0xFFFFu;
      }
#pragma unroll
      for (uint32_t a = 1; a < kPayloadMax; ++a)
#pragma unroll
        for (uint32_t b = a; b > 0; --b)
          if (key[b] < key[b - 1]) {
            const uint32_t x = key[b];
            key[b] = key[b - 1];
            key[b - 1] = x;
          }
      for (uint32_t k = 0; k < n_ent; ++k) jobs[at + k] = ((t * 64u + (key[k] >> 6)) << 6) | (key[k] & 63u);
    }
    for (uint64_t dm = ballot(is_dense); dm; dm &= dm - 1) {
      const int k = __ffsll(static_cast<unsigned long long>(dm)) - 1;
      const uint64_t slot = (read_lane(pl, k) >> 13) & 0xFFFFFFFFull;
      const uint32_t tk = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(t), k));
      const uint32_t base = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(at), k));
      const uint64_t word = dense[slot * kTile + lane];  // row `lane` of the tile
      uint64_t colmask = 0;                               // rows of column `lane`
      for (int c = 0; c < kWave; ++c) {
        const uint64_t m = ballot((word >> c) & 1);
        if (lane == static_cast<uint32_t>(c)) colmask = m;
      }
      // Exclusive prefix of the columns' bit counts across the wave.
      uint32_t pos = __popcll(colmask);
#pragma unroll
      for (int m = 1; m < kWave; m <<= 1) {
        const uint32_t o = static_cast<uint32_t>(__shfl_up(static_cast<int>(pos), m, kWave));
        if (lane >= static_cast<uint32_t>(m)) pos += o;
      }
      pos -= __popcll(colmask);
      for (uint64_t r = colmask; r; r &= r - 1)
        jobs[base + pos++] = ((tk * 64u + lane) << 6) | static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(r)) - 1);
    }
  }
}

template <int G>
__global__ void __launch_bounds__(kSegBlock) walk_seg_kernel(const Chunk* __restrict__ chunks,
                                                             const PanelLevel* __restrict__ lv) {
  constexpr int SW = 64 / G;         // words of a row in one segment
  constexpr int LJ = SW / 2;         // lanes per job (two words each)
  constexpr int J = kWave / LJ;      // jobs per wave step
  constexpr int NG = kSegWaves * J;  // lane groups per workgroup
  constexpr int kAcc = 64 * SW;      // words of one strip accumulator
  static_assert(SW >= 2 && kAcc % kSegBlock == 0, "segment shape");
  __shared__ unsigned long long acc[2][kAcc];
  __shared__ unsigned long long s_mask[3][4];  // N, V, Z, S of the strip in each slot
  __shared__ unsigned long long s_cnt[kBlockCounters];
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t cid = blockIdx.x / G;
  const uint32_t seg = blockIdx.x % G;
  const Chunk ch = chunks[cid];
  const PanelLevel& L = lv[ch.panel];
  const bool owner = (cid % G) == seg;  // structural sweep, tile counts and edges
  const gptr<const uint32_t> __restrict__ off = as_global(L.off);
  const gptr<const uint32_t> __restrict__ col = as_global(L.col);
  const gptr<const uint32_t> __restrict__ jobs = as_global(L.jobs);
  const gptr<const uint32_t> __restrict__ job_off = as_global(L.job_off);
  const gptr<const uint64_t> __restrict__ v_in = as_global(L.v_in);
  const gptr<const uint64_t> __restrict__ s_in = as_global(L.s_in);
  const gptr<uint64_t> x_out = as_global(L.x_out);
  const uint64_t alive_in = L.alive_in;
  const uint64_t alive_next = L.alive_next;
  const uint32_t sub = lane % LJ;
  const uint32_t gi = wave * J + lane / LJ;
  const uint32_t w0 = seg * SW + 2 * sub;  // first of this lane's two words of a row
  const uint64_t keep0 = ((alive_in >> w0) & 1) ? ~0ull : 0ull;
  const uint64_t keep1 = ((alive_in >> (w0 + 1)) & 1) ? ~0ull : 0ull;
  const gptr<const char> xb = reinterpret_cast<gptr<const char>>(as_global(L.x_in)) + 8u * w0;
  // Per-strip values of the chunk, one strip per lane: row pointers, the job offsets of the
  // strip boundaries (clamped to the chunk's tile range), previous Z and N masks, S_k & alive.
  const uint32_t n_cs = ch.strip_end - ch.strip_begin;
  const uint32_t ys = ch.strip_begin + lane;
  const uint32_t off_l = lane <= n_cs ? off[ys] : 0u;
  const uint32_t cl_l = off_l < ch.tile_begin ? ch.tile_begin : (off_l > ch.tile_end ? ch.tile_end : off_l);
  const uint32_t jo_l = lane <= n_cs ? job_off[cl_l] : 0u;
  const uint64_t zp_l = lane < n_cs ? as_global(L.z_out)[ys] : 0ull;
  const uint64_t np_l = lane < n_cs ? as_global(L.n_out)[ys] : 0ull;
  const uint64_t left_l = (owner && lane < n_cs) ? (s_in[ys] & alive_in) : 0ull;
  uint32_t cnt[kBlockCounters];
#pragma unroll
  for (int i = 0; i < kBlockCounters; ++i) cnt[i] = 0;
  if (threadIdx.x < kBlockCounters) s_cnt[threadIdx.x] = 0;
  if (threadIdx.x < 12) (&s_mask[0][0])[threadIdx.x] = 0;
  for (int i = threadIdx.x; i < 2 * kAcc; i += kSegBlock) (&acc[0][0])[i] = 0;
  __syncthreads();

  // Stores strip i's masks (slot i % 3), once every wave has contributed; thread 0 only.
  auto flush_masks = [&](uint32_t i) {
    const uint32_t y = ch.strip_begin + i;
    const uint32_t slot = i % 3;
    const uint32_t r0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(off_l), static_cast<int>(i)));
    const uint32_t r1 =
        static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(off_l), static_cast<int>(i + 1)));
    const bool split = r0 < ch.tile_begin || r1 > ch.tile_end;
    const uint64_t sym = s_mask[slot][3];
    const gptr<uint64_t> part = as_global(L.part) + (static_cast<uint64_t>(seg) * L.n_strips + y) * 3;
    if (!split) {
      part[0] = s_mask[slot][0];
      part[1] = s_mask[slot][1];
      part[2] = s_mask[slot][2];
      cnt[bNumeric] += __popcll(s_mask[slot][0]);
      if (owner) {
        as_global(L.s_out)[y] = sym;
        cnt[bOut] += __popcll(sym);
        cnt[bLeft] += __popcll(read_lane(left_l, static_cast<int>(i)));
      }
    } else {
      part[0] = 0;
      part[1] = 0;
      part[2] = 0;
      if (owner && sym) atomicOr((unsigned long long*)(as_global(L.s_out) + y), static_cast<unsigned long long>(sym));
    }
    s_mask[slot][0] = 0;
    s_mask[slot][1] = 0;
    s_mask[slot][2] = 0;
    s_mask[slot][3] = 0;
  };

  for (uint32_t i = 0; i < n_cs; ++i) {
    const uint32_t b = i & 1;
    const uint32_t slot = i % 3;
    const uint32_t y = ch.strip_begin + i;
    const uint32_t r0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(off_l), static_cast<int>(i)));
    const uint32_t r1 =
        static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(off_l), static_cast<int>(i + 1)));
    const uint32_t t0 = r0 > ch.tile_begin ? r0 : ch.tile_begin;
    const uint32_t t1 = r1 < ch.tile_end ? r1 : ch.tile_end;
    const bool split = (t0 != r0) || (t1 != r1);
    const uint32_t ja = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(jo_l), static_cast<int>(i)));
    const uint32_t jb =
        static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(jo_l), static_cast<int>(i + 1)));
    char* const accb = reinterpret_cast<char*>(acc[b]) + 16u * sub;
    // Gathers: group gi takes jobs ja + gi, ja + gi + NG, ...
    for (uint32_t base = ja + gi; base < jb; base += NG * kSegUnroll) {
      uint32_t jv[kSegUnroll];
#pragma unroll
      for (int k = 0; k < kSegUnroll; ++k) {
        const uint32_t j = base + NG * k;
        jv[k] = j < jb ? __builtin_nontemporal_load(jobs + j) : kNoJob;
      }
      uint64_t vw[kSegUnroll];
#pragma unroll
      for (int k = 0; k < kSegUnroll; ++k) vw[k] = jv[k] != kNoJob ? v_in[jv[k] >> 12] : 0ull;
      u64x2 val[kSegUnroll];
#pragma unroll
      for (int k = 0; k < kSegUnroll; ++k) {
        const uint32_t u = jv[k] >> 6;
        const bool ok = (vw[k] >> (u & 63u)) & 1;
        if (owner && sub == 0) cnt[bEdges] += ok;
        val[k] = ok ? *reinterpret_cast<gptr<const u64x2>>(xb + static_cast<uint64_t>(u) * 512u) : u64x2{0ull, 0ull};
      }
#pragma unroll
      for (int k = 0; k < kSegUnroll; ++k) {
        const uint64_t v0 = val[k].x & keep0, v1 = val[k].y & keep1;
        if (v0 | v1) {
          unsigned long long* a = reinterpret_cast<unsigned long long*>(accb + (jv[k] & 63u) * (8u * SW));
          atomicOr(a, v0);
          atomicOr(a + 1, v1);
        }
      }
    }
    // Structural sweep of the strip's tiles (owner segment only).
    if (owner) {
      uint64_t sym = 0;
      for (uint32_t p = t0 + threadIdx.x; p < t1; p += kSegBlock) {
        const uint64_t sm = s_in[__builtin_nontemporal_load(col + p)] & alive_in;
        sym |= sm;
        cnt[bTP] += __popcll(sm);
        cnt[bUsed] += sm != 0;
      }
      sym = wave_or(sym);
      if (lane == 0 && sym) atomicOr(&s_mask[slot][3], static_cast<unsigned long long>(sym));
    }
    __syncthreads();
    if (threadIdx.x == 0 && i > 0) flush_masks(i - 1);
    // Write-out of this segment of the strip; the accumulator is cleared for strip i + 2.
    const uint64_t zp = read_lane(zp_l, static_cast<int>(i));
    const uint64_t np = read_lane(np_l, static_cast<int>(i));
    const gptr<uint64_t> dst = x_out + static_cast<uint64_t>(y) * 64u * 64u + seg * SW;
    uint64_t nmask = 0, vmask = 0, zmask = 0;
    uint32_t pc = 0;
#pragma unroll
    for (int m = 0; m < kAcc / kSegBlock; ++m) {
      const uint32_t idx = threadIdx.x + m * kSegBlock;
      const uint32_t r = idx / SW, w = seg * SW + idx % SW;
      const uint64_t val = acc[b][idx];
      acc[b][idx] = 0;
      if (!split) {
        if (val) {
          nmask |= 1ull << w;
          zmask |= 1ull << r;
          if ((alive_next >> w) & 1) vmask |= 1ull << r;
          pc += __popcll(val);
        }
        if (val != 0 || (((zp >> r) & 1) && ((np >> w) & 1))) __builtin_nontemporal_store(val, dst + r * 64u + (idx % SW));
      } else if (val) {
        atomicOr((unsigned long long*)(dst + r * 64u + (idx % SW)), static_cast<unsigned long long>(val));
      }
    }
    if (!split) {
      cnt[bPop] += pc;
      nmask = wave_or(nmask);
      vmask = wave_or(vmask);
      zmask = wave_or(zmask);
      if (lane == 0) {
        if (nmask) atomicOr(&s_mask[slot][0], static_cast<unsigned long long>(nmask));
        if (vmask) atomicOr(&s_mask[slot][1], static_cast<unsigned long long>(vmask));
        if (zmask) atomicOr(&s_mask[slot][2], static_cast<unsigned long long>(zmask));
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && n_cs > 0) flush_masks(n_cs - 1);
  flush_block_counters(cnt, s_cnt, as_global(L.counters), nullptr);
}

// --- numeric, 64 words per vertex, from the job list, whole rows, eight waves ----------
// The job-list kernel with one segment. The gathers of whole 512-byte source rows are bound
// by how many waves keep loads in flight on a CU, not by the loads each wave keeps in
// flight, so this kernel is built for the most resident waves: eight waves share one
// 64 x 64-word strip accumulator (32 KiB of LDS: four workgroups, 32 waves per CU) and the
// kernel fits in 64 VGPRs (eight waves per SIMD). A wave gathers one job at a time with all
// 64 lanes (lane = word of the row): the job is read out of a register holding 64 jobs with
// v_readlane, so the row address and the target row are scalar, and kLeanUnroll gathers
// per register set, two sets alternating, stay in flight.
//
// Per strip: the structural sweep of its tiles (S mask, tile products, used tiles), the jobs
// [job_off[t0], job_off[t1]) dealt to the waves in runs of kLeanRun (run R goes to wave
// R mod 8, so a light strip still spreads over the waves), source rows whose V bit is clear
// skipped, one barrier, and the write-out shared with the sweep kernel.
#ifndef MG_LEAN_UNROLL
#define MG_LEAN_UNROLL 2
#endif
constexpr int kLeanWaves = 8;
constexpr int kLeanBlock = kLeanWaves * kWave;
constexpr int kLeanUnroll = MG_LEAN_UNROLL;  // row gathers per register set
constexpr uint32_t kLeanRun = 8;             // consecutive jobs of a strip dealt to one wave
constexpr uint32_t kLeanSpan = kLeanRun * kLeanWaves * (kWave / kLeanRun);  // jobs per batch round
static_assert(kWave % kLeanRun == 0, "a batch holds whole runs");

struct LeanSet {
  uint64_t v[kLeanUnroll];
  uint64_t rows[kLeanUnroll];  // target rows of the gathered source row (0: none)
};

// Groups of a batch (`job`: one job per lane, `ok`: the lanes to gather). Jobs of one source
// that follow each other in a run of kLeanRun lanes (the job list is column-major within a
// tile) form one group, gathered once; `rows` becomes, in the group's first lane, the target
// rows of the whole group. Returns the first lanes of the groups. Vector work only: the
// scalar unit, which issues one instruction per cycle for the whole CU, is kept for the
// gathers' own bookkeeping.
__device__ inline uint64_t lean_groups(uint32_t job, uint64_t ok, uint32_t lane, uint64_t& rows) {
  const uint32_t u = job >> 6;
  rows = 1ull << (job & 63u);
  const uint32_t u_prev = static_cast<uint32_t>(__shfl_up(static_cast<int>(u), 1, kWave));
  const bool cont = (lane % kLeanRun) != 0 && ((ok >> lane) & 1) && ((ok >> (lane - 1)) & 1) && u_prev == u;
  const uint64_t cm = ballot(cont);
  if (cm) {
    const bool next_cont = lane + 1 < kWave && ((cm >> (lane + 1)) & 1);
    // One step per further job of the longest group: each lane ORs in its successor's rows.
    for (uint64_t ck = cm; ck; ck &= ck >> 1) {
      const uint64_t nxt = shfl64(rows, lane + 1 < kWave ? lane + 1 : lane);
      if (next_cont) rows |= nxt;
    }
  }
  return ok & ~cm;
}

// Issues the gathers of the next kLeanUnroll groups of `starts` (wave-uniform: first lanes of
// the groups still to gather): the row address (`row_addr`, one per lane of the batch, made by
// vector instructions) is read out as a scalar, and each lane loads word `word` of the row (its
// own word, or a live one when its own is dead).
__device__ inline void lean_issue(LeanSet& s, uint64_t& starts, uint64_t row_addr, uint64_t rows, uint32_t word) {
#pragma unroll
  for (int k = 0; k < kLeanUnroll; ++k) {
    s.rows[k] = 0;
    s.v[k] = 0;
    if (starts) {
      const int src = __ffsll(static_cast<unsigned long long>(starts)) - 1;
      starts &= starts - 1;
      const gptr<const uint64_t> row = reinterpret_cast<gptr<const uint64_t>>(read_lane(row_addr, src));
      s.v[k] = row[word];
      s.rows[k] = read_lane(rows, src);
    }
  }
}

// ORs each gathered row into its target rows; with count_rows, adds the jobs whose source row
// is non-empty in a live word to `jobs_ok` (wave-uniform).
__device__ inline void lean_commit(const LeanSet& s, char* __restrict__ acc_lane, uint64_t keep, bool count_rows,
                                   uint32_t& jobs_ok) {
#pragma unroll
  for (int k = 0; k < kLeanUnroll; ++k) {
    const uint64_t r = s.rows[k];
    if (r) {
      const uint64_t val = s.v[k] & keep;
      if (count_rows && ballot(val != 0)) jobs_ok += static_cast<uint32_t>(__popcll(r));
      atomicOr(reinterpret_cast<unsigned long long*>(acc_lane + 512u * (__ffsll(static_cast<unsigned long long>(r)) - 1)),
               val);
      for (uint64_t rest = r & (r - 1); rest; rest &= rest - 1)
        atomicOr(reinterpret_cast<unsigned long long*>(acc_lane + 512u * (__ffsll(static_cast<unsigned long long>(rest)) - 1)),
                 val);
    }
  }
}

// kPartial: some panel of the level has dead words (its gathers then read only the cache lines of
// live words); false keeps the plain lane-per-word loads and their registers.
template <bool kPartial>
__global__ void __launch_bounds__(kLeanBlock) __attribute__((amdgpu_waves_per_eu(kLeanWaves, kLeanWaves)))
    walk_lean_kernel(const Chunk* __restrict__ chunks, const PanelLevel* __restrict__ lv) {
  constexpr int kWords = 64 * 64;
  constexpr uint32_t kNoTile = 0xFFFFFFFFu;
  __shared__ unsigned long long acc[kWords];
  // Per strip parity: N, V, Z and S masks being reduced (thread 0 stores and clears a slot
  // after the strip's second barrier; the other slot is in use by then).
  __shared__ unsigned long long s_mask[2][4];
  __shared__ unsigned long long s_cnt[kBlockCounters];
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / kWave;
  const Chunk ch = chunks[blockIdx.x];
  const PanelLevel& L = lv[ch.panel];
  const gptr<const uint32_t> __restrict__ off = as_global(L.off);
  const gptr<const uint32_t> __restrict__ col = as_global(L.col);
  const gptr<const uint32_t> __restrict__ jobs = as_global(L.jobs);
  const gptr<const uint32_t> __restrict__ job_off = as_global(L.job_off);
  const gptr<const uint64_t> __restrict__ v_in = as_global(L.v_in);
  const gptr<const uint64_t> __restrict__ s_in = as_global(L.s_in);
  const gptr<const uint64_t> x_in = as_global(L.x_in);
  char* const acc_lane = reinterpret_cast<char*>(acc) + 8u * lane;
  const uint64_t alive_in = L.alive_in;
  const uint64_t alive_next = L.alive_next;
  const uint64_t keep = ((alive_in >> lane) & 1) ? ~0ull : 0ull;
  // Word this lane gathers: its own if live, else the first live word, so that a row of a
  // panel with few live words touches only their cache lines (the value is masked by keep).
  const uint32_t gather_word =
      (!kPartial || keep) ? lane
                          : (alive_in ? static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(alive_in)) - 1) : 0u);
  const bool lane_next = (alive_next >> lane) & 1;  // this lane's word propagates further
  // Per-strip values of the chunk, one strip per lane: row pointers and the job offsets of
  // the strip boundaries (clamped to the chunk's tile range).
  const uint32_t n_cs = ch.strip_end - ch.strip_begin;
  const uint32_t ys = ch.strip_begin + lane;
  const uint32_t off_l = lane <= n_cs ? off[ys] : 0u;
  const uint32_t cl_l = off_l < ch.tile_begin ? ch.tile_begin : (off_l > ch.tile_end ? ch.tile_end : off_l);
  const uint32_t jo_l = lane <= n_cs ? job_off[cl_l] : 0u;
  // This wave's batches over the whole chunk, in order: batch (strip i, base) holds the jobs
  // base + 64 * (l / kLeanRun) + l % kLeanRun below jb(i) (lane l: run l / kLeanRun of the
  // batch); the next one is base + kLeanSpan in the same strip, else the first batch of the
  // next strip with more than kLeanRun * wave jobs (strip n_cs: none left).
  const uint32_t jo_next_l = static_cast<uint32_t>(__shfl_down(static_cast<int>(jo_l), 1, kWave));
  const uint64_t wave_strips = ballot(lane < n_cs && jo_next_l - jo_l > kLeanRun * wave);
  auto first_batch = [&](uint32_t from, uint32_t& strip, uint32_t& base) {
    const uint64_t m = from < 64 ? (wave_strips & (~0ull << from)) : 0ull;
    strip = m ? static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(m)) - 1) : n_cs;
    base = m ? static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(jo_l), static_cast<int>(strip))) +
                   kLeanRun * wave
             : 0u;
  };
  auto next_batch = [&](uint32_t strip, uint32_t base, uint32_t& n_strip, uint32_t& n_base) {
    const uint32_t end =
        static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(jo_l), static_cast<int>(strip + 1)));
    if (base + kLeanSpan < end) {
      n_strip = strip;
      n_base = base + kLeanSpan;
    } else {
      first_batch(strip + 1, n_strip, n_base);
    }
  };
  auto load_batch = [&](uint32_t strip, uint32_t base) -> uint32_t {
    if (strip >= n_cs) return kNoJob;
    const uint32_t end =
        static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(jo_l), static_cast<int>(strip + 1)));
    const uint32_t j = base + kLeanRun * kLeanWaves * (lane / kLeanRun) + lane % kLeanRun;
    return j < end ? __builtin_nontemporal_load(jobs + j) : kNoJob;
  };
  // Source rows are filtered by their V bits unless at least two thirds of X_k's structural
  // tiles are numerically non-empty (the previous level's counters): then nearly every source
  // row is non-empty, gathering the few empty ones costs less than loading V words, and the
  // gathered rows that are non-empty in a live word (the V bit's definition) are counted.
  bool check_v;
  {
    const gptr<const unsigned long long> prev = as_global(static_cast<const unsigned long long*>(L.counters)) -
                                                kCounterStride + lane * kNumCounters;
    const uint64_t numeric = wave_sum(prev[kOutNumeric]);
    const uint64_t structural = wave_sum(prev[kOut]);
    check_v = !(structural > 0 && 3 * numeric >= 2 * structural);
  }
  uint32_t p_strip, p_base;
  first_batch(0, p_strip, p_base);
  uint32_t job = load_batch(p_strip, p_base);
  uint64_t vw = (check_v && job != kNoJob) ? v_in[job >> 12] : 0ull;
  uint32_t edges = 0;  // wave-uniform
  uint32_t cnt[kBlockCounters];
#pragma unroll
  for (int i = 0; i < kBlockCounters; ++i) cnt[i] = 0;
  if (threadIdx.x < kBlockCounters) s_cnt[threadIdx.x] = 0;
  if (threadIdx.x < 8) (&s_mask[0][0])[threadIdx.x] = 0;
  for (int i = threadIdx.x; i < kWords; i += kLeanBlock) acc[i] = 0;
  __syncthreads();
  // Tile products and used tiles of the hop, per tile column: column t holds col_count[t]
  // tiles, each with S_k[t] & alive as its source mask. The chunk adds up its columns.
  {
    const gptr<const uint32_t> cc = as_global(L.col_count);
    uint64_t tp = 0, used = 0;
    for (uint32_t t = ch.col_begin + threadIdx.x; t < ch.col_end; t += kLeanBlock) {
      const uint64_t sm = s_in[t] & alive_in;
      const uint64_t c = cc[t];
      tp += static_cast<uint64_t>(__popcll(sm)) * c;
      used += sm ? c : 0ull;
    }
    tp = wave_sum(tp);
    used = wave_sum(used);
    if (lane == 0) {
      if (tp) atomicAdd(&s_cnt[bTP], static_cast<unsigned long long>(tp));
      if (used) atomicAdd(&s_cnt[bUsed], static_cast<unsigned long long>(used));
    }
  }

  for (uint32_t i = 0; i < n_cs; ++i) {
    const uint32_t y = ch.strip_begin + i;
    const uint32_t sl = i & 1;
    const uint32_t r0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(off_l), static_cast<int>(i)));
    const uint32_t r1 =
        static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(off_l), static_cast<int>(i + 1)));
    const uint32_t t0 = r0 > ch.tile_begin ? r0 : ch.tile_begin;
    const uint32_t t1 = r1 < ch.tile_end ? r1 : ch.tile_end;
    const bool split = (t0 != r0) || (t1 != r1);
    // The strip's first kLeanBlock tile columns, for the structural mask after the gathers.
    const uint32_t p0 = t0 + threadIdx.x;
    const uint32_t c0 = p0 < t1 ? __builtin_nontemporal_load(col + p0) : kNoTile;
    // Gathers of this wave's batches in strip i. The next batch's jobs are loaded before this
    // batch's gathers start and its V words right after the first register set is issued,
    // so every load has a whole batch to arrive.
    while (p_strip == i) {
      uint32_t n_strip, n_base;
      next_batch(p_strip, p_base, n_strip, n_base);
      const uint32_t job_n = load_batch(n_strip, n_base);
      const bool ok = check_v ? ((vw >> ((job >> 6) & 63u)) & 1) : job != kNoJob;
      const uint64_t okm = ballot(ok);
      if (check_v) edges += static_cast<uint32_t>(__popcll(okm));
      const bool count_rows = !check_v;
      uint64_t rows;
      uint64_t m = lean_groups(job, okm, lane, rows);
      const uint64_t row_addr = reinterpret_cast<uint64_t>(x_in) + static_cast<uint64_t>(job >> 6) * 512u;
      LeanSet a, b;
      lean_issue(a, m, row_addr, rows, gather_word);
      const uint64_t vw_n = (check_v && job_n != kNoJob) ? v_in[job_n >> 12] : 0ull;
      for (;;) {
        if (!m) {
          lean_commit(a, acc_lane, keep, count_rows, edges);
          break;
        }
        lean_issue(b, m, row_addr, rows, gather_word);
        lean_commit(a, acc_lane, keep, count_rows, edges);
        if (!m) {
          lean_commit(b, acc_lane, keep, count_rows, edges);
          break;
        }
        lean_issue(a, m, row_addr, rows, gather_word);
        lean_commit(b, acc_lane, keep, count_rows, edges);
      }
      job = job_n;
      vw = vw_n;
      p_strip = n_strip;
      p_base = n_base;
    }
    // Structural mask of the strip: OR of its tiles' S_k[t] & alive. It stops as soon as it
    // holds every live word (exact: no tile can add to it), which at dense hops is after the
    // first tiles; a wave that finds the mask full when its gathers end loads nothing more.
    {
      volatile unsigned long long* const sym = &s_mask[sl][3];
      for (uint32_t p = p0; p < t1; p += kLeanBlock) {
        if (read_first_lane64(*sym) == alive_in) break;
        const uint32_t t = p == p0 ? c0 : __builtin_nontemporal_load(col + p);
        const uint64_t sm = s_in[t] & alive_in;
        if (sm) atomicOr(const_cast<unsigned long long*>(sym), static_cast<unsigned long long>(sm));
      }
    }
    __syncthreads();
    // Write-out: wave w stores rows w, w + 8, ... (lane = word) and clears them. An unsplit
    // strip reduces its masks by ballots (N: non-zero words, V: rows non-zero in a word that
    // propagates, Z: non-zero rows) and skips zero words outside the previous Z x N; a piece
    // of a split strip ORs into memory zeroed before the launch.
    const gptr<uint64_t> dst = as_global(L.x_out) + static_cast<uint64_t>(y) * kWords + lane;
    if (!split) {
      const uint64_t z_prev = as_global(L.z_out)[y];
      const bool n_prev_lane = (as_global(L.n_out)[y] >> lane) & 1;
      uint64_t zmask = 0, vmask = 0;
      bool nz_lane = false;
      uint32_t pc = 0;
#pragma unroll 2
      for (uint32_t r = wave; r < 64; r += kLeanWaves) {
        const uint64_t val = acc[r * 64 + lane];
        acc[r * 64 + lane] = 0;
        const bool nz = val != 0;
        const uint64_t rb = ballot(nz);
        const uint64_t rv = ballot(nz && lane_next);
        zmask |= rb ? (1ull << r) : 0ull;
        vmask |= rv ? (1ull << r) : 0ull;
        nz_lane |= nz;
        pc += __popcll(val);
        if (nz || (((z_prev >> r) & 1) && n_prev_lane)) __builtin_nontemporal_store(val, dst + r * 64u);
      }
      const uint64_t nmask = ballot(nz_lane);
      cnt[bPop] += pc;
      if (lane == 0) {
        if (nmask) atomicOr(&s_mask[sl][0], static_cast<unsigned long long>(nmask));
        if (vmask) atomicOr(&s_mask[sl][1], static_cast<unsigned long long>(vmask));
        if (zmask) atomicOr(&s_mask[sl][2], static_cast<unsigned long long>(zmask));
      }
    } else {
#pragma unroll 2
      for (uint32_t r = wave; r < 64; r += kLeanWaves) {
        const uint64_t val = acc[r * 64 + lane];
        acc[r * 64 + lane] = 0;
        if (val) atomicOr((unsigned long long*)(dst + r * 64u), static_cast<unsigned long long>(val));
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint64_t sym = s_mask[sl][3];
      if (!split) {
        as_global(L.n_out)[y] = s_mask[sl][0];
        as_global(L.v_out)[y] = s_mask[sl][1];
        as_global(L.z_out)[y] = s_mask[sl][2];
        as_global(L.s_out)[y] = sym;
        cnt[bNumeric] += __popcll(s_mask[sl][0]);
        cnt[bOut] += __popcll(sym);
        cnt[bLeft] += __popcll(s_in[y] & alive_in);
      } else if (sym) {
        atomicOr((unsigned long long*)(as_global(L.s_out) + y), static_cast<unsigned long long>(sym));
      }
      s_mask[sl][0] = 0;
      s_mask[sl][1] = 0;
      s_mask[sl][2] = 0;
      s_mask[sl][3] = 0;
    }
  }
  if (lane == 0) cnt[bEdges] += edges;
  flush_block_counters(cnt, s_cnt, as_global(L.counters), nullptr);
}

// N, V, Z of every unsplit strip from the G segments' partial masks (split strips get zeros
// here and their masks from finalize_strips_kernel, which runs after this).
__global__ void __launch_bounds__(kBlock) seg_combine_kernel(const PanelLevel* __restrict__ lv, uint32_t n_segs) {
  const PanelLevel& L = lv[blockIdx.y];
  const uint32_t ns = L.n_strips;
  for (uint32_t y = blockIdx.x * blockDim.x + threadIdx.x; y < ns; y += gridDim.x * blockDim.x) {
    uint64_t n = 0, v = 0, z = 0;
    for (uint32_t sg = 0; sg < n_segs; ++sg) {
      const uint64_t* p = L.part + (static_cast<uint64_t>(sg) * ns + y) * 3;
      n |= p[0];
      v |= p[1];
      z |= p[2];
    }
    L.n_out[y] = n;
    L.v_out[y] = v;
    L.z_out[y] = z;
  }
}

// Zero the strips that are written by several chunks (atomic OR targets) and their
// structural masks, for every level at once, at the start of a run (nothing reads or writes a
// level's output before its own launch).
template <int W>
__global__ void __launch_bounds__(kBlock) zero_split_kernel(const ZeroItem* __restrict__ items) {
  constexpr int kWords = 64 * W;
  const ZeroItem it = items[blockIdx.x];
  for (int i = threadIdx.x; i < kWords; i += kBlock) it.x[i] = 0;
  if (threadIdx.x == 0) *it.s = 0;
}

// Masks, popcount and tile counts of the split strips once every piece has been OR-ed in.
template <int W>
__global__ void __launch_bounds__(kBlock) finalize_strips_kernel(const StripItem* __restrict__ items,
                                                                 const PanelLevel* __restrict__ lv) {
  constexpr int kWords = 64 * W;
  __shared__ unsigned long long s_n, s_v, s_z, s_pop;
  if (threadIdx.x == 0) { s_n = 0; s_v = 0; s_z = 0; s_pop = 0; }
  __syncthreads();
  const StripItem it = items[blockIdx.x];
  const PanelLevel& L = lv[it.panel];
  const uint64_t* src = L.x_out + static_cast<uint64_t>(it.strip) * kWords;
  uint64_t nmask = 0, vmask = 0, zmask = 0, pc = 0;
  for (int i = threadIdx.x; i < kWords; i += kBlock) {
    const uint64_t val = src[i];
    if (val) {
      nmask |= 1ull << (i % W);
      zmask |= 1ull << (i / W);
      if ((L.alive_next >> (i % W)) & 1) vmask |= 1ull << (i / W);
      pc += __popcll(val);
    }
  }
  nmask = wave_or(nmask);
  vmask = wave_or(vmask);
  zmask = wave_or(zmask);
  pc = wave_sum(pc);
  if (lane_id() == 0) {
    if (nmask) atomicOr(&s_n, static_cast<unsigned long long>(nmask));
    if (vmask) atomicOr(&s_v, static_cast<unsigned long long>(vmask));
    if (zmask) atomicOr(&s_z, static_cast<unsigned long long>(zmask));
    if (pc) atomicAdd(&s_pop, static_cast<unsigned long long>(pc));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    L.n_out[it.strip] = s_n;
    L.v_out[it.strip] = s_v;
    L.z_out[it.strip] = s_z;
    unsigned long long* const slot = L.counters + (blockIdx.x % kCounterSlots) * kNumCounters;
    if (s_pop) atomicAdd(slot + kPop, s_pop);
    if (s_n) atomicAdd(slot + kOutNumeric, static_cast<unsigned long long>(__popcll(s_n)));
    const uint64_t so = L.s_out[it.strip];
    if (so) atomicAdd(slot + kOut, static_cast<unsigned long long>(__popcll(so)));
    const uint64_t left = L.s_in[it.strip] & L.alive_in;
    if (left) atomicAdd(slot + kLeft, static_cast<unsigned long long>(__popcll(left)));
  }
}

// --- first hop by push from the source strips ----------------------------------
// X_0 is one-hot: only the A^T tile columns of the strips that hold sources carry numeric
// work into X_1. The first hop of a walk can therefore read just those columns (through a
// column index of A^T) and OR each source row into its target rows with global atomics,
// instead of sweeping all strips with the LDS accumulator. Three or four launches produce the
// same X_1, masks and counters as the sweep: push_reset (clear the previous run's words of
// X_1, reset its masks), S_1 and the tile counts (either push_structural over the row-major
// index, or push_first_hop from the source columns plus per-column counts in push_finalize),
// push_first_hop (the jobs) and push_finalize (N, V masks, popcount and strip counts, read
// back from the rows the pushes marked in Z).

// Positions [begin, end) of the A^T column index that belong to source strip `strip`.
struct PushItem {
  uint32_t strip, begin, end, panel;
};
constexpr uint32_t kPushItemTiles = 512;   // column positions per item
constexpr int kPushCap = 128;              // jobs per wave list
constexpr int kPushUnroll = 8;             // source rows loaded per step

// One wave per strip of panel panels[blockIdx.y]: zero X_1 words inside the previous Z x N
// masks, reset X_1's masks.
template <int W>
__global__ void __launch_bounds__(kBlock) push_reset_kernel(const PanelLevel* __restrict__ lv,
                                                            const uint32_t* __restrict__ panels) {
  const PanelLevel& L = lv[panels[blockIdx.y]];
  const uint32_t lane = lane_id();
  const gptr<uint64_t> x = as_global(L.x_out);
  const uint32_t n_waves = gridDim.x * kWaves;
  for (uint32_t y = read_first_lane(blockIdx.x * kWaves + threadIdx.x / kWave); y < L.n_strips; y += n_waves) {
    const uint64_t z = as_global(L.z_out)[y];
    const uint64_t n = as_global(L.n_out)[y];
    if (z != 0 && n != 0) {
      const bool mine = lane < W && ((n >> lane) & 1);
      for (uint64_t m = z; m; m &= m - 1) {
        const uint64_t r = static_cast<uint64_t>(__ffsll(static_cast<unsigned long long>(m)) - 1);
        if (mine) x[(static_cast<uint64_t>(y) * 64 + r) * W + lane] = 0;
      }
    }
    if (lane == 0) {
      as_global(L.n_out)[y] = 0;
      as_global(L.v_out)[y] = 0;
      as_global(L.z_out)[y] = 0;
      as_global(L.s_out)[y] = 0;
    }
  }
}

// S_1 and the tile-product counts of the first hop when the pushed columns hold many of the
// tiles (batches of large queries): per A^T tile (y, t), S_1[y] |= S_0[t] & alive. One
// workgroup per scheduler chunk; the chunk's strips are reduced in LDS and stored (pieces of
// split strips OR into S_1, zeroed by push_reset). With few pushed tiles (a walk from random
// sources) push_first_hop ORs S_1 from the source columns instead and push_finalize counts
// the tile products per column.
__global__ void __launch_bounds__(kBlock) push_structural_kernel(const Chunk* __restrict__ chunks,
                                                                 const PanelLevel* __restrict__ lv) {
  __shared__ unsigned long long s_sym[kChunkStrips];
  __shared__ unsigned long long s_cnt[kBlockCounters];
  const Chunk ch = chunks[blockIdx.x];
  const PanelLevel& L = lv[ch.panel];
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / kWave;
  const gptr<const uint32_t> col = as_global(L.col);
  const gptr<const uint32_t> row = as_global(L.row);
  const gptr<const uint64_t> s_in = as_global(L.s_in);
  const uint64_t alive = L.alive_in;
  uint32_t cnt[kBlockCounters];
#pragma unroll
  for (int i = 0; i < kBlockCounters; ++i) cnt[i] = 0;
  if (threadIdx.x < kBlockCounters) s_cnt[threadIdx.x] = 0;
  if (threadIdx.x < kChunkStrips) s_sym[threadIdx.x] = 0;
  __syncthreads();
  // kGroups groups of 64 tiles per wave step, their index loads issued together.
  constexpr int kGroups = 4;
  for (uint32_t g0 = ch.tile_begin + 64 * wave; g0 < ch.tile_end; g0 += 64 * kWaves * kGroups) {
    uint32_t t[kGroups], y[kGroups];
    uint64_t sv[kGroups];
#pragma unroll
    for (int q = 0; q < kGroups; ++q) {
      const uint32_t p = g0 + 64 * kWaves * q + lane;
      const bool v = p < ch.tile_end;
      t[q] = v ? __builtin_nontemporal_load(col + p) : 0xFFFFFFFFu;
      y[q] = v ? __builtin_nontemporal_load(row + p) : 0u;
    }
#pragma unroll
    for (int q = 0; q < kGroups; ++q) sv[q] = t[q] != 0xFFFFFFFFu ? (s_in[t[q]] & alive) : 0ull;
    // A group mostly lies in one strip: lanes in lane 0's strip are reduced across the wave
    // first, so the LDS ORs do not all queue on one address.
#pragma unroll
    for (int q = 0; q < kGroups; ++q) {
      cnt[bTP] += __popcll(sv[q]);
      cnt[bUsed] += sv[q] != 0;
      const uint32_t y0 = read_first_lane(y[q]);
      const bool same = y[q] == y0;
      const uint64_t red = wave_or(same ? sv[q] : 0ull);
      if (lane == 0 && red) atomicOr(&s_sym[y0 - ch.strip_begin], static_cast<unsigned long long>(red));
      if (!same && sv[q]) atomicOr(&s_sym[y[q] - ch.strip_begin], static_cast<unsigned long long>(sv[q]));
    }
  }
  __syncthreads();
  const uint32_t n_cs = ch.strip_end - ch.strip_begin;
  if (threadIdx.x < n_cs) {
    const uint32_t y = ch.strip_begin + threadIdx.x;
    const gptr<const uint32_t> off = as_global(L.off);
    const bool split = ch.tile_begin != off[y] || ch.tile_end != off[y + 1];
    if (split) {
      if (s_sym[threadIdx.x]) atomicOr((unsigned long long*)(L.s_out + y), s_sym[threadIdx.x]);
    } else {
      as_global(L.s_out)[y] = s_sym[threadIdx.x];
    }
  }
  flush_block_counters(cnt, s_cnt, as_global(L.counters), nullptr);
}

// Pushes the first n jobs of a wave list; a job is (target vertex << 32) | source vertex.
// Each source row is loaded whole (lane = word) and its non-zero words are OR-ed into the
// target row; the target row is marked in Z.
template <int W>
__device__ inline void push_drain(unsigned long long* __restrict__ jobs, uint32_t n, gptr<const uint64_t> x_in,
                                  gptr<uint64_t> x_out, gptr<uint64_t> z_out, uint32_t lane,
                                  unsigned long long pad_job) {
  const uint32_t padded = (n + kPushUnroll - 1) / kPushUnroll * kPushUnroll;
  if (lane < padded - n) jobs[n + lane] = pad_job;
  wave_lds_sync();
  for (uint32_t j = 0; j < padded; j += kPushUnroll) {
    unsigned long long e[kPushUnroll];
    uint64_t val[kPushUnroll];
#pragma unroll
    for (int k = 0; k < kPushUnroll; ++k) {
      e[k] = jobs[j + k];
      val[k] = lane < W ? x_in[(e[k] & 0xFFFFFFFFull) * W + lane] : 0ull;
    }
#pragma unroll
    for (int k = 0; k < kPushUnroll; ++k) {
      const uint64_t v = e[k] >> 32;
      if (val[k]) atomicOr((unsigned long long*)(x_out + v * W + lane), static_cast<unsigned long long>(val[k]));
      const uint64_t any = ballot(val[k] != 0);
      if (lane == 0 && any != 0) atomicOr((unsigned long long*)(z_out + (v >> 6)), 1ull << (v & 63));
    }
  }
  wave_lds_sync();
}

// One workgroup per item; its waves take the item's column positions in groups of 64
// (lane = tile). Per set bit of a tile whose source vertex is non-empty in X_0: one job. With
// s_from_columns, every tile (y, t) of the column also ORs S_0[t] & alive into S_1[y]
// (structural X_1; tiles of other columns add nothing): a hub strip y is reached from most
// source columns, so S_1[y] is read first and only a missing bit costs an atomic.
template <int W>
__global__ void __launch_bounds__(kBlock) push_first_hop_kernel(const PushItem* __restrict__ items,
                                                                const PanelLevel* __restrict__ lv,
                                                                bool s_from_columns) {
  __shared__ unsigned long long jobs_all[kWaves][kPushCap];
  __shared__ unsigned long long s_cnt[kBlockCounters];
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / kWave;
  const PushItem it = items[blockIdx.x];
  const PanelLevel& L = lv[it.panel];
  const gptr<const uint32_t> csc_tile = as_global(L.csc_tile);
  const gptr<const uint32_t> csc_row = as_global(L.csc_row);
  const gptr<const uint64_t> csc_payload = as_global(L.csc_payload);
  const gptr<const uint64_t> dense_rows = as_global(L.dense);
  const gptr<const uint64_t> rmask = as_global(L.rmask);
  const gptr<const uint64_t> x_in = as_global(L.x_in);
  const gptr<uint64_t> x_out = as_global(L.x_out);
  const gptr<uint64_t> z_out = as_global(L.z_out);
  const uint32_t t = it.strip;
  const uint64_t vm = as_global(L.v_in)[t];
  const uint64_t sm = s_from_columns ? (as_global(L.s_in)[t] & L.alive_in) : 0ull;
  const unsigned long long pad_job = static_cast<unsigned long long>(L.n_strips) * 64;  // zero source row
  unsigned long long* const jobs = jobs_all[wave];
  uint32_t cnt[kBlockCounters];
#pragma unroll
  for (int i = 0; i < kBlockCounters; ++i) cnt[i] = 0;
  if (threadIdx.x < kBlockCounters) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  uint32_t len = 0;
  for (uint32_t g = it.begin + 64 * wave; g < it.end; g += 64 * kWaves) {
    const uint32_t p = g + lane;
    const bool valid = p < it.end;
    const uint32_t i = valid ? __builtin_nontemporal_load(csc_tile + p) : 0u;
    const uint32_t y = valid ? __builtin_nontemporal_load(csc_row + p) : 0u;
    const uint64_t pl = valid ? __builtin_nontemporal_load(csc_payload + p) : 0ull;
    const uint64_t rm_l = (pl & kPayloadDense) ? rmask[i] : 0ull;
    if (valid && sm) {
      unsigned long long* const s1 = reinterpret_cast<unsigned long long*>(L.s_out + y);
      const uint64_t cur = __hip_atomic_load(s1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if ((cur | sm) != cur) atomicOr(s1, static_cast<unsigned long long>(sm));
    }
    const bool act = valid && vm != 0 && pl != 0;
    const bool dense = act && (pl & kPayloadDense) != 0;
    const uint64_t ent = (act && !dense) ? pl : 0ull;
    const uint32_t n_ent = static_cast<uint32_t>(ent >> 48) & 7u;
#pragma unroll
    for (uint32_t k = 0; k < kPayloadMax; ++k) {
      const uint32_t e = static_cast<uint32_t>(ent >> (12 * k)) & 0xFFFu;
      const bool ok = k < n_ent && ((vm >> (e & 63u)) & 1);
      const uint64_t has = ballot(ok);
      if (has) {
        if (len + 64 > kPushCap) {
          push_drain<W>(jobs, len, x_in, x_out, z_out, lane, pad_job);
          len = 0;
        }
        if (ok)
          jobs[len + lanes_below(has)] = (static_cast<unsigned long long>(y * 64 + (e >> 6)) << 32) | (t * 64 + (e & 63u));
        len += __popcll(has);
        if (lane == 0) cnt[bEdges] += __popcll(has);
      }
    }
    // Dense tiles: the whole wave loads them (lane = row), kDenseBatch at a time, with the
    // row masks loaded per lane beside the payloads.
    for (uint64_t dm = ballot(dense); dm;) {
      uint64_t bits[kDenseBatch];
      uint32_t yb[kDenseBatch];
#pragma unroll
      for (int q = 0; q < kDenseBatch; ++q) {
        const bool any = dm != 0;
        const int k = any ? __ffsll(static_cast<unsigned long long>(dm)) - 1 : 0;
        dm &= dm - 1;
          const uint64_t slot = (read_lane(pl, k) >> 13) & 0xFFFFFFFFull;
        yb[q] = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(y), k));
        const uint64_t rm = any ? read_lane(rm_l, k) : 0ull;
        bits[q] = ((rm >> lane) & 1) ? (dense_rows[slot * kTile + lane] & vm) : 0ull;
      }
#pragma unroll
      for (int q = 0; q < kDenseBatch; ++q) {
        for (uint64_t has = ballot(bits[q] != 0); has; has = ballot(bits[q] != 0)) {
          if (len + 64 > kPushCap) {
            push_drain<W>(jobs, len, x_in, x_out, z_out, lane, pad_job);
            len = 0;
          }
          if (bits[q]) {
            const uint32_t c = static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(bits[q])) - 1);
            jobs[len + lanes_below(has)] = (static_cast<unsigned long long>(yb[q] * 64 + lane) << 32) | (t * 64 + c);
            bits[q] &= bits[q] - 1;
          }
          len += __popcll(has);
          if (lane == 0) cnt[bEdges] += __popcll(has);
        }
      }
    }
  }
  if (len) push_drain<W>(jobs, len, x_in, x_out, z_out, lane, pad_job);
  flush_block_counters(cnt, s_cnt, as_global(L.counters), nullptr);
}

// One wave per strip: N and V masks and popcount of X_1 from the rows marked in Z, and the
// per-strip tile counts (T_L from S_0, T_C from S_1, numeric tiles from N_1); with
// column_counts, per tile column t = the strip, the tile products and used tiles: col_count[t]
// tiles with S_0[t] & alive.
template <int W>
__global__ void __launch_bounds__(kBlock) push_finalize_kernel(const PanelLevel* __restrict__ lv,
                                                               const uint32_t* __restrict__ panels,
                                                               bool column_counts) {
  __shared__ unsigned long long s_cnt[kBlockCounters];
  const PanelLevel& L = lv[panels[blockIdx.y]];
  const uint32_t lane = lane_id();
  const gptr<const uint64_t> x = as_global(L.x_out);
  uint32_t cnt[kBlockCounters];
#pragma unroll
  for (int i = 0; i < kBlockCounters; ++i) cnt[i] = 0;
  if (threadIdx.x < kBlockCounters) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t n_waves = gridDim.x * kWaves;
  uint64_t tp = 0, used = 0;  // lane 0
  for (uint32_t y = read_first_lane(blockIdx.x * kWaves + threadIdx.x / kWave); y < L.n_strips; y += n_waves) {
    const uint64_t z = as_global(L.z_out)[y];
    uint64_t nb = 0, vb = 0;
    // Rows of Z four at a time, their loads issued together.
    for (uint64_t m = z; m;) {
      uint64_t val[4];
      uint32_t r[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const bool any = m != 0;
        r[q] = any ? static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(m)) - 1) : 0u;
        m &= m - 1;
        val[q] = (any && lane < W) ? x[(static_cast<uint64_t>(y) * 64 + r[q]) * W + lane] : 0ull;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        cnt[bPop] += __popcll(val[q]);
        const uint64_t words = ballot(val[q] != 0);
        nb |= words;
        if (words & L.alive_next) vb |= 1ull << r[q];
      }
    }
    if (lane == 0) {
      if (z != 0) {
        as_global(L.n_out)[y] = nb;
        as_global(L.v_out)[y] = vb;
      }
      cnt[bNumeric] += __popcll(nb);
      cnt[bOut] += __popcll(as_global(L.s_out)[y]);
      const uint64_t left = as_global(L.s_in)[y] & L.alive_in;
      cnt[bLeft] += __popcll(left);
      const uint64_t cc = column_counts ? as_global(L.col_count)[y] : 0ull;
      tp += static_cast<uint64_t>(__popcll(left)) * cc;
      used += left ? cc : 0ull;
    }
  }
  if (lane == 0) {
    if (tp) atomicAdd(&s_cnt[bTP], static_cast<unsigned long long>(tp));
    if (used) atomicAdd(&s_cnt[bUsed], static_cast<unsigned long long>(used));
  }
  flush_block_counters(cnt, s_cnt, as_global(L.counters), nullptr);
}

// Digest of words [w0, w0+n_words) of a vertex-major frontier (word index relative to w0).
// Row of X that holds vertex v: strips are stored in the engine's order (map: vertex strip ->
// engine strip; null keeps the vertex order).
__device__ inline uint64_t engine_row(const uint32_t* __restrict__ map, uint64_t v) {
  return map ? static_cast<uint64_t>(map[v >> 6]) * 64 + (v & 63) : v;
}

__global__ void __launch_bounds__(kBlock) frontier_digest_kernel(const uint64_t* __restrict__ x, uint64_t n_vertices,
                                                                 uint32_t W, uint32_t w0, uint32_t n_words,
                                                                 const uint32_t* __restrict__ map,
                                                                 unsigned long long* __restrict__ out) {
  uint64_t acc = 0;
  const uint64_t total = n_vertices * n_words;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride) {
    const uint64_t v = i / n_words;
    const uint32_t w = static_cast<uint32_t>(i % n_words);
    const uint64_t val = x[engine_row(map, v) * W + w0 + w];
    if (val) acc += digest_term(v * 64 + w, val);
  }
  acc = wave_sum(acc);
  if (lane_id() == 0 && acc) atomicAdd(out, static_cast<unsigned long long>(acc));
}

// Numeric occupancy of a hop's left operand X_k (per-strip numeric masks n_prev, words of
// live queries `alive`): out[0] += A^T tiles whose column strip holds a numerically non-empty
// X_k tile (T_R,used counted numerically), out[1] += numerically non-empty X_k tiles.
__global__ void __launch_bounds__(kBlock) numeric_use_kernel(const uint32_t* __restrict__ col, uint32_t n_tiles,
                                                             const uint64_t* __restrict__ n_prev, uint32_t n_strips,
                                                             uint64_t alive, unsigned long long* __restrict__ out) {
  uint64_t used = 0, left = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_tiles; i += stride)
    used += (n_prev[col[i]] & alive) != 0;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n_strips; t += stride)
    left += __popcll(n_prev[t] & alive);
  used = wave_sum(used);
  left = wave_sum(left);
  if (lane_id() == 0) {
    if (used) atomicAdd(out, static_cast<unsigned long long>(used));
    if (left) atomicAdd(out + 1, static_cast<unsigned long long>(left));
  }
}

// Copy words [w0, w0+n_words) of every vertex, in vertex order, into a compact
// (n_vertices x n_words) array.
__global__ void __launch_bounds__(kBlock) frontier_slice_kernel(const uint64_t* __restrict__ x, uint64_t n_vertices,
                                                                uint32_t W, uint32_t w0, uint32_t n_words,
                                                                const uint32_t* __restrict__ map,
                                                                uint64_t* __restrict__ out) {
  const uint64_t total = n_vertices * n_words;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride)
    out[i] = x[engine_row(map, i / n_words) * W + w0 + (i % n_words)];
}

// Set bits of each chunk's A^T tiles (from the tile payloads): those of the chunk's panel
// when `lv` is given, else those of `payload0` (one graph's plan).
__global__ void __launch_bounds__(kBlock) chunk_bits_kernel(const Chunk* __restrict__ chunks,
                                                            const PanelLevel* __restrict__ lv,
                                                            const uint64_t* __restrict__ payload0,
                                                            uint32_t* __restrict__ bits_out) {
  __shared__ unsigned long long s_bits;
  if (threadIdx.x == 0) s_bits = 0;
  __syncthreads();
  const Chunk ch = chunks[blockIdx.x];
  const uint64_t* payload = lv ? lv[ch.panel].payload : payload0;
  uint64_t b = 0;
  for (uint32_t p = ch.tile_begin + threadIdx.x; p < ch.tile_end; p += kBlock) {
    const uint64_t pl = payload[p];
    b += (pl & kPayloadDense) ? (pl & 0x1FFFu) : ((pl >> 48) & 7u);
  }
  b = wave_sum(b);
  if (lane_id() == 0 && b) atomicAdd(&s_bits, static_cast<unsigned long long>(b));
  __syncthreads();
  if (threadIdx.x == 0) bits_out[blockIdx.x] = static_cast<uint32_t>(s_bits);
}

// Export tile (xr, y) of X_k in the reference orientation: X tile word c (source xr*64+c)
// bit r (vertex y*64+r) = bit c of vertex-major word w0+xr of vertex y*64+r.
__global__ void __launch_bounds__(kBlock) export_tiles_kernel(const uint32_t* __restrict__ row_idx,
                                                              const uint32_t* __restrict__ col_idx, uint32_t n_tiles,
                                                              const uint64_t* __restrict__ x, uint32_t W, uint32_t w0,
                                                              const uint32_t* __restrict__ map,
                                                              uint64_t* __restrict__ out) {
  const uint32_t lane = lane_id();
  const uint32_t n_waves = gridDim.x * kWaves;
  for (uint32_t i = read_first_lane(blockIdx.x * kWaves + threadIdx.x / kWave); i < n_tiles; i += n_waves) {
    const uint32_t xr = row_idx[i], y = col_idx[i];
    const uint64_t w = x[engine_row(map, static_cast<uint64_t>(y) * 64 + lane) * W + w0 + xr];
    uint64_t mine = 0;
    for (int c = 0; c < 64; ++c) {
      const uint64_t m = ballot((w >> c) & 1);
      if (lane == static_cast<uint32_t>(c)) mine = m;
    }
    out[static_cast<uint64_t>(i) * kTile + lane] = mine;
  }
}

// Scheduler: chunks over one graph's A^T tile stream (panel index filled in later).
std::vector<Chunk> build_chunks(const std::vector<uint32_t>& off, uint32_t n_strips, uint32_t chunk_tiles,
                                std::vector<uint32_t>* split) {
  std::vector<Chunk> chunks;
  uint32_t cur_begin = 0, cur_tiles = 0, cur_strips = 0;
  auto close = [&](uint32_t end_strip) {
    if (end_strip > cur_begin) chunks.push_back({0, cur_begin, end_strip, off[cur_begin], off[end_strip]});
    cur_begin = end_strip;
    cur_tiles = 0;
    cur_strips = 0;
  };
  for (uint32_t y = 0; y < n_strips; ++y) {
    const uint32_t sz = off[y + 1] - off[y];
    if (sz > chunk_tiles) {
      close(y);
      split->push_back(y);
      for (uint32_t b = off[y]; b < off[y + 1]; b += chunk_tiles)
        chunks.push_back({0, y, y + 1, b, std::min(off[y + 1], b + chunk_tiles)});
      cur_begin = y + 1;
      continue;
    }
    if (cur_tiles + sz > chunk_tiles || cur_strips == kChunkStrips) close(y);
    cur_tiles += sz;
    ++cur_strips;
  }
  close(n_strips);
  // Tile columns, split evenly over the chunks.
  const uint64_t nc = chunks.size();
  for (uint64_t i = 0; i < nc; ++i) {
    chunks[i].col_begin = static_cast<uint32_t>(i * n_strips / nc);
    chunks[i].col_end = static_cast<uint32_t>((i + 1) * n_strips / nc);
  }
  return chunks;
}

// Per-graph scheduler data, built once from the A^T row pointer.
//
// Engine order of the vertices: the walk stores strip y of every frontier at position
// map[y] (a fixed pseudo-random permutation of the strips; vertex v = 64y + r lives in row
// 64 map[y] + r) and walks A^T with its strips and tile columns relabelled the same way, so
// the hop kernels see an isomorphic graph. R-MAT vertex ids are biased bit by bit (hubs have
// few one bits), and the frontier rows of hubs, gathered most, would otherwise crowd onto
// the same memory channels: gathering the R-MAT-22 hop's 65 M rows takes 4.79 ms in vertex
// order and 4.10 ms with strips permuted (3.64 ms at four times the grid). Sources are
// mapped on the way in, frontiers, digests and exports on the way out; counts are unchanged.
struct GraphPlan {
  const mg_tiled_matrix* at = nullptr;
  uint32_t n_strips = 0;
  // map[y] = engine strip of vertex strip y; empty when the vertex order is kept.
  std::vector<uint32_t> map;
  DevBuf<uint32_t> d_map;
  // A^T's index in engine order (r_*: owned copies when relabelled); orig[i] = position of
  // engine tile i in at->data (null when the order is kept).
  DevBuf<uint32_t> r_off, r_col, r_row, r_orig;
  const uint32_t* off = nullptr;
  const uint32_t* col = nullptr;
  const uint32_t* row = nullptr;
  const uint32_t* orig = nullptr;
  uint32_t engine_vertex(uint32_t v) const { return map.empty() ? v : map[v >> 6] * 64u + (v & 63u); }
  const uint32_t* device_map() const { return map.empty() ? nullptr : d_map.get(); }
  std::vector<Chunk> chunks;
  std::vector<uint64_t> chunk_cost;  // launch-order key per chunk (chunk_cost_of)
  std::vector<uint32_t> split;
  DevBuf<uint64_t> rmask;    // row-occupancy mask per A^T tile, taken when the plan is made
  DevBuf<uint64_t> payload;  // compact bit list per A^T tile (tile_payload_kernel)
  DevBuf<uint64_t> dense;    // rows of the dense tiles (dense_compact_kernel)
  // Column index of A^T (first-hop push): tile indices grouped by tile column, each with a
  // copy of its strip and payload so the push streams them, and the column offsets into it
  // (host copy, n_strips + 1). Empty unless build_column_index ran.
  DevBuf<uint32_t> csc_tile;
  DevBuf<uint32_t> csc_row;
  DevBuf<uint64_t> csc_payload;
  std::vector<uint32_t> host_csc_off;
  uint64_t total_bits = 0;   // set bits of A^T (chunk_costs)
  // Job list of A^T for the segmented hop kernel (build_job_list): empty unless built.
  DevBuf<uint32_t> jobs;
  DevBuf<uint32_t> job_off;
  DevBuf<uint32_t> col_count;  // tiles per tile column of A^T (with the column index or the job list)
  bool has_jobs = false;
};

// Launch-order key of each chunk, in hundredths: its set bits + 0.35 per tile + 24 per
// strip. Set bits count the jobs a full frontier would give, tiles the row-mask and column
// reads, strips the write-out; heaviest first leaves light chunks for the tail of a launch.
// Within one allocation this order beat ordering by tiles alone by 2-4 % on hops 2-4 of the
// 4096-source walk on R-MAT 22 (tiles alone beat strip order by 4 % and reversed by 35 %).
// *total_bits receives the set bits of all chunks (every tile once).
std::vector<uint64_t> chunk_costs(const GraphPlan& g, hipStream_t st, uint64_t* total_bits) {
  *total_bits = 0;
  const size_t n = g.chunks.size();
  std::vector<uint64_t> cost(n, 0);
  if (!n || !g.at->n_nz_tile) return cost;
  DevBuf<Chunk> dch(n);
  DevBuf<uint32_t> bits(n);
  MG_HIP(hipMemcpyAsync(dch.get(), g.chunks.data(), sizeof(Chunk) * n, hipMemcpyHostToDevice, st));
  chunk_bits_kernel<<<static_cast<uint32_t>(n), kBlock, 0, st>>>(dch.get(), nullptr, g.payload.get(), bits.get());
  MG_LAUNCH_CHECK();
  std::vector<uint32_t> hb(n);
  MG_HIP(hipMemcpyAsync(hb.data(), bits.get(), sizeof(uint32_t) * n, hipMemcpyDeviceToHost, st));
  MG_HIP(hipStreamSynchronize(st));
  for (size_t i = 0; i < n; ++i) {
    const Chunk& c = g.chunks[i];
    cost[i] = 100ull * hb[i] + 35ull * (c.tile_end - c.tile_begin) + 2400ull * (c.strip_end - c.strip_begin);
    *total_bits += hb[i];
  }
  return cost;
}

// Whether the segmented hop kernel can run on this graph: every job must fit in a u32
// (source vertex below 2^26) and every job position, plus a workgroup's look-ahead past the
// end of a range, in a u32.
bool job_list_fits(const GraphPlan& g) {
  return static_cast<uint64_t>(g.n_strips) * 64 <= kJobMaxVertices && g.total_bits <= 0xFFFF0000ull;
}

// Tiles per tile column of A^T (count zeroed by the caller).
__global__ void __launch_bounds__(kBlock) column_count_kernel(const uint32_t* __restrict__ col, uint32_t n,
                                                              uint32_t* __restrict__ count) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    atomicAdd(count + col[i], 1u);
}

// Builds the job list of A^T (once per graph): per-tile bit counts from the payloads, their
// exclusive scan as job_off, then job_fill_kernel. 4 bytes per set bit plus 4 per tile.
void build_job_list(GraphPlan& g, hipStream_t st) {
  if (g.has_jobs) return;
  MG_REQUIRE(job_list_fits(g), "walk: graph too large for the job-list hop kernel");
  const uint32_t n = g.at->n_nz_tile;
  g.job_off.alloc(n + 1ull);
  g.jobs.alloc(std::max<uint64_t>(1, g.total_bits));
  DevBuf<uint32_t> bits;
  if (n) {
    bits.alloc(n);
    tile_bits_kernel<<<grid_for(n, kBlock), kBlock, 0, st>>>(g.payload.get(), n, bits.get());
    MG_LAUNCH_CHECK();
    exclusive_scan_u32_on(st, bits.get(), g.job_off.get(), n);
    job_fill_kernel<<<grid_for(n, kWaves * 64, 65536), kBlock, 0, st>>>(g.payload.get(), g.col,
                                                                        g.dense.get(), g.job_off.get(), n,
                                                                        g.jobs.get());
    MG_LAUNCH_CHECK();
  } else {
    MG_HIP(hipMemsetAsync(g.job_off.get(), 0, sizeof(uint32_t), st));
  }
  if (!g.col_count.get()) {  // else the column index built it
    const uint32_t ns = std::max<uint32_t>(1, g.n_strips);
    g.col_count.alloc(ns);
    MG_HIP(hipMemsetAsync(g.col_count.get(), 0, sizeof(uint32_t) * ns, st));
    if (n) {
      column_count_kernel<<<grid_for(n, kBlock), kBlock, 0, st>>>(g.col, n, g.col_count.get());
      MG_LAUNCH_CHECK();
    }
  }
  MG_HIP(hipStreamSynchronize(st));  // bits is freed on return
  g.has_jobs = true;
}

__global__ void __launch_bounds__(kBlock) relabel_keys_kernel(const uint32_t* __restrict__ row,
                                                              const uint32_t* __restrict__ col,
                                                              const uint32_t* __restrict__ map, uint32_t n,
                                                              uint64_t* __restrict__ keys,
                                                              uint32_t* __restrict__ vals) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    keys[i] = (static_cast<uint64_t>(map[row[i]]) << 32) | map[col[i]];
    vals[i] = i;
  }
}

__global__ void __launch_bounds__(kBlock) split_keys_kernel(const uint64_t* __restrict__ keys, uint32_t n,
                                                            uint32_t* __restrict__ row, uint32_t* __restrict__ col) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    row[i] = static_cast<uint32_t>(keys[i] >> 32);
    col[i] = static_cast<uint32_t>(keys[i]);
  }
}

// The engine's strip permutation for n strips: a Fisher-Yates shuffle driven by SplitMix64
// with a fixed key, so a graph always gets the same order.
std::vector<uint32_t> strip_permutation(uint32_t n) {
  std::vector<uint32_t> p(n);
  std::iota(p.begin(), p.end(), 0u);
  const uint64_t key = stream_key(0x5354524950ull /* "STRIP" */, n);
  for (uint32_t i = n; i > 1; --i) {
    const uint32_t j = static_cast<uint32_t>((sm64_at(key, i) >> 11) % i);
    std::swap(p[i - 1], p[j]);
  }
  return p;
}

// A^T's index in engine order: tile (y, t) becomes (map[y], map[t]); the tiles are sorted by
// the new (row, column) (radix sort of 64-bit keys with the old positions as values). The
// row pointer follows from the old one on the host. The tile data is not moved.
void relabel_index(GraphPlan& g, const std::vector<uint32_t>& off, std::vector<uint32_t>* r_off_host,
                   hipStream_t st) {
  const mg_tiled_matrix* at = g.at;
  const uint32_t ns = at->n_strips, n = at->n_nz_tile;
  g.map = strip_permutation(ns);
  std::vector<uint32_t> inv(ns);
  for (uint32_t y = 0; y < ns; ++y) inv[g.map[y]] = y;
  r_off_host->assign(ns + 1ull, 0);
  for (uint32_t yp = 0; yp < ns; ++yp) (*r_off_host)[yp + 1] = (*r_off_host)[yp] + (off[inv[yp] + 1] - off[inv[yp]]);
  g.d_map.alloc(std::max<uint32_t>(1, ns));
  if (ns) MG_HIP(hipMemcpyAsync(g.d_map.get(), g.map.data(), sizeof(uint32_t) * ns, hipMemcpyHostToDevice, st));
  g.r_off.alloc(ns + 1ull);
  MG_HIP(hipMemcpyAsync(g.r_off.get(), r_off_host->data(), sizeof(uint32_t) * (ns + 1ull), hipMemcpyHostToDevice,
                        st));
  g.r_col.alloc(std::max<uint32_t>(1, n));
  g.r_row.alloc(std::max<uint32_t>(1, n));
  g.r_orig.alloc(std::max<uint32_t>(1, n));
  if (n) {
    MG_REQUIRE(n < (1u << 31), "walk: more than 2^31 tiles");
    DevBuf<uint64_t> keys(n), keys_out(n);
    DevBuf<uint32_t> vals(n);
    relabel_keys_kernel<<<grid_for(n, kBlock), kBlock, 0, st>>>(at->row_idx.get(), at->col_idx.get(), g.d_map.get(), n,
                                                              keys.get(), vals.get());
    MG_LAUNCH_CHECK();
    int bits = 1;
    while ((1ull << bits) < ns) ++bits;
    size_t temp_bytes = 0;
    MG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys.get(), keys_out.get(), vals.get(),
                                              g.r_orig.get(), static_cast<int>(n), 0, 32 + bits, st));
    DevBuf<uint8_t> temp(temp_bytes ? temp_bytes : 1);
    MG_HIP(hipcub::DeviceRadixSort::SortPairs(temp.get(), temp_bytes, keys.get(), keys_out.get(), vals.get(),
                                              g.r_orig.get(), static_cast<int>(n), 0, 32 + bits, st));
    split_keys_kernel<<<grid_for(n, kBlock), kBlock, 0, st>>>(keys_out.get(), n, g.r_row.get(), g.r_col.get());
    MG_LAUNCH_CHECK();
    MG_HIP(hipStreamSynchronize(st));  // the temporaries are freed on return
  }
  g.off = g.r_off.get();
  g.col = g.r_col.get();
  g.row = g.r_row.get();
  g.orig = g.r_orig.get();
}

// Every launch and copy is ordered on `st`, the caller's context stream, so a matrix still
// being filled by earlier asynchronous work on that stream is complete before it is read.
// relabel: walk the graph in the engine's strip order (see GraphPlan).
GraphPlan plan_graph(const mg_tiled_matrix* at, uint32_t chunk_tiles, hipStream_t st, bool relabel) {
  GraphPlan g;
  g.at = at;
  g.n_strips = at->n_strips;
  std::vector<uint32_t> off(at->n_strips + 1ull);
  MG_HIP(hipMemcpyAsync(off.data(), at->offset_row.get(), sizeof(uint32_t) * off.size(), hipMemcpyDeviceToHost, st));
  MG_HIP(hipStreamSynchronize(st));
  g.off = at->offset_row.get();
  g.col = at->col_idx.get();
  g.row = at->row_idx.get();
  if (relabel && at->n_strips > 1) {
    std::vector<uint32_t> r_off;
    relabel_index(g, off, &r_off, st);
    off.swap(r_off);
  }
  g.chunks = build_chunks(off, at->n_strips, chunk_tiles, &g.split);
  g.rmask.alloc(std::max<uint32_t>(1, at->n_nz_tile));
  g.payload.alloc(std::max<uint32_t>(1, at->n_nz_tile));
  if (at->n_nz_tile) {
    row_mask_kernel<<<grid_for(at->n_nz_tile, kWaves, 65536), kBlock, 0, st>>>(at->data.get(), g.orig,
                                                                               at->n_nz_tile, g.rmask.get());
    MG_LAUNCH_CHECK();
    tile_payload_kernel<<<grid_for(at->n_nz_tile, kWaves, 65536), kBlock, 0, st>>>(at->data.get(), g.orig,
                                                                                   at->n_nz_tile, g.payload.get());
    MG_LAUNCH_CHECK();
    const uint32_t n = at->n_nz_tile;
    DevBuf<uint32_t> flag(n), pos(n + 1ull);
    dense_flag_kernel<<<grid_for(n, kBlock), kBlock, 0, st>>>(g.payload.get(), n, flag.get());
    MG_LAUNCH_CHECK();
    exclusive_scan_u32_on(st, flag.get(), pos.get(), n);
    uint32_t n_dense = 0;
    MG_HIP(hipMemcpyAsync(&n_dense, pos.get() + n, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    MG_HIP(hipStreamSynchronize(st));
    g.dense.alloc(std::max<uint64_t>(1, static_cast<uint64_t>(n_dense) * kTile));
    if (n_dense) {
      dense_compact_kernel<<<grid_for(n, kWaves, 65536), kBlock, 0, st>>>(at->data.get(), g.orig, n, pos.get(),
                                                                          g.payload.get(), g.dense.get());
      MG_LAUNCH_CHECK();
    }
    // flag and pos are freed on return: their last readers must have finished.
    MG_HIP(hipStreamSynchronize(st));
  }
  g.chunk_cost = chunk_costs(g, st, &g.total_bits);
  return g;
}

// Order within a column is whatever the atomics give: every consumer ORs or sums.
__global__ void __launch_bounds__(kBlock) column_scatter_kernel(const uint32_t* __restrict__ col, uint32_t n,
                                                                uint32_t* __restrict__ cursor,
                                                                uint32_t* __restrict__ out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[atomicAdd(cursor + col[i], 1u)] = i;
}

__global__ void __launch_bounds__(kBlock) column_copy_kernel(const uint32_t* __restrict__ tile, uint32_t n,
                                                             const uint32_t* __restrict__ row,
                                                             const uint64_t* __restrict__ payload,
                                                             uint32_t* __restrict__ out_row,
                                                             uint64_t* __restrict__ out_payload) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const uint32_t i = tile[j];
    out_row[j] = row[i];
    out_payload[j] = payload[i];
  }
}

// Counting sort of the A^T tiles by column (16 bytes per tile on the device).
void build_column_index(GraphPlan& g, hipStream_t st) {
  const mg_tiled_matrix* at = g.at;
  const uint32_t ns = at->n_strips, n = at->n_nz_tile;
  DevBuf<uint32_t> cursor;
  cursor.alloc(std::max<uint32_t>(1, ns));
  MG_HIP(hipMemsetAsync(cursor.get(), 0, sizeof(uint32_t) * std::max<uint32_t>(1, ns), st));
  if (n) {
    column_count_kernel<<<grid_for(n, kBlock), kBlock, 0, st>>>(g.col, n, cursor.get());
    MG_LAUNCH_CHECK();
  }
  std::vector<uint32_t> count(ns);
  if (ns) MG_HIP(hipMemcpyAsync(count.data(), cursor.get(), sizeof(uint32_t) * ns, hipMemcpyDeviceToHost, st));
  MG_HIP(hipStreamSynchronize(st));
  g.host_csc_off.assign(ns + 1ull, 0);
  for (uint32_t t = 0; t < ns; ++t) g.host_csc_off[t + 1] = g.host_csc_off[t] + count[t];
  if (!g.col_count.get()) {
    g.col_count.alloc(std::max<uint32_t>(1, ns));
    if (ns) MG_HIP(hipMemcpyAsync(g.col_count.get(), count.data(), sizeof(uint32_t) * ns, hipMemcpyHostToDevice, st));
  }
  if (ns)
    MG_HIP(hipMemcpyAsync(cursor.get(), g.host_csc_off.data(), sizeof(uint32_t) * ns, hipMemcpyHostToDevice, st));
  g.csc_tile.alloc(std::max<uint32_t>(1, n));
  g.csc_row.alloc(std::max<uint32_t>(1, n));
  g.csc_payload.alloc(std::max<uint32_t>(1, n));
  if (n) {
    column_scatter_kernel<<<grid_for(n, kBlock), kBlock, 0, st>>>(g.col, n, cursor.get(), g.csc_tile.get());
    MG_LAUNCH_CHECK();
    column_copy_kernel<<<grid_for(n, kBlock), kBlock, 0, st>>>(g.csc_tile.get(), n, g.row,
                                                               g.payload.get(), g.csc_row.get(), g.csc_payload.get());
    MG_LAUNCH_CHECK();
  }
  // cursor is freed on return and host_csc_off's upload reads host memory.
  MG_HIP(hipStreamSynchronize(st));
}

}  // namespace

// One panel's device state: frontiers and masks for levels 0..hops.
struct Panel {
  uint32_t graph = 0;
  uint32_t hops = 0;
  uint32_t n_strips = 0;
  uint64_t words_used = 0;  // words of X holding sources
  std::vector<uint64_t> alive;  // alive[k]: words that propagate from X_k to X_{k+1}
  std::vector<DevBuf<uint64_t>> x, s, n, v, z;
  DevBuf<uint64_t> part;  // segmented hop kernel: partial masks, Engine::seg x 3 per strip
};

// Device launch plan for a set of panels over levels 0..max_hops-1.
struct Engine {
  uint32_t W = 64;
  // Hop kernel of the 64-word levels from level seg_from on: 0 = tile-decoding sweep
  // (walk_numeric64_kernel), G = 1, 2, 4 or 8 = job-list kernel with G row segments.
  uint32_t seg = 0;
  uint32_t seg_from = 1;
  std::vector<GraphPlan> graphs;
  std::vector<Panel> panels;
  uint32_t max_hops = 0;
  DevBuf<unsigned long long> counters;  // panels * (max_hops+1) * kCounterStride
  // per level
  std::vector<std::vector<Chunk>> host_chunks;  // launch order of each level
  std::vector<DevBuf<ChunkLog>> logs;           // per-chunk outcomes (when logging)
  bool logging = false;
  std::vector<DevBuf<PanelLevel>> lv;
  std::vector<std::vector<PanelLevel>> host_lv;  // host copies of lv
  std::vector<uint8_t> partial;  // per level: some active panel has dead words (alive_in != all)
  std::vector<DevBuf<Chunk>> chunks;
  std::vector<uint32_t> n_chunks;
  std::vector<DevBuf<StripItem>> split;
  std::vector<uint32_t> n_split;
  DevBuf<ZeroItem> zero_items;  // the split strips of every level
  uint32_t n_zero_items = 0;
  // X_0 setup
  DevBuf<PanelInit> inits;
  DevBuf<uint32_t> panel_strips;
  DevBuf<FillRange> fills;
  uint32_t n_fills = 0;
  DevBuf<SourceItem> sources;
  uint32_t n_sources = 0;
  int prune = 0;
  // First hop by push: panels whose level 0 is pushed (their level-0 chunks go to push_chunks,
  // which only the structural pass sweeps), their column items and the panel list.
  std::vector<uint8_t> push0;
  DevBuf<Chunk> push_chunks;
  uint32_t n_push_chunks = 0;
  bool push_s_from_columns = false;
  DevBuf<PushItem> push_items;
  uint32_t n_push_items = 0;
  DevBuf<uint32_t> push_panels;
  uint32_t n_push_panels = 0;
  uint32_t push_max_strips = 0;

  unsigned long long* ctr(uint32_t p, uint32_t level) {
    return counters.get() + (static_cast<size_t>(p) * (max_hops + 1) + level) * kCounterStride;
  }
};

namespace {

// Allocate frontiers, build the per-level descriptor tables, chunk lists and items.
// All copies and fills are ordered on `st` (the context stream); host sources stay alive
// until the stream has consumed them.
void build_engine(Engine* e, hipStream_t st) {
  const uint32_t P = static_cast<uint32_t>(e->panels.size());
  e->max_hops = 0;
  for (auto& p : e->panels) e->max_hops = std::max(e->max_hops, p.hops);
  for (auto& p : e->panels) {
    const uint64_t nv = static_cast<uint64_t>(p.n_strips) * 64;
    p.x.resize(p.hops + 1);
    p.s.resize(p.hops + 1);
    p.n.resize(p.hops + 1);
    p.v.resize(p.hops + 1);
    p.z.resize(p.hops + 1);
    for (uint32_t k = 0; k <= p.hops; ++k) {
      p.x[k].alloc((nv + 1) * e->W);  // plus one all-zero vertex row (drain padding reads it)
      p.s[k].alloc(p.n_strips);
      p.n[k].alloc(p.n_strips);
      p.v[k].alloc(p.n_strips);
      p.z[k].alloc(p.n_strips);
      // All strips start zero and known to be zero (Z and N masks 0); later runs rely on
      // that to skip stores of words that stay zero.
      MG_HIP(hipMemsetAsync(p.x[k].get(), 0, p.x[k].bytes(), st));
      MG_HIP(hipMemsetAsync(p.n[k].get(), 0, p.n[k].bytes(), st));
      MG_HIP(hipMemsetAsync(p.s[k].get(), 0, p.s[k].bytes(), st));
      MG_HIP(hipMemsetAsync(p.v[k].get(), 0, p.v[k].bytes(), st));
      MG_HIP(hipMemsetAsync(p.z[k].get(), 0, p.z[k].bytes(), st));
    }
  }
  e->counters.alloc(static_cast<size_t>(P) * (e->max_hops + 1) * kCounterStride);
  e->lv.resize(e->max_hops);
  e->host_lv.clear();
  e->partial.clear();
  e->chunks.resize(e->max_hops);
  e->host_chunks.assign(e->max_hops, {});
  e->n_chunks.assign(e->max_hops, 0);
  e->split.resize(e->max_hops);
  e->n_split.assign(e->max_hops, 0);
  std::vector<ZeroItem> zero_items;
  for (uint32_t k = 0; k < e->max_hops; ++k) {
    std::vector<PanelLevel> lv(P);
    std::vector<Chunk> ch, pch;
    std::vector<uint64_t> cost;  // per entry of ch
    std::vector<StripItem> sp;
    for (uint32_t pi = 0; pi < P; ++pi) {
      Panel& p = e->panels[pi];
      if (k >= p.hops) continue;
      const GraphPlan& g = e->graphs[p.graph];
      PanelLevel& L = lv[pi];
      L.off = g.off;
      L.col = g.col;
      L.row = g.row;
      L.orig = g.orig;
      L.data = g.at->data.get();
      L.rmask = g.rmask.get();
      L.payload = g.payload.get();
      L.dense = g.dense.get();
      L.csc_tile = g.csc_tile.get();
      L.col_count = g.col_count.get();
      L.csc_row = g.csc_row.get();
      L.csc_payload = g.csc_payload.get();
      L.x_in = p.x[k].get();
      L.v_in = p.v[k].get();
      L.s_in = e->prune ? p.n[k].get() : p.s[k].get();
      L.x_out = p.x[k + 1].get();
      L.n_out = p.n[k + 1].get();
      L.v_out = p.v[k + 1].get();
      L.z_out = p.z[k + 1].get();
      L.s_out = p.s[k + 1].get();
      L.counters = e->ctr(pi, k + 1);
      L.alive_in = p.alive[k];
      L.alive_next = (k + 1 < p.hops) ? p.alive[k + 1] : 0;
      L.n_strips = p.n_strips;
      const bool pushed = k == 0 && pi < e->push0.size() && e->push0[pi];
      for (size_t i = 0; i < g.chunks.size(); ++i) {
        Chunk c = g.chunks[i];
        c.panel = pi;
        if (pushed) {
          pch.push_back(c);
        } else {
          ch.push_back(c);
          cost.push_back(g.chunk_cost[i]);
        }
      }
      if (!pushed)
        for (uint32_t y : g.split) {
          sp.push_back({pi, y});
          zero_items.push_back({p.x[k + 1].get() + static_cast<uint64_t>(y) * 64 * e->W, p.s[k + 1].get() + y});
        }
    }
    // Heaviest chunks first (chunk_costs) so the tail of the launch is made of light ones.
    // Among equal costs, the same chunk of every panel of one graph is adjacent, so the
    // panels' reads of those A^T tiles follow each other closely and the repeats are served
    // from cache.
    {
      std::vector<uint32_t> idx(ch.size());
      std::iota(idx.begin(), idx.end(), 0u);
      std::sort(idx.begin(), idx.end(), [&](uint32_t i, uint32_t j) {
        const Chunk &a = ch[i], &b = ch[j];
        if (cost[i] != cost[j]) return cost[i] > cost[j];
        const uint32_t ga = e->panels[a.panel].graph, gb = e->panels[b.panel].graph;
        if (ga != gb) return ga < gb;
        if (a.strip_begin != b.strip_begin) return a.strip_begin < b.strip_begin;
        if (a.tile_begin != b.tile_begin) return a.tile_begin < b.tile_begin;
        return a.panel < b.panel;
      });
      std::vector<Chunk> sorted(ch.size());
      for (size_t i = 0; i < idx.size(); ++i) sorted[i] = ch[idx[i]];
      ch.swap(sorted);
    }
    e->lv[k].alloc(P);
    MG_HIP(hipMemcpyAsync(e->lv[k].get(), lv.data(), sizeof(PanelLevel) * P, hipMemcpyHostToDevice, st));
    e->host_lv.push_back(lv);
    bool partial = false;
    for (const auto& L : lv) partial |= L.x_in && L.alive_in != ~0ull;
    e->partial.push_back(partial ? 1 : 0);
    e->n_chunks[k] = static_cast<uint32_t>(ch.size());
    e->host_chunks[k] = ch;
    e->chunks[k].alloc(std::max<size_t>(1, ch.size()));
    if (!ch.empty()) MG_HIP(hipMemcpyAsync(e->chunks[k].get(), ch.data(), sizeof(Chunk) * ch.size(), hipMemcpyHostToDevice, st));
    if (k == 0) {
      e->n_push_chunks = static_cast<uint32_t>(pch.size());
      e->push_chunks.alloc(std::max<size_t>(1, pch.size()));
      if (!pch.empty())
        MG_HIP(hipMemcpyAsync(e->push_chunks.get(), pch.data(), sizeof(Chunk) * pch.size(), hipMemcpyHostToDevice, st));
    }
    e->n_split[k] = static_cast<uint32_t>(sp.size());
    e->split[k].alloc(std::max<size_t>(1, sp.size()));
    if (!sp.empty()) MG_HIP(hipMemcpyAsync(e->split[k].get(), sp.data(), sizeof(StripItem) * sp.size(), hipMemcpyHostToDevice, st));
    MG_HIP(hipStreamSynchronize(st));  // lv, ch, pch and sp are about to go out of scope
  }
  e->n_zero_items = static_cast<uint32_t>(zero_items.size());
  e->zero_items.alloc(std::max<size_t>(1, zero_items.size()));
  if (!zero_items.empty())
    MG_HIP(hipMemcpyAsync(e->zero_items.get(), zero_items.data(), sizeof(ZeroItem) * zero_items.size(),
                          hipMemcpyHostToDevice, st));
  // X_0 fills: frontier and masks of every panel, plus all counters.
  std::vector<PanelInit> inits(P);
  std::vector<uint32_t> strips(P);
  std::vector<FillRange> fills;
  for (uint32_t pi = 0; pi < P; ++pi) {
    Panel& p = e->panels[pi];
    inits[pi] = {p.x[0].get(), p.s[0].get(), p.v[0].get(), p.n[0].get()};
    strips[pi] = p.n_strips;
  }
  fills.push_back({reinterpret_cast<uint64_t*>(e->counters.get()), e->counters.size()});
  e->inits.alloc(P);
  MG_HIP(hipMemcpyAsync(e->inits.get(), inits.data(), sizeof(PanelInit) * P, hipMemcpyHostToDevice, st));
  e->panel_strips.alloc(P);
  MG_HIP(hipMemcpyAsync(e->panel_strips.get(), strips.data(), sizeof(uint32_t) * P, hipMemcpyHostToDevice, st));
  e->n_fills = static_cast<uint32_t>(fills.size());
  e->fills.alloc(fills.size());
  MG_HIP(hipMemcpyAsync(e->fills.get(), fills.data(), sizeof(FillRange) * fills.size(), hipMemcpyHostToDevice, st));
  MG_HIP(hipStreamSynchronize(st));
}

// Selects the hop kernel of the 64-word levels (Engine::seg): builds the job lists of the
// graphs and the panels' partial-mask buffers on first use and points the level tables at
// them. G = 0 keeps the tile-decoding sweep. Not called while a run may be in flight.
void apply_seg(Engine* e, uint32_t G, hipStream_t st) {
  MG_REQUIRE(G == 0 || G == 1 || G == 2 || G == 4 || G == 8, "walk: row segments must be 0, 1, 2, 4 or 8");
  if (G && e->W != 64) G = 0;  // narrower panels keep walk_numeric_kernel<W>
  if (G)
    for (auto& g : e->graphs)
      if (!job_list_fits(g)) G = 0;
  if (G) {
    for (auto& g : e->graphs) build_job_list(g, st);
    for (auto& p : e->panels) {
      if (G == 1) break;  // the one-segment kernel writes the masks itself
      const uint64_t need = static_cast<uint64_t>(G) * 3 * std::max<uint32_t>(1, p.n_strips);
      if (p.part.size() < need) {
        p.part.release();
        p.part.alloc(need);
      }
    }
  }
  e->seg = G;
  for (uint32_t k = 0; k < e->host_lv.size(); ++k) {
    auto& lv = e->host_lv[k];
    for (uint32_t pi = 0; pi < lv.size(); ++pi) {
      if (!lv[pi].x_in) continue;  // panel not active at this level
      const GraphPlan& g = e->graphs[e->panels[pi].graph];
      lv[pi].jobs = g.has_jobs ? g.jobs.get() : nullptr;
      lv[pi].job_off = g.has_jobs ? g.job_off.get() : nullptr;
      lv[pi].col_count = g.col_count.get();
      lv[pi].part = e->panels[pi].part.get();
    }
    MG_HIP(hipMemcpyAsync(e->lv[k].get(), lv.data(), sizeof(PanelLevel) * lv.size(), hipMemcpyHostToDevice, st));
  }
  MG_HIP(hipStreamSynchronize(st));
}

// Whether walks and batches use the engine's strip order (GraphPlan): yes unless the caller's
// flag asks for the vertex order, or MG_WALK_RELABEL=0 is set in the environment.
bool use_relabel(bool identity_flag) {
  if (identity_flag) return false;
  const char* env = std::getenv("MG_WALK_RELABEL");
  return !(env && std::atoi(env) == 0);
}

// Default hop kernel of 64-word levels (kDefaultSegments); MG_WALK_SEGMENTS (0, 1, 2, 4 or 8)
// overrides it.
#ifndef MG_DEFAULT_SEGMENTS
#define MG_DEFAULT_SEGMENTS 1
#endif
constexpr uint32_t kDefaultSegments = MG_DEFAULT_SEGMENTS;
uint32_t default_segments() {
  const char* env = std::getenv("MG_WALK_SEGMENTS");
  return env ? static_cast<uint32_t>(std::atoi(env)) : kDefaultSegments;
}

template <int W>
void launch_level_w(hipStream_t st, Engine* e, uint32_t k, hipEvent_t ev[4]) {
  // The structural pass runs inside the numeric kernel: its event pair brackets nothing.
  if (ev) MG_HIP(hipEventRecord(ev[0], st));
  if (ev) MG_HIP(hipEventRecord(ev[1], st));
  if (ev) MG_HIP(hipEventRecord(ev[2], st));
  const uint32_t G = (W == 64 && e->seg && !e->logging && k >= e->seg_from) ? e->seg : 0;
  if (e->n_chunks[k]) {
    ChunkLog* const log = e->logging ? e->logs[k].get() : nullptr;
    if constexpr (W == 64) {
      const uint32_t nb = e->n_chunks[k];
      switch (G) {
        case 0: walk_numeric64_kernel<<<nb, kNumBlock, 0, st>>>(e->chunks[k].get(), e->lv[k].get(), log); break;
        case 1:
          if (e->partial[k])
            walk_lean_kernel<true><<<nb, kLeanBlock, 0, st>>>(e->chunks[k].get(), e->lv[k].get());
          else
            walk_lean_kernel<false><<<nb, kLeanBlock, 0, st>>>(e->chunks[k].get(), e->lv[k].get());
          break;
        case 2: walk_seg_kernel<2><<<nb * 2, kSegBlock, 0, st>>>(e->chunks[k].get(), e->lv[k].get()); break;
        case 4: walk_seg_kernel<4><<<nb * 4, kSegBlock, 0, st>>>(e->chunks[k].get(), e->lv[k].get()); break;
        case 8: walk_seg_kernel<8><<<nb * 8, kSegBlock, 0, st>>>(e->chunks[k].get(), e->lv[k].get()); break;
        default: fail(MG_ERR_INTERNAL, "walk: bad segment count");
      }
    } else {
      walk_numeric_kernel<W><<<e->n_chunks[k], kNumBlock, 0, st>>>(e->chunks[k].get(), e->lv[k].get(), log);
    }
    MG_LAUNCH_CHECK();
  }
  if (ev) MG_HIP(hipEventRecord(ev[3], st));
  if (G > 1) {
    const uint32_t n_panels = static_cast<uint32_t>(e->panels.size());
    uint32_t max_strips = 1;
    for (const auto& p : e->panels) max_strips = std::max(max_strips, p.n_strips);
    seg_combine_kernel<<<dim3(grid_for(max_strips, kBlock, 1024), n_panels), kBlock, 0, st>>>(e->lv[k].get(), G);
    MG_LAUNCH_CHECK();
  }
  if (e->n_split[k]) {
    finalize_strips_kernel<W><<<e->n_split[k], kBlock, 0, st>>>(e->split[k].get(), e->lv[k].get());
    MG_LAUNCH_CHECK();
  }
}

void launch_level(hipStream_t st, Engine* e, uint32_t k, hipEvent_t ev[4]) {
  switch (e->W) {
    case 1: launch_level_w<1>(st, e, k, ev); break;
    case 2: launch_level_w<2>(st, e, k, ev); break;
    case 4: launch_level_w<4>(st, e, k, ev); break;
    case 8: launch_level_w<8>(st, e, k, ev); break;
    case 16: launch_level_w<16>(st, e, k, ev); break;
    case 32: launch_level_w<32>(st, e, k, ev); break;
    case 64: launch_level_w<64>(st, e, k, ev); break;
    default: fail(MG_ERR_INTERNAL, "walk: unsupported words per vertex " + std::to_string(e->W));
  }
}

// First hop by push for the panels listed in `panels` (level 0): `items` are their source
// columns, `chunks` their level-0 scheduler chunks (swept by push_structural unless
// s_from_columns); all already on the device. max_strips bounds the strips of any listed panel.
struct PushLaunch {
  const PushItem* items;
  uint32_t n_items;
  const Chunk* chunks;
  uint32_t n_chunks;
  const uint32_t* panels;
  uint32_t n_panels;
  uint32_t max_strips;
  bool s_from_columns;
};

// S_1 from the source columns (atomics, one per missing bit) pays while the pushed columns hold
// at most this fraction of the panels' tiles; above it the structural sweep of all tiles does.
constexpr uint64_t kPushColumnS = 8;  // pushed tiles * kPushColumnS <= tiles

template <int W>
void launch_push_first_hop_w(hipStream_t st, Engine* e, const PushLaunch& pl, hipEvent_t ev[4]) {
  if (ev) {
    MG_HIP(hipEventRecord(ev[0], st));
    MG_HIP(hipEventRecord(ev[1], st));
    MG_HIP(hipEventRecord(ev[2], st));
  }
  if (pl.n_panels) {
    push_reset_kernel<W><<<dim3(grid_for(pl.max_strips, kWaves, 4096), pl.n_panels), kBlock, 0, st>>>(e->lv[0].get(),
                                                                                                    pl.panels);
    MG_LAUNCH_CHECK();
  }
  if (!pl.s_from_columns && pl.n_chunks) {
    push_structural_kernel<<<pl.n_chunks, kBlock, 0, st>>>(pl.chunks, e->lv[0].get());
    MG_LAUNCH_CHECK();
  }
  if (pl.n_items) {
    push_first_hop_kernel<W><<<pl.n_items, kBlock, 0, st>>>(pl.items, e->lv[0].get(), pl.s_from_columns);
    MG_LAUNCH_CHECK();
  }
  if (pl.n_panels) {
    push_finalize_kernel<W><<<dim3(grid_for(pl.max_strips, kWaves, 16384), pl.n_panels), kBlock, 0, st>>>(
        e->lv[0].get(), pl.panels, pl.s_from_columns);
    MG_LAUNCH_CHECK();
  }
  if (ev) MG_HIP(hipEventRecord(ev[3], st));
}

void launch_push_first_hop(hipStream_t st, Engine* e, const PushLaunch& pl, hipEvent_t ev[4]) {
  switch (e->W) {
    case 1: launch_push_first_hop_w<1>(st, e, pl, ev); break;
    case 2: launch_push_first_hop_w<2>(st, e, pl, ev); break;
    case 4: launch_push_first_hop_w<4>(st, e, pl, ev); break;
    case 8: launch_push_first_hop_w<8>(st, e, pl, ev); break;
    case 16: launch_push_first_hop_w<16>(st, e, pl, ev); break;
    case 32: launch_push_first_hop_w<32>(st, e, pl, ev); break;
    case 64: launch_push_first_hop_w<64>(st, e, pl, ev); break;
    default: fail(MG_ERR_INTERNAL, "walk: unsupported words per vertex " + std::to_string(e->W));
  }
}

// Column items of one panel's source strips (sorted, distinct), cut into kPushItemTiles
// positions; returns the number of A^T tiles in those columns.
uint64_t push_items_for(const GraphPlan& g, uint32_t panel, const std::vector<uint32_t>& strips,
                        std::vector<PushItem>* out) {
  const std::vector<uint32_t>& off = g.host_csc_off;
  uint64_t tiles = 0;
  for (uint32_t t : strips) {
    tiles += off[t + 1] - off[t];
    if (out)
      for (uint32_t b = off[t]; b < off[t + 1]; b += kPushItemTiles)
        out->push_back({t, b, std::min(off[t + 1], b + kPushItemTiles), panel});
  }
  return tiles;
}

// X_0 fills + source bits (sources already on the device).
template <int W>
void launch_clear_x0(hipStream_t st, Engine* e) {
  clear_x0_kernel<W><<<2048, kBlock, 0, st>>>(e->inits.get(), e->panel_strips.get(),
                                              static_cast<uint32_t>(e->panels.size()));
  MG_LAUNCH_CHECK();
  if (e->n_zero_items) {
    zero_split_kernel<W><<<e->n_zero_items, kBlock, 0, st>>>(e->zero_items.get());
    MG_LAUNCH_CHECK();
  }
}

void launch_init(hipStream_t st, Engine* e, uint32_t n_sources) {
  fill_zero_kernel<<<64, kBlock, 0, st>>>(e->fills.get(), e->n_fills);
  MG_LAUNCH_CHECK();
  switch (e->W) {
    case 1: launch_clear_x0<1>(st, e); break;
    case 2: launch_clear_x0<2>(st, e); break;
    case 4: launch_clear_x0<4>(st, e); break;
    case 8: launch_clear_x0<8>(st, e); break;
    case 16: launch_clear_x0<16>(st, e); break;
    case 32: launch_clear_x0<32>(st, e); break;
    case 64: launch_clear_x0<64>(st, e); break;
    default: fail(MG_ERR_INTERNAL, "walk: unsupported words per vertex");
  }
  if (n_sources) {
    init_sources_kernel<<<grid_for(n_sources, 256), 256, 0, st>>>(e->sources.get(), n_sources, e->inits.get(), e->W);
    MG_LAUNCH_CHECK();
  }
}

uint32_t words_for(uint32_t n_sources) {
  uint32_t W = 1;
  while (W * 64 < n_sources) W <<= 1;
  return W;
}

// c: the kCounterSlots slots of one (panel, level).
void hop_stats_from(const unsigned long long* c_slots, uint32_t rows, mg_hop_stats* out) {
  unsigned long long c[kNumCounters] = {};
  for (uint32_t sl = 0; sl < kCounterSlots; ++sl)
    for (uint32_t i = 0; i < kNumCounters; ++i) c[i] += c_slots[sl * kNumCounters + i];
  std::memset(out, 0, sizeof(*out));
  out->tile_products = c[kTP];
  out->left_tiles = c[kLeft];
  out->right_tiles_used = c[kRightUsed];
  out->out_tiles = c[kOut];
  out->out_tiles_numeric = c[kOutNumeric];
  out->popcount = c[kPop];
  out->edges_visited = c[kEdges];
  out->alg_bytes = 520ull * (c[kLeft] + c[kRightUsed] + c[kOut]) + 4ull * (2ull * rows + 2);
}

uint64_t words_mask(uint32_t w0, uint32_t n) {
  if (n >= 64) return ~0ull;
  return ((1ull << n) - 1) << w0;
}

}  // namespace
}  // namespace mgamd

using namespace mgamd;

// ----------------------------------------------------------------------------
// Single walks: one panel.
// ----------------------------------------------------------------------------
struct mg_walk {
  int device = 0;
  Engine e;
  uint32_t max_sources = 0;
  uint32_t flags = 0;
  uint32_t n_sources = 0;  // of the last run
  uint32_t hops = 0;       // of the last run
  bool timing = false;
  PinnedBuf<SourceItem> staging;
  // First hop by push: per-run items (source strips' column ranges) and their staging.
  PinnedBuf<PushItem> push_staging;
  DevBuf<PushItem> push_items;
  uint32_t max_push_items = 0;
  bool last_push = false;              // the last run's first hop was pushed
  hipEvent_t staging_read = nullptr;   // the last run's uploads have left the staging buffers
  // Timing log: per timed run, 4 events per hop (symbolic start/end, numeric kernel
  // start/end), kept for every run since timing was enabled.
  std::vector<std::vector<hipEvent_t>> run_events;
  std::vector<hipEvent_t> event_pool;
  ~mg_walk() {
    for (auto& r : run_events)
      for (auto ev : r) (void)hipEventDestroy(ev);
    for (auto ev : event_pool) (void)hipEventDestroy(ev);
    if (staging_read) (void)hipEventDestroy(staging_read);
  }
  hipEvent_t take_event() {
    if (event_pool.empty()) {
      hipEvent_t ev;
      MG_HIP(hipEventCreate(&ev));
      return ev;
    }
    hipEvent_t ev = event_pool.back();
    event_pool.pop_back();
    return ev;
  }
  void reset_log() {
    for (auto& r : run_events)
      for (auto ev : r) event_pool.push_back(ev);
    run_events.clear();
  }
};

extern "C" int mg_walk_create(mg_context* ctx, const mg_tiled_matrix* at, uint32_t max_sources, uint32_t max_hops,
                              uint32_t flags, mg_walk** out) {
  return guarded([&] {
    MG_REQUIRE(ctx && at && out, "mg_walk_create: NULL argument");
    *out = nullptr;
    MG_REQUIRE(max_sources >= 1 && max_sources <= 4096, "mg_walk_create: max_sources must be in [1, 4096]");
    MG_REQUIRE(max_hops >= 1 && max_hops <= 64, "mg_walk_create: max_hops must be in [1, 64]");
    MG_REQUIRE((flags & ~(MG_WALK_PRUNE | MG_WALK_SWEEP_FIRST_HOP | MG_WALK_VERTEX_ORDER)) == 0,
               "mg_walk_create: unknown flags");
    DeviceGuard g(ctx->device);
    auto w = std::make_unique<mg_walk>();
    w->device = ctx->device;
    w->max_sources = max_sources;
    w->flags = flags;
    Engine& e = w->e;
    e.W = words_for(max_sources);
    e.prune = (flags & MG_WALK_PRUNE) ? 1 : 0;
    e.graphs.push_back(plan_graph(at, kDefaultChunkTiles, ctx->stream, use_relabel(flags & MG_WALK_VERTEX_ORDER)));
    Panel p;
    p.graph = 0;
    p.hops = max_hops;
    p.n_strips = at->n_strips;
    p.words_used = e.W;
    p.alive.assign(max_hops, words_mask(0, e.W));
    e.panels.push_back(std::move(p));
    // The column index first: the level tables built next point at it.
    if (!(flags & MG_WALK_SWEEP_FIRST_HOP)) build_column_index(e.graphs[0], ctx->stream);
    build_engine(&e, ctx->stream);
    apply_seg(&e, default_segments(), ctx->stream);
    e.sources.alloc(max_sources);
    w->staging.alloc(max_sources);
    if (!(flags & MG_WALK_SWEEP_FIRST_HOP)) {
      // Each source strip's column is cut into pieces of kPushItemTiles positions.
      w->max_push_items = max_sources + at->n_nz_tile / kPushItemTiles + 1;
      w->push_items.alloc(w->max_push_items);
      w->push_staging.alloc(w->max_push_items);
      const uint32_t zero = 0;
      e.push_panels.alloc(1);
      MG_HIP(hipMemcpyAsync(e.push_panels.get(), &zero, sizeof(zero), hipMemcpyHostToDevice, ctx->stream));
      MG_HIP(hipStreamSynchronize(ctx->stream));
    }
    MG_HIP(hipEventCreateWithFlags(&w->staging_read, hipEventDisableTiming));
    *out = w.release();
  });
}

extern "C" int mg_walk_set_timing(mg_walk* w, int enable) {
  return guarded([&] {
    MG_REQUIRE(w != nullptr, "mg_walk_set_timing: NULL walk");
    const bool on = enable != 0;
    if (on && !w->timing) w->reset_log();
    w->timing = on;
  });
}

extern "C" int mg_walk_set_hop_kernel(mg_context* ctx, mg_walk* w, uint32_t segments) {
  return guarded([&] {
    MG_REQUIRE(ctx && w, "mg_walk_set_hop_kernel: NULL argument");
    MG_REQUIRE(ctx->device == w->device, "mg_walk_set_hop_kernel: walk belongs to another device");
    DeviceGuard g(ctx->device);
    MG_HIP(hipStreamSynchronize(ctx->stream));  // no run may still read the level tables
    apply_seg(&w->e, segments, ctx->stream);
  });
}

extern "C" int mg_walk_hop_kernel(const mg_walk* w, uint32_t* segments) {
  return guarded([&] {
    MG_REQUIRE(w && segments, "mg_walk_hop_kernel: NULL argument");
    *segments = w->e.seg;
  });
}

extern "C" int mg_walk_run(mg_context* ctx, mg_walk* w, uint32_t n_sources, const uint32_t* sources, uint32_t hops) {
  return guarded([&] {
    MG_REQUIRE(ctx && w, "mg_walk_run: NULL argument");
    MG_REQUIRE(ctx->device == w->device, "mg_walk_run: walk belongs to another device");
    MG_REQUIRE(n_sources >= 1 && n_sources <= w->max_sources, "mg_walk_run: n_sources out of range");
    MG_REQUIRE(hops >= 1 && hops <= w->e.max_hops, "mg_walk_run: hops out of range");
    MG_REQUIRE(sources != nullptr, "mg_walk_run: sources is NULL");
    const uint64_t n_vertices = static_cast<uint64_t>(w->e.panels[0].n_strips) * 64;
    {
      std::vector<uint32_t> sorted(sources, sources + n_sources);
      std::sort(sorted.begin(), sorted.end());
      MG_REQUIRE(sorted.back() < n_vertices, "mg_walk_run: source vertex out of range");
      MG_REQUIRE(std::adjacent_find(sorted.begin(), sorted.end()) == sorted.end(),
                 "mg_walk_run: sources must be distinct");
    }
    DeviceGuard g(ctx->device);
    hipStream_t st = ctx->stream;
    w->n_sources = n_sources;
    w->hops = hops;
    // The previous run's uploads read the staging buffers asynchronously.
    MG_HIP(hipEventSynchronize(w->staging_read));
    SourceItem* stg = w->staging.get();
    const GraphPlan& gp = w->e.graphs[0];
    for (uint32_t i = 0; i < n_sources; ++i) stg[i] = {0, i, gp.engine_vertex(sources[i])};
    MG_HIP(hipMemcpyAsync(w->e.sources.get(), stg, sizeof(SourceItem) * n_sources, hipMemcpyHostToDevice, st));
    // First hop: push from the source strips when their A^T columns hold at most half of the
    // tiles (the sweep reads every tile once; the push reads the source columns).
    uint32_t n_items = 0;
    bool push = false, s_cols = false;
    if (w->max_push_items) {
      std::vector<uint32_t> strips(n_sources);
      for (uint32_t i = 0; i < n_sources; ++i) strips[i] = gp.engine_vertex(sources[i]) / 64;
      std::sort(strips.begin(), strips.end());
      strips.erase(std::unique(strips.begin(), strips.end()), strips.end());
      const GraphPlan& g = w->e.graphs[0];
      const uint64_t pushed_tiles = push_items_for(g, 0, strips, nullptr);
      push = 2 * pushed_tiles <= g.at->n_nz_tile;
      s_cols = kPushColumnS * pushed_tiles <= g.at->n_nz_tile;
      if (push) {
        std::vector<PushItem> items;
        push_items_for(g, 0, strips, &items);
        n_items = static_cast<uint32_t>(items.size());
        MG_REQUIRE(n_items <= w->max_push_items, "mg_walk_run: internal push item overflow");
        std::copy(items.begin(), items.end(), w->push_staging.get());
        if (n_items)
          MG_HIP(hipMemcpyAsync(w->push_items.get(), w->push_staging.get(), sizeof(PushItem) * n_items,
                                hipMemcpyHostToDevice, st));
      }
    }
    MG_HIP(hipEventRecord(w->staging_read, st));
    launch_init(st, &w->e, n_sources);
    std::vector<hipEvent_t> ev;
    if (w->timing) {
      ev.resize(4 * static_cast<size_t>(hops));
      for (auto& x : ev) x = w->take_event();
    }
    for (uint32_t k = 0; k < hops; ++k) {
      hipEvent_t* const evk = w->timing ? &ev[4 * k] : nullptr;
      if (k == 0 && push)
        launch_push_first_hop(st, &w->e,
                              PushLaunch{w->push_items.get(), n_items, w->e.chunks[0].get(), w->e.n_chunks[0],
                                         w->e.push_panels.get(), 1u, w->e.panels[0].n_strips, s_cols},
                              evk);
      else
        launch_level(st, &w->e, k, evk);
    }
    w->last_push = push;
    if (w->timing) w->run_events.push_back(std::move(ev));
  });
}

extern "C" int mg_walk_hop_stats(mg_context* ctx, mg_walk* w, uint32_t hop, mg_hop_stats* out) {
  return guarded([&] {
    MG_REQUIRE(ctx && w && out, "mg_walk_hop_stats: NULL argument");
    MG_REQUIRE(hop >= 1 && hop <= w->hops, "mg_walk_hop_stats: hop out of range");
    DeviceGuard g(ctx->device);
    unsigned long long c[kCounterStride];
    MG_HIP(hipMemcpyAsync(c, w->e.ctr(0, hop), sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
    MG_HIP(hipStreamSynchronize(ctx->stream));
    hop_stats_from(c, (w->n_sources + 63) / 64, out);
    {
      // Numeric occupancy of X_{hop-1}, from its per-strip numeric masks.
      const GraphPlan& gp = w->e.graphs[0];
      const Panel& pn = w->e.panels[0];
      unsigned long long* d = ctx->counters.get();
      MG_HIP(hipMemsetAsync(d, 0, 2 * sizeof(unsigned long long), ctx->stream));
      const uint32_t nt = gp.at->n_nz_tile;
      numeric_use_kernel<<<grid_for(std::max<uint32_t>(nt, pn.n_strips), kBlock, 4096), kBlock, 0, ctx->stream>>>(
          gp.col, nt, pn.n[hop - 1].get(), pn.n_strips, words_mask(0, (w->n_sources + 63) / 64), d);
      MG_LAUNCH_CHECK();
      unsigned long long h[2] = {0, 0};
      MG_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
      MG_HIP(hipStreamSynchronize(ctx->stream));
      out->right_tiles_used_numeric = h[0];
      out->left_tiles_numeric = h[1];
      out->alg_bytes_numeric =
          520ull * (h[0] + h[1] + out->out_tiles_numeric) + 4ull * (2ull * ((w->n_sources + 63) / 64) + 2);
    }
    if (hop == 1 && w->last_push) {
      out->kernel = MG_HOP_PUSH;
    } else if (w->e.W == 64 && w->e.seg && hop - 1 >= w->e.seg_from) {
      out->kernel = MG_HOP_JOBS;
      out->segments = w->e.seg;
    } else {
      out->kernel = MG_HOP_SWEEP;
    }
    if (!w->run_events.empty() && w->run_events.back().size() >= 4ull * hop) {
      const auto& ev = w->run_events.back();
      float a = 0, b = 0;
      MG_HIP(hipEventElapsedTime(&a, ev[4 * (hop - 1)], ev[4 * (hop - 1) + 1]));
      MG_HIP(hipEventElapsedTime(&b, ev[4 * (hop - 1) + 2], ev[4 * (hop - 1) + 3]));
      out->symbolic_ms = a;
      out->numeric_ms = b;
    }
  });
}

extern "C" int mg_walk_kernel_times(mg_context* ctx, mg_walk* w, uint32_t hop, double* symbolic_ms,
                                    double* numeric_ms, uint32_t* n_runs) {
  return guarded([&] {
    MG_REQUIRE(ctx && w && symbolic_ms && numeric_ms && n_runs, "mg_walk_kernel_times: NULL argument");
    MG_REQUIRE(hop >= 1, "mg_walk_kernel_times: hop must be >= 1");
    DeviceGuard g(ctx->device);
    MG_HIP(hipStreamSynchronize(ctx->stream));
    double a_sum = 0, b_sum = 0;
    uint32_t n = 0;
    for (const auto& ev : w->run_events) {
      if (ev.size() < 4ull * hop) continue;
      float a = 0, b = 0;
      MG_HIP(hipEventElapsedTime(&a, ev[4 * (hop - 1)], ev[4 * (hop - 1) + 1]));
      MG_HIP(hipEventElapsedTime(&b, ev[4 * (hop - 1) + 2], ev[4 * (hop - 1) + 3]));
      a_sum += a;
      b_sum += b;
      ++n;
    }
    *symbolic_ms = a_sum;
    *numeric_ms = b_sum;
    *n_runs = n;
  });
}

extern "C" int mg_walk_words_per_vertex(const mg_walk* w, uint32_t* words) {
  return guarded([&] {
    MG_REQUIRE(w && words, "mg_walk_words_per_vertex: NULL argument");
    *words = w->e.W;
  });
}

extern "C" int mg_walk_frontier_download(mg_context* ctx, const mg_walk* w, uint32_t hop, uint64_t* xt) {
  return guarded([&] {
    MG_REQUIRE(ctx && w && xt, "mg_walk_frontier_download: NULL argument");
    MG_REQUIRE(hop <= w->hops, "mg_walk_frontier_download: hop out of range");
    DeviceGuard g(ctx->device);
    const auto& x = w->e.panels[0].x[hop];
    const uint64_t nv = static_cast<uint64_t>(w->e.panels[0].n_strips) * 64;
    const uint64_t words = nv * w->e.W;
    const uint32_t* map = w->e.graphs[0].device_map();
    if (!map) {
      MG_HIP(hipMemcpyAsync(xt, x.get(), words * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    } else {
      DevBuf<uint64_t> tmp(words);  // X_hop in vertex order
      frontier_slice_kernel<<<grid_for(words, kBlock * 8, 16384), kBlock, 0, ctx->stream>>>(x.get(), nv, w->e.W, 0,
                                                                                           w->e.W, map, tmp.get());
      MG_LAUNCH_CHECK();
      MG_HIP(hipMemcpyAsync(xt, tmp.get(), words * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    }
    MG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

namespace {
uint64_t frontier_digest(mg_context* ctx, const uint64_t* x, uint64_t nv, uint32_t W, uint32_t w0, uint32_t n_words,
                         const uint32_t* map) {
  unsigned long long* d = ctx->counters.get();
  MG_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), ctx->stream));
  if (n_words) {
    frontier_digest_kernel<<<grid_for(nv * n_words, kBlock * 16, 16384), kBlock, 0, ctx->stream>>>(x, nv, W, w0,
                                                                                                   n_words, map, d);
    MG_LAUNCH_CHECK();
  }
  unsigned long long h = 0;
  MG_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  MG_HIP(hipStreamSynchronize(ctx->stream));
  return h;
}

// g: the plan of the frontier's graph (its strip order).
mg_tiled_matrix* export_frontier(mg_context* ctx, const uint64_t* x, const uint64_t* s_dev, uint32_t ns, uint32_t W,
                                 uint32_t w0, uint32_t rows, const GraphPlan& g) {
  std::vector<uint64_t> s_eng(ns), s(ns);
  MG_HIP(hipMemcpyAsync(s_eng.data(), s_dev, sizeof(uint64_t) * ns, hipMemcpyDeviceToHost, ctx->stream));
  MG_HIP(hipStreamSynchronize(ctx->stream));
  for (uint32_t y = 0; y < ns; ++y) s[y] = s_eng[g.map.empty() ? y : g.map[y]];  // vertex strip order
  // Tile index in row-major (x, y) order over the padded-square grid.
  std::vector<uint32_t> off(ns + 1ull, 0), ri, ci;
  for (uint32_t xr = 0; xr < rows; ++xr) {
    for (uint32_t y = 0; y < ns; ++y)
      if ((s[y] >> (w0 + xr)) & 1) {
        ri.push_back(xr);
        ci.push_back(y);
      }
    off[xr + 1] = static_cast<uint32_t>(ri.size());
  }
  for (uint32_t xr = rows; xr < ns; ++xr) off[xr + 1] = off[xr];
  auto m = std::make_unique<mg_tiled_matrix>();
  m->device = ctx->device;
  const uint32_t nt = static_cast<uint32_t>(ri.size());
  alloc_tiled_matrix(m.get(), ns, nt);
  MG_HIP(hipMemcpyAsync(m->offset_row.get(), off.data(), sizeof(uint32_t) * off.size(), hipMemcpyHostToDevice,
                        ctx->stream));
  if (nt) {
    MG_HIP(hipMemcpyAsync(m->row_idx.get(), ri.data(), sizeof(uint32_t) * nt, hipMemcpyHostToDevice, ctx->stream));
    MG_HIP(hipMemcpyAsync(m->col_idx.get(), ci.data(), sizeof(uint32_t) * nt, hipMemcpyHostToDevice, ctx->stream));
    export_tiles_kernel<<<grid_for(nt, kWaves, 65536), kBlock, 0, ctx->stream>>>(m->row_idx.get(), m->col_idx.get(),
                                                                                 nt, x, W, w0, g.device_map(),
                                                                                 m->data.get());
    MG_LAUNCH_CHECK();
  }
  MG_HIP(hipStreamSynchronize(ctx->stream));
  return m.release();
}
}  // namespace

extern "C" int mg_walk_frontier_digest(mg_context* ctx, const mg_walk* w, uint32_t hop, uint64_t* digest) {
  return guarded([&] {
    MG_REQUIRE(ctx && w && digest, "mg_walk_frontier_digest: NULL argument");
    MG_REQUIRE(hop <= w->hops, "mg_walk_frontier_digest: hop out of range");
    DeviceGuard g(ctx->device);
    const Panel& p = w->e.panels[0];
    *digest = frontier_digest(ctx, p.x[hop].get(), static_cast<uint64_t>(p.n_strips) * 64, w->e.W, 0,
                              (w->n_sources + 63) / 64, w->e.graphs[0].device_map());
  });
}

extern "C" int mg_walk_export(mg_context* ctx, const mg_walk* w, uint32_t hop, mg_tiled_matrix** out) {
  return guarded([&] {
    MG_REQUIRE(ctx && w && out, "mg_walk_export: NULL argument");
    *out = nullptr;
    MG_REQUIRE(hop <= w->hops, "mg_walk_export: hop out of range");
    DeviceGuard g(ctx->device);
    const Panel& p = w->e.panels[0];
    *out = export_frontier(ctx, p.x[hop].get(), p.s[hop].get(), p.n_strips, w->e.W, 0, (w->n_sources + 63) / 64,
                           w->e.graphs[0]);
  });
}

extern "C" int mg_walk_destroy(mg_walk* w) {
  return guarded([&] {
    if (!w) return;
    DeviceGuard g(w->device);
    delete w;
  });
}

// ----------------------------------------------------------------------------
// Batched walks: many queries over several graphs, packed into 4096-source panels,
// all hop levels of all panels issued as one captured HIP graph.
// ----------------------------------------------------------------------------
struct QueryPlace {
  uint32_t panel = 0, w0 = 0, n_words = 0, n_sources = 0, hops = 0, graph = 0;
};

struct mg_walk_batch {
  int device = 0;
  Engine e;
  std::vector<QueryPlace> queries;
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int captured_on = -1;
  ~mg_walk_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

namespace {

void record_batch(hipStream_t st, mg_walk_batch* b) {
  Engine& e = b->e;
  launch_init(st, &e, e.n_sources);
  for (uint32_t k = 0; k < e.max_hops; ++k) {
    if (k == 0 && e.n_push_panels)
      launch_push_first_hop(st, &e,
                            PushLaunch{e.push_items.get(), e.n_push_items, e.push_chunks.get(), e.n_push_chunks,
                                       e.push_panels.get(), e.n_push_panels, e.push_max_strips,
                                       e.push_s_from_columns},
                            nullptr);
    launch_level(st, &e, k, nullptr);  // the panels whose first hop is swept
  }
}

}  // namespace

extern "C" int mg_walk_batch_create(mg_context* ctx, const mg_tiled_matrix* const* graphs_at, uint32_t n_graphs,
                                    const mg_walk_query* queries, uint32_t n_queries, const mg_batch_options* opt,
                                    mg_walk_batch** out) {
  return guarded([&] {
    MG_REQUIRE(ctx && graphs_at && queries && out, "mg_walk_batch_create: NULL argument");
    *out = nullptr;
    MG_REQUIRE(n_graphs >= 1 && n_queries >= 1, "mg_walk_batch_create: empty batch");
    const uint32_t chunk_tiles = (opt && opt->chunk_tiles) ? opt->chunk_tiles : kDefaultChunkTiles;
    const uint32_t packing = opt ? opt->packing : MG_PACK_BY_HOPS;
    const uint32_t bflags = opt ? opt->flags : 0u;
    MG_REQUIRE(packing <= MG_PACK_ONE_PER_PANEL, "mg_walk_batch_create: unknown packing");
    MG_REQUIRE((bflags & ~(MG_BATCH_SWEEP_FIRST_HOP | MG_BATCH_VERTEX_ORDER)) == 0,
               "mg_walk_batch_create: unknown flags");
    DeviceGuard g(ctx->device);
    auto b = std::make_unique<mg_walk_batch>();
    b->device = ctx->device;
    b->use_graph = !opt || opt->use_graph;
    Engine& e = b->e;
    e.W = 64;
    for (uint32_t gi = 0; gi < n_graphs; ++gi) {
      MG_REQUIRE(graphs_at[gi] != nullptr, "mg_walk_batch_create: NULL graph");
      e.graphs.push_back(plan_graph(graphs_at[gi], chunk_tiles, ctx->stream, use_relabel(bflags & MG_BATCH_VERTEX_ORDER)));
    }
    // Validate queries.
    b->queries.resize(n_queries);
    for (uint32_t q = 0; q < n_queries; ++q) {
      const mg_walk_query& Q = queries[q];
      MG_REQUIRE(Q.graph < n_graphs, "mg_walk_batch_create: query graph out of range");
      MG_REQUIRE(Q.n_sources >= 1 && Q.n_sources <= 4096, "mg_walk_batch_create: query size must be in [1, 4096]");
      MG_REQUIRE(Q.hops >= 1 && Q.hops <= 64, "mg_walk_batch_create: query hops must be in [1, 64]");
      MG_REQUIRE(Q.sources != nullptr, "mg_walk_batch_create: query sources NULL");
      const uint64_t nv = static_cast<uint64_t>(e.graphs[Q.graph].n_strips) * 64;
      std::vector<uint32_t> sorted(Q.sources, Q.sources + Q.n_sources);
      std::sort(sorted.begin(), sorted.end());
      MG_REQUIRE(sorted.back() < nv, "mg_walk_batch_create: source vertex out of range");
      MG_REQUIRE(std::adjacent_find(sorted.begin(), sorted.end()) == sorted.end(),
                 "mg_walk_batch_create: sources of a query must be distinct");
      b->queries[q].graph = Q.graph;
      b->queries[q].n_sources = Q.n_sources;
      b->queries[q].n_words = (Q.n_sources + 63) / 64;
      b->queries[q].hops = Q.hops;
    }
    // Packing: per graph, queries in descending hop order (ties: larger first), first fit
    // into 64-word panels; MG_PACK_ONE_PER_PANEL gives every query its own panel.
    std::vector<uint32_t> order(n_queries);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) {
      const auto& A = b->queries[a];
      const auto& B = b->queries[c];
      if (A.graph != B.graph) return A.graph < B.graph;
      if (A.hops != B.hops) return A.hops > B.hops;
      return A.n_words > B.n_words;
    });
    std::vector<uint32_t> panel_fill;  // words used per panel
    uint32_t first_panel_of_graph = 0;
    uint32_t cur_graph = ~0u;
    for (uint32_t q : order) {
      QueryPlace& P = b->queries[q];
      if (P.graph != cur_graph) {
        cur_graph = P.graph;
        first_panel_of_graph = static_cast<uint32_t>(e.panels.size());
      }
      uint32_t chosen = ~0u;
      if (packing == MG_PACK_BY_HOPS) {
        for (uint32_t pi = first_panel_of_graph; pi < e.panels.size(); ++pi)
          if (panel_fill[pi] + P.n_words <= 64) {
            chosen = pi;
            break;
          }
      }
      if (chosen == ~0u) {
        Panel p;
        p.graph = P.graph;
        p.hops = 0;
        p.n_strips = e.graphs[P.graph].n_strips;
        e.panels.push_back(std::move(p));
        panel_fill.push_back(0);
        chosen = static_cast<uint32_t>(e.panels.size() - 1);
      }
      P.panel = chosen;
      P.w0 = panel_fill[chosen];
      panel_fill[chosen] += P.n_words;
      Panel& pn = e.panels[chosen];
      pn.hops = std::max(pn.hops, P.hops);
    }
    for (uint32_t pi = 0; pi < e.panels.size(); ++pi) {
      e.panels[pi].words_used = panel_fill[pi];
      e.panels[pi].alive.assign(e.panels[pi].hops, 0);
    }
    for (const QueryPlace& P : b->queries)
      for (uint32_t k = 0; k < P.hops; ++k) e.panels[P.panel].alive[k] |= words_mask(P.w0, P.n_words);
    // First hop by push, per panel: when the A^T columns of its source strips hold at most
    // half of its graph's tiles. The column indexes come first: the level tables point at them.
    std::vector<PushItem> push_items;
    std::vector<uint32_t> push_panels;
    if (!(bflags & MG_BATCH_SWEEP_FIRST_HOP)) {
      std::vector<std::vector<uint32_t>> strips(e.panels.size());
      for (uint32_t q = 0; q < n_queries; ++q) {
        const QueryPlace& P = b->queries[q];
        const GraphPlan& gq = e.graphs[P.graph];
        for (uint32_t i = 0; i < P.n_sources; ++i) strips[P.panel].push_back(gq.engine_vertex(queries[q].sources[i]) / 64);
      }
      std::vector<uint8_t> indexed(e.graphs.size(), 0);
      e.push0.assign(e.panels.size(), 0);
      uint64_t pushed_sum = 0, pushed_panel_tiles = 0;
      for (uint32_t pi = 0; pi < e.panels.size(); ++pi) {
        std::vector<uint32_t>& st = strips[pi];
        std::sort(st.begin(), st.end());
        st.erase(std::unique(st.begin(), st.end()), st.end());
        const uint32_t gi = e.panels[pi].graph;
        GraphPlan& g = e.graphs[gi];
        if (!indexed[gi]) {
          build_column_index(g, ctx->stream);
          indexed[gi] = 1;
        }
        const uint64_t pushed_tiles = push_items_for(g, pi, st, nullptr);
        if (2 * pushed_tiles <= g.at->n_nz_tile) {
          e.push0[pi] = 1;
          pushed_sum += pushed_tiles;
          pushed_panel_tiles += g.at->n_nz_tile;
          push_items_for(g, pi, st, &push_items);
          push_panels.push_back(pi);
          e.push_max_strips = std::max(e.push_max_strips, g.n_strips);
        }
      }
      e.push_s_from_columns = kPushColumnS * pushed_sum <= pushed_panel_tiles;
    }
    build_engine(&e, ctx->stream);
    apply_seg(&e, default_segments(), ctx->stream);
    e.n_push_items = static_cast<uint32_t>(push_items.size());
    e.push_items.alloc(std::max<size_t>(1, push_items.size()));
    if (!push_items.empty())
      MG_HIP(hipMemcpyAsync(e.push_items.get(), push_items.data(), sizeof(PushItem) * push_items.size(),
                            hipMemcpyHostToDevice, ctx->stream));
    e.n_push_panels = static_cast<uint32_t>(push_panels.size());
    e.push_panels.alloc(std::max<size_t>(1, push_panels.size()));
    if (!push_panels.empty())
      MG_HIP(hipMemcpyAsync(e.push_panels.get(), push_panels.data(), sizeof(uint32_t) * push_panels.size(),
                            hipMemcpyHostToDevice, ctx->stream));
    // Sources: all queries' bits, uploaded once.
    std::vector<SourceItem> items;
    for (uint32_t q = 0; q < n_queries; ++q) {
      const QueryPlace& P = b->queries[q];
      const GraphPlan& gq = e.graphs[P.graph];
      for (uint32_t i = 0; i < P.n_sources; ++i)
        items.push_back({P.panel, P.w0 * 64 + i, gq.engine_vertex(queries[q].sources[i])});
    }
    e.n_sources = static_cast<uint32_t>(items.size());
    e.sources.alloc(items.size());
    MG_HIP(hipMemcpyAsync(e.sources.get(), items.data(), sizeof(SourceItem) * items.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    MG_HIP(hipStreamSynchronize(ctx->stream));  // push_items, push_panels and items are host locals
    *out = b.release();
  });
}

extern "C" int mg_walk_batch_run(mg_context* ctx, mg_walk_batch* b) {
  return guarded([&] {
    MG_REQUIRE(ctx && b, "mg_walk_batch_run: NULL argument");
    MG_REQUIRE(ctx->device == b->device, "mg_walk_batch_run: batch belongs to another device");
    DeviceGuard g(ctx->device);
    hipStream_t st = ctx->stream;
    if (!b->use_graph) {
      record_batch(st, b);
      return;
    }
    if (!b->exec) {
      MG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      try {
        record_batch(st, b);
      } catch (...) {
        hipGraph_t junk = nullptr;
        (void)hipStreamEndCapture(st, &junk);
        if (junk) (void)hipGraphDestroy(junk);
        throw;
      }
      MG_HIP(hipStreamEndCapture(st, &b->graph));
      MG_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    MG_HIP(hipGraphLaunch(b->exec, st));
  });
}

extern "C" int mg_walk_batch_pushed_panels(const mg_walk_batch* b, uint32_t* n_panels) {
  return guarded([&] {
    MG_REQUIRE(b && n_panels, "mg_walk_batch_pushed_panels: NULL argument");
    *n_panels = b->e.n_push_panels;
  });
}

extern "C" int mg_walk_batch_info(const mg_walk_batch* b, uint32_t* n_panels, uint32_t* max_hops, uint64_t* n_chunks) {
  return guarded([&] {
    MG_REQUIRE(b != nullptr, "mg_walk_batch_info: NULL batch");
    if (n_panels) *n_panels = static_cast<uint32_t>(b->e.panels.size());
    if (max_hops) *max_hops = b->e.max_hops;
    if (n_chunks) {
      uint64_t c = 0;
      for (uint32_t k = 0; k < b->e.max_hops; ++k) c += b->e.n_chunks[k];
      *n_chunks = c;
    }
  });
}

extern "C" int mg_walk_batch_query_stats(mg_context* ctx, const mg_walk_batch* b, uint32_t q, uint32_t hop,
                                         mg_hop_stats* out) {
  return guarded([&] {
    MG_REQUIRE(ctx && b && out, "mg_walk_batch_query_stats: NULL argument");
    MG_REQUIRE(q < b->queries.size(), "mg_walk_batch_query_stats: query out of range");
    const QueryPlace& P = b->queries[q];
    MG_REQUIRE(hop >= 1 && hop <= b->e.panels[P.panel].hops, "mg_walk_batch_query_stats: hop out of range");
    DeviceGuard g(ctx->device);
    unsigned long long c[kCounterStride];
    MG_HIP(hipMemcpyAsync(c, const_cast<Engine&>(b->e).ctr(P.panel, hop), sizeof(c), hipMemcpyDeviceToHost,
                          ctx->stream));
    MG_HIP(hipStreamSynchronize(ctx->stream));
    hop_stats_from(c, b->e.panels[P.panel].words_used, out);
  });
}

extern "C" int mg_walk_batch_totals(mg_context* ctx, const mg_walk_batch* b, uint64_t* tile_products,
                                    uint64_t* alg_bytes, uint64_t* edges_visited) {
  return guarded([&] {
    MG_REQUIRE(ctx && b, "mg_walk_batch_totals: NULL argument");
    DeviceGuard g(ctx->device);
    std::vector<unsigned long long> c(b->e.counters.size());
    MG_HIP(hipMemcpyAsync(c.data(), b->e.counters.get(), b->e.counters.bytes(), hipMemcpyDeviceToHost, ctx->stream));
    MG_HIP(hipStreamSynchronize(ctx->stream));
    uint64_t tp = 0, by = 0, ed = 0;
    for (uint32_t pi = 0; pi < b->e.panels.size(); ++pi)
      for (uint32_t k = 1; k <= b->e.panels[pi].hops; ++k) {
        const unsigned long long* x = c.data() + (static_cast<size_t>(pi) * (b->e.max_hops + 1) + k) * kCounterStride;
        mg_hop_stats s;
        hop_stats_from(x, b->e.panels[pi].words_used, &s);
        tp += s.tile_products;
        by += s.alg_bytes;
        ed += s.edges_visited;
      }
    if (tile_products) *tile_products = tp;
    if (alg_bytes) *alg_bytes = by;
    if (edges_visited) *edges_visited = ed;
  });
}

extern "C" int mg_walk_batch_query_digest(mg_context* ctx, const mg_walk_batch* b, uint32_t q, uint64_t* digest) {
  return guarded([&] {
    MG_REQUIRE(ctx && b && digest, "mg_walk_batch_query_digest: NULL argument");
    MG_REQUIRE(q < b->queries.size(), "mg_walk_batch_query_digest: query out of range");
    DeviceGuard g(ctx->device);
    const QueryPlace& P = b->queries[q];
    const Panel& p = b->e.panels[P.panel];
    *digest = frontier_digest(ctx, p.x[P.hops].get(), static_cast<uint64_t>(p.n_strips) * 64, b->e.W, P.w0, P.n_words,
                              b->e.graphs[P.graph].device_map());
  });
}

extern "C" int mg_walk_batch_query_download(mg_context* ctx, const mg_walk_batch* b, uint32_t q, uint64_t* out) {
  return guarded([&] {
    MG_REQUIRE(ctx && b && out, "mg_walk_batch_query_download: NULL argument");
    MG_REQUIRE(q < b->queries.size(), "mg_walk_batch_query_download: query out of range");
    DeviceGuard g(ctx->device);
    const QueryPlace& P = b->queries[q];
    const Panel& p = b->e.panels[P.panel];
    const uint64_t nv = static_cast<uint64_t>(p.n_strips) * 64;
    DevBuf<uint64_t> tmp(nv * P.n_words);
    frontier_slice_kernel<<<grid_for(nv * P.n_words, kBlock * 8, 16384), kBlock, 0, ctx->stream>>>(
        p.x[P.hops].get(), nv, b->e.W, P.w0, P.n_words, b->e.graphs[P.graph].device_map(), tmp.get());
    MG_LAUNCH_CHECK();
    MG_HIP(hipMemcpyAsync(out, tmp.get(), tmp.bytes(), hipMemcpyDeviceToHost, ctx->stream));
    MG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

extern "C" int mg_walk_batch_query_export(mg_context* ctx, const mg_walk_batch* b, uint32_t q, mg_tiled_matrix** out) {
  return guarded([&] {
    MG_REQUIRE(ctx && b && out, "mg_walk_batch_query_export: NULL argument");
    *out = nullptr;
    MG_REQUIRE(q < b->queries.size(), "mg_walk_batch_query_export: query out of range");
    DeviceGuard g(ctx->device);
    const QueryPlace& P = b->queries[q];
    const Panel& p = b->e.panels[P.panel];
    *out = export_frontier(ctx, p.x[P.hops].get(), p.s[P.hops].get(), p.n_strips, b->e.W, P.w0, P.n_words,
                           b->e.graphs[P.graph]);
  });
}

extern "C" int mg_walk_batch_destroy(mg_walk_batch* b) {
  return guarded([&] {
    if (!b) return;
    DeviceGuard g(b->device);
    delete b;
  });
}

// ----------------------------------------------------------------------------
// Scheduler control (cost models): logging, chunk records, panel info, launch order.
// ----------------------------------------------------------------------------
namespace {
void drop_graph(mg_walk_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}
}  // namespace

extern "C" int mg_walk_batch_set_hop_kernel(mg_context* ctx, mg_walk_batch* b, uint32_t segments) {
  return guarded([&] {
    MG_REQUIRE(ctx && b, "mg_walk_batch_set_hop_kernel: NULL argument");
    MG_REQUIRE(ctx->device == b->device, "mg_walk_batch_set_hop_kernel: batch belongs to another device");
    DeviceGuard g(ctx->device);
    MG_HIP(hipDeviceSynchronize());  // no run of this batch may still read its tables
    drop_graph(b);                   // kernel choice changes: capture again
    apply_seg(&b->e, segments, ctx->stream);
  });
}

extern "C" int mg_walk_batch_set_logging(mg_walk_batch* b, int enable) {
  return guarded([&] {
    MG_REQUIRE(b != nullptr, "mg_walk_batch_set_logging: NULL batch");
    DeviceGuard g(b->device);
    MG_HIP(hipDeviceSynchronize());  // no run of this batch may still be reading its tables
    Engine& e = b->e;
    const bool on = enable != 0;
    if (on && e.logs.empty()) {
      e.logs.resize(e.max_hops);
      for (uint32_t k = 0; k < e.max_hops; ++k) {
        e.logs[k].alloc(std::max<uint32_t>(1, e.n_chunks[k]));
        MG_HIP(hipMemset(e.logs[k].get(), 0, e.logs[k].bytes()));
      }
      MG_HIP(hipDeviceSynchronize());  // the fills run on the null stream; runs use the context's
    }
    if (on != e.logging) drop_graph(b);  // kernel arguments change: capture again
    e.logging = on;
  });
}

extern "C" int mg_walk_batch_level_chunks(mg_context* ctx, const mg_walk_batch* b, uint32_t level, uint64_t* n,
                                          mg_chunk_record* out) {
  return guarded([&] {
    MG_REQUIRE(ctx && b && n, "mg_walk_batch_level_chunks: NULL argument");
    const Engine& e = b->e;
    MG_REQUIRE(level < e.max_hops, "mg_walk_batch_level_chunks: level out of range");
    const auto& ch = e.host_chunks[level];
    *n = ch.size();
    if (!out || ch.empty()) return;
    DeviceGuard g(ctx->device);
    DevBuf<uint32_t> bits(ch.size());
    chunk_bits_kernel<<<static_cast<uint32_t>(ch.size()), kBlock, 0, ctx->stream>>>(
        e.chunks[level].get(), e.lv[level].get(), nullptr, bits.get());
    MG_LAUNCH_CHECK();
    std::vector<uint32_t> hb(ch.size());
    std::vector<ChunkLog> hl(ch.size(), ChunkLog{0, 0});
    MG_HIP(hipMemcpyAsync(hb.data(), bits.get(), sizeof(uint32_t) * ch.size(), hipMemcpyDeviceToHost, ctx->stream));
    if (!e.logs.empty())
      MG_HIP(hipMemcpyAsync(hl.data(), e.logs[level].get(), sizeof(ChunkLog) * ch.size(), hipMemcpyDeviceToHost,
                            ctx->stream));
    MG_HIP(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < ch.size(); ++i)
      out[i] = {ch[i].panel, ch[i].strip_begin, ch[i].strip_end, ch[i].tile_begin, ch[i].tile_end, hb[i],
                hl[i].jobs, hl[i].out_tiles};
  });
}

extern "C" int mg_walk_batch_panel_info(const mg_walk_batch* b, uint32_t p, uint32_t* graph, uint32_t* hops,
                                        uint32_t* words_used, uint64_t* alive) {
  return guarded([&] {
    MG_REQUIRE(b != nullptr, "mg_walk_batch_panel_info: NULL batch");
    MG_REQUIRE(p < b->e.panels.size(), "mg_walk_batch_panel_info: panel out of range");
    const Panel& P = b->e.panels[p];
    if (graph) *graph = P.graph;
    if (hops) *hops = P.hops;
    if (words_used) *words_used = static_cast<uint32_t>(P.words_used);
    if (alive)
      for (uint32_t k = 0; k < P.hops; ++k) alive[k] = P.alive[k];
  });
}

extern "C" int mg_walk_batch_set_level_order(mg_walk_batch* b, uint32_t level, const uint32_t* order, uint64_t n) {
  return guarded([&] {
    MG_REQUIRE(b && order, "mg_walk_batch_set_level_order: NULL argument");
    Engine& e = b->e;
    MG_REQUIRE(level < e.max_hops, "mg_walk_batch_set_level_order: level out of range");
    auto& ch = e.host_chunks[level];
    MG_REQUIRE(n == ch.size(), "mg_walk_batch_set_level_order: order must list every chunk of the level");
    std::vector<char> seen(ch.size(), 0);
    std::vector<Chunk> next(ch.size());
    for (uint64_t i = 0; i < n; ++i) {
      MG_REQUIRE(order[i] < n && !seen[order[i]], "mg_walk_batch_set_level_order: order is not a permutation");
      seen[order[i]] = 1;
      next[i] = ch[order[i]];
    }
    DeviceGuard g(b->device);
    MG_HIP(hipDeviceSynchronize());  // no run of this batch may still be reading the chunk list
    ch.swap(next);
    if (!ch.empty())
      MG_HIP(hipMemcpy(e.chunks[level].get(), ch.data(), sizeof(Chunk) * ch.size(), hipMemcpyHostToDevice));
    if (!e.logs.empty()) MG_HIP(hipMemset(e.logs[level].get(), 0, e.logs[level].bytes()));
    MG_HIP(hipDeviceSynchronize());  // null-stream fill; runs use the context's stream
  });
}
