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
// Per hop:
//   symbolic   S_{k+1}[y] = OR_{t in A^T.row(y)} S_k[t]   (u64 per strip: tile rows x of X
//              structurally present), with tile-product / tile counts for B_alg;
//   numeric    the gather/OR kernel below, streaming A^T tiles in scheduler chunks;
//   finalize   per-strip numeric masks for strips split across chunks.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "mg_device.hpp"
#include "mg_internal.hpp"

namespace mgamd {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;
constexpr uint32_t kChunkTiles = 1024;   // scheduler target tiles per chunk
constexpr uint32_t kChunkStrips = 32;    // at most this many strips per chunk
constexpr int kLoadsInFlight = 8;        // frontier gathers per wave flush
constexpr int kTileBatch = 8;            // A^T tile loads in flight per wave

enum Counter { kTP = 0, kLeft, kRightUsed, kOut, kOutNumeric, kPop, kEdges, kNumCounters };

// Chunk of the A^T tile stream: strips [strip_begin, strip_end) restricted to tiles
// [tile_begin, tile_end). A strip whose tiles exceed one chunk is split into pieces.
struct Chunk {
  uint32_t strip_begin, strip_end, tile_begin, tile_end;
};

// --- X_0 --------------------------------------------------------------------
__global__ void init_sources_kernel(const uint32_t* __restrict__ src, uint32_t n_sources, uint32_t W,
                                    unsigned long long* __restrict__ x0, unsigned long long* __restrict__ s0,
                                    unsigned long long* __restrict__ v0) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sources) return;
  const uint32_t v = src[s];
  atomicOr(x0 + static_cast<uint64_t>(v) * W + s / 64, 1ull << (s % 64));
  atomicOr(s0 + v / 64, 1ull << (s / 64));
  atomicOr(v0 + v / 64, 1ull << (v % 64));
}

// --- symbolic ---------------------------------------------------------------
// One wave per strip y. Also accumulates T_L (structural tiles of X_k, counted once
// per strip), tile products and T_R,used.
__global__ void __launch_bounds__(kBlock) walk_symbolic_kernel(uint32_t n_strips, const uint32_t* __restrict__ off,
                                                               const uint32_t* __restrict__ col,
                                                               const uint64_t* __restrict__ s_in,
                                                               uint64_t* __restrict__ s_out,
                                                               unsigned long long* __restrict__ counters) {
  __shared__ unsigned long long s_acc[4];
  if (threadIdx.x < 4) s_acc[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t lane = lane_id();
  const uint32_t n_waves = gridDim.x * kWaves;
  uint64_t tp = 0, used = 0, left = 0, out = 0;
  for (uint32_t y = read_first_lane(blockIdx.x * kWaves + threadIdx.x / kWave); y < n_strips; y += n_waves) {
    const uint32_t p0 = off[y], p1 = off[y + 1];
    uint64_t s = 0;
    for (uint32_t p = p0 + lane; p < p1; p += kWave) {
      const uint64_t m = s_in[col[p]];
      s |= m;
      tp += __popcll(m);
      used += (m != 0);
    }
    s = wave_or(s);
    if (lane == 0) {
      s_out[y] = s;
      out += __popcll(s);
      left += __popcll(s_in[y]);
    }
  }
  tp = wave_sum(tp);
  used = wave_sum(used);
  if (lane == 0) {
    atomicAdd(&s_acc[0], static_cast<unsigned long long>(tp));
    atomicAdd(&s_acc[1], static_cast<unsigned long long>(used));
  }
  if (lane == 0) {
    atomicAdd(&s_acc[2], static_cast<unsigned long long>(left));
    atomicAdd(&s_acc[3], static_cast<unsigned long long>(out));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(counters + kTP, s_acc[0]);
    atomicAdd(counters + kRightUsed, s_acc[1]);
    atomicAdd(counters + kLeft, s_acc[2]);
    atomicAdd(counters + kOut, s_acc[3]);
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

template <int W>
__device__ inline void flush_queue(GatherQueue& q, const uint64_t* __restrict__ x_in,
                                   unsigned long long* __restrict__ acc, uint32_t lane) {
  using S = QueueShape<W>;
  uint64_t v[S::kLoads];
  if constexpr (W == 64) {
    // One job per load: its source vertex is wave-uniform.
#pragma unroll
    for (int j = 0; j < S::kLoads; ++j) {
      v[j] = 0;
      if (static_cast<uint32_t>(j) < q.len) {
        const uint32_t u = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(q.u), j));
        v[j] = x_in[static_cast<uint64_t>(u) * W + lane];
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
      v[j] = (job < q.len) ? x_in[static_cast<uint64_t>(u) * W + w] : 0;
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
                                const uint64_t* __restrict__ x_in, unsigned long long* __restrict__ acc) {
  if (lane == q.len) {
    q.u = u;
    q.rows = rows;
  }
  if (++q.len == QueueShape<W>::kCapacity) flush_queue<W>(q, x_in, acc, lane);
}

// Extract the gather jobs of one A^T tile: for each source column t' with a set bit
// (restricted to non-empty sources vm), job (t*64+t', rows with bit t').
template <int W>
__device__ inline void tile_jobs(uint32_t t, uint64_t b, uint64_t vm, GatherQueue& q, uint32_t lane,
                                 const uint64_t* __restrict__ x_in, unsigned long long* __restrict__ acc,
                                 uint64_t& edges) {
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
    push_job<W>(q, t * 64u + static_cast<uint32_t>(tp), rows, lane, x_in, acc);
  }
}

// One workgroup per scheduler chunk; per strip a 64 x W-word LDS accumulator.
template <int W>
__global__ void __launch_bounds__(kBlock) walk_numeric_kernel(
    const Chunk* __restrict__ chunks, const uint32_t* __restrict__ off, const uint32_t* __restrict__ col,
    const uint64_t* __restrict__ data, const uint64_t* __restrict__ x_in, const uint64_t* __restrict__ v_in,
    uint64_t* __restrict__ x_out, uint64_t* __restrict__ n_out, uint64_t* __restrict__ v_out,
    unsigned long long* __restrict__ counters) {
  constexpr int kWords = 64 * W;  // accumulator words per strip
  __shared__ unsigned long long acc[kWords];
  __shared__ unsigned long long s_n, s_v, s_pop;
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / kWave;
  const Chunk ch = chunks[blockIdx.x];
  uint64_t edges = 0, pop_total = 0, numeric_tiles = 0;

  for (uint32_t y = ch.strip_begin; y < ch.strip_end; ++y) {
    const uint32_t r0 = off[y], r1 = off[y + 1];
    const uint32_t seg0 = r0 > ch.tile_begin ? r0 : ch.tile_begin;
    const uint32_t seg1 = r1 < ch.tile_end ? r1 : ch.tile_end;
    const bool split = (seg0 != r0) || (seg1 != r1);
    for (int i = threadIdx.x; i < kWords; i += kBlock) acc[i] = 0;
    if (threadIdx.x == 0) { s_n = 0; s_v = 0; s_pop = 0; }
    __syncthreads();

    GatherQueue q{0, 0, 0};
    // Wave `wave` scans tiles p = seg0 + wave + kWaves*(64*g + lane): one coalesced load of
    // 64 tile columns and one gather of their source-strip masks per 64 tiles; the active
    // tiles then go in groups of kTileBatch whose 512-B loads are in flight together.
    for (uint32_t g0 = seg0 + wave; g0 < seg1; g0 += kWaves * 64) {
      const uint32_t p = g0 + kWaves * lane;
      const bool valid = p < seg1;
      const uint32_t t_l = valid ? col[p] : 0;
      const uint64_t vm_l = valid ? v_in[t_l] : 0;
      uint64_t act = ballot(vm_l != 0);
      while (act) {
        int ks[kTileBatch];
        int cnt = 0;
#pragma unroll
        for (int j = 0; j < kTileBatch; ++j) {
          ks[j] = 0;
          if (act) {
            ks[j] = __ffsll(static_cast<unsigned long long>(act)) - 1;
            act &= act - 1;
            cnt = j + 1;
          }
        }
        // A^T tiles are streamed once per hop: non-temporal loads keep them from
        // evicting the frontier words that the gathers re-read.
        uint64_t b[kTileBatch];
#pragma unroll
        for (int j = 0; j < kTileBatch; ++j)
          b[j] = (j < cnt) ? __builtin_nontemporal_load(
                                 data + static_cast<uint64_t>(g0 + kWaves * static_cast<uint32_t>(ks[j])) * kTile + lane)
                           : 0;
#pragma unroll
        for (int j = 0; j < kTileBatch; ++j) {
          if (j < cnt) {
            const uint32_t t = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(t_l), ks[j]));
            const uint64_t vm = read_lane(vm_l, ks[j]);
            tile_jobs<W>(t, b[j], vm, q, lane, x_in, acc, edges);
          }
        }
      }
    }
    if (q.len) flush_queue<W>(q, x_in, acc, lane);
    __syncthreads();

    // Write the strip out. Unsplit strips also reduce their numeric masks here.
    uint64_t* dst = x_out + static_cast<uint64_t>(y) * kWords;
    if (!split) {
      uint64_t nmask = 0, vmask = 0, pc = 0;
      for (int i = threadIdx.x; i < kWords; i += kBlock) {
        const uint64_t val = acc[i];
        dst[i] = val;
        if (val) {
          nmask |= 1ull << (i % W);
          vmask |= 1ull << (i / W);
          pc += __popcll(val);
        }
      }
      nmask = wave_or(nmask);
      vmask = wave_or(vmask);
      pc = wave_sum(pc);
      if (lane == 0) {
        if (nmask) atomicOr(&s_n, static_cast<unsigned long long>(nmask));
        if (vmask) atomicOr(&s_v, static_cast<unsigned long long>(vmask));
        if (pc) atomicAdd(&s_pop, static_cast<unsigned long long>(pc));
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        n_out[y] = s_n;
        v_out[y] = s_v;
        pop_total += s_pop;
        numeric_tiles += __popcll(s_n);
      }
    } else {
      for (int i = threadIdx.x; i < kWords; i += kBlock) {
        const uint64_t val = acc[i];
        if (val) atomicOr(reinterpret_cast<unsigned long long*>(dst + i), static_cast<unsigned long long>(val));
      }
    }
    __syncthreads();
  }
  edges = wave_sum(edges);
  if (lane == 0 && edges) atomicAdd(counters + kEdges, static_cast<unsigned long long>(edges));
  if (threadIdx.x == 0) {
    if (pop_total) atomicAdd(counters + kPop, static_cast<unsigned long long>(pop_total));
    if (numeric_tiles) atomicAdd(counters + kOutNumeric, static_cast<unsigned long long>(numeric_tiles));
  }
}

// Zero the strips that are written by several chunks (atomic OR targets).
template <int W>
__global__ void __launch_bounds__(kBlock) zero_strips_kernel(const uint32_t* __restrict__ strips, uint64_t* __restrict__ x) {
  constexpr int kWords = 64 * W;
  uint64_t* dst = x + static_cast<uint64_t>(strips[blockIdx.x]) * kWords;
  for (int i = threadIdx.x; i < kWords; i += kBlock) dst[i] = 0;
}

// Masks and popcount of the split strips once every piece has been OR-ed in.
template <int W>
__global__ void __launch_bounds__(kBlock) finalize_strips_kernel(const uint32_t* __restrict__ strips,
                                                                 const uint64_t* __restrict__ x,
                                                                 uint64_t* __restrict__ n_out,
                                                                 uint64_t* __restrict__ v_out,
                                                                 unsigned long long* __restrict__ counters) {
  constexpr int kWords = 64 * W;
  __shared__ unsigned long long s_n, s_v, s_pop;
  if (threadIdx.x == 0) { s_n = 0; s_v = 0; s_pop = 0; }
  __syncthreads();
  const uint32_t y = strips[blockIdx.x];
  const uint64_t* src = x + static_cast<uint64_t>(y) * kWords;
  uint64_t nmask = 0, vmask = 0, pc = 0;
  for (int i = threadIdx.x; i < kWords; i += kBlock) {
    const uint64_t val = src[i];
    if (val) {
      nmask |= 1ull << (i % W);
      vmask |= 1ull << (i / W);
      pc += __popcll(val);
    }
  }
  nmask = wave_or(nmask);
  vmask = wave_or(vmask);
  pc = wave_sum(pc);
  if (lane_id() == 0) {
    if (nmask) atomicOr(&s_n, static_cast<unsigned long long>(nmask));
    if (vmask) atomicOr(&s_v, static_cast<unsigned long long>(vmask));
    if (pc) atomicAdd(&s_pop, static_cast<unsigned long long>(pc));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    n_out[y] = s_n;
    v_out[y] = s_v;
    if (s_pop) atomicAdd(counters + kPop, s_pop);
    if (s_n) atomicAdd(counters + kOutNumeric, static_cast<unsigned long long>(__popcll(s_n)));
  }
}

// Digest of a vertex-major frontier: non-zero words of the first n_words source words.
__global__ void __launch_bounds__(kBlock) frontier_digest_kernel(const uint64_t* __restrict__ x, uint64_t n_vertices,
                                                                 uint32_t W, uint32_t n_words,
                                                                 unsigned long long* __restrict__ out) {
  uint64_t acc = 0;
  const uint64_t total = n_vertices * W;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride) {
    const uint32_t w = static_cast<uint32_t>(i % W);
    const uint64_t val = x[i];
    if (val && w < n_words) acc += digest_term((i / W) * 64 + w, val);
  }
  acc = wave_sum(acc);
  if (lane_id() == 0 && acc) atomicAdd(out, static_cast<unsigned long long>(acc));
}

// Export X_k tile (x, y) in the reference orientation: X tile word c (source x*64+c)
// bit r (vertex y*64+r) = bit c of vertex-major word x of vertex y*64+r.
__global__ void __launch_bounds__(kBlock) export_tiles_kernel(const uint32_t* __restrict__ row_idx,
                                                              const uint32_t* __restrict__ col_idx, uint32_t n_tiles,
                                                              const uint64_t* __restrict__ x, uint32_t W,
                                                              uint64_t* __restrict__ out) {
  const uint32_t lane = lane_id();
  const uint32_t n_waves = gridDim.x * kWaves;
  for (uint32_t i = read_first_lane(blockIdx.x * kWaves + threadIdx.x / kWave); i < n_tiles; i += n_waves) {
    const uint32_t xr = row_idx[i], y = col_idx[i];
    const uint64_t w = x[(static_cast<uint64_t>(y) * 64 + lane) * W + xr];
    uint64_t mine = 0;
    for (int c = 0; c < 64; ++c) {
      const uint64_t m = ballot((w >> c) & 1);
      if (lane == static_cast<uint32_t>(c)) mine = m;
    }
    out[static_cast<uint64_t>(i) * kTile + lane] = mine;
  }
}

std::vector<Chunk> build_chunks(const std::vector<uint32_t>& off, uint32_t n_strips, std::vector<uint32_t>* split) {
  std::vector<Chunk> chunks;
  uint32_t cur_begin = 0, cur_tiles = 0, cur_strips = 0;
  auto close = [&](uint32_t end_strip) {
    if (end_strip > cur_begin) chunks.push_back({cur_begin, end_strip, off[cur_begin], off[end_strip]});
    cur_begin = end_strip;
    cur_tiles = 0;
    cur_strips = 0;
  };
  for (uint32_t y = 0; y < n_strips; ++y) {
    const uint32_t sz = off[y + 1] - off[y];
    if (sz > kChunkTiles) {
      close(y);
      split->push_back(y);
      for (uint32_t b = off[y]; b < off[y + 1]; b += kChunkTiles)
        chunks.push_back({y, y + 1, b, std::min(off[y + 1], b + kChunkTiles)});
      cur_begin = y + 1;
      continue;
    }
    if (cur_tiles + sz > kChunkTiles || cur_strips == kChunkStrips) close(y);
    cur_tiles += sz;
    ++cur_strips;
  }
  close(n_strips);
  // Heaviest chunks first so the tail of the launch is made of light ones.
  std::stable_sort(chunks.begin(), chunks.end(), [](const Chunk& a, const Chunk& b) {
    return (a.tile_end - a.tile_begin) > (b.tile_end - b.tile_begin);
  });
  return chunks;
}

}  // namespace
}  // namespace mgamd

using namespace mgamd;

struct mg_walk {
  int device = 0;
  const mg_tiled_matrix* at = nullptr;
  uint32_t n_strips = 0;
  uint32_t W = 64;
  uint32_t max_sources = 0;
  uint32_t max_hops = 0;
  uint32_t flags = 0;
  uint32_t n_sources = 0;  // of the last run
  uint32_t hops = 0;       // of the last run
  bool timing = false;
  std::vector<DevBuf<uint64_t>> x;    // hops+1 frontiers, vertex-major
  std::vector<DevBuf<uint64_t>> s;    // structural tile-row masks per strip
  std::vector<DevBuf<uint64_t>> nmask;  // numeric tile-row masks per strip
  std::vector<DevBuf<uint64_t>> vmask;  // non-empty vertex masks per strip
  DevBuf<Chunk> chunks;
  uint32_t n_chunks = 0;
  DevBuf<uint32_t> split_strips;
  uint32_t n_split = 0;
  DevBuf<unsigned long long> counters;  // (max_hops+1) * kNumCounters
  DevBuf<uint32_t> sources;
  PinnedBuf<uint32_t> staging;
  // Timing log: per timed run, 4 events per hop (symbolic start/end, numeric kernel
  // start/end), kept for every run since timing was enabled.
  std::vector<std::vector<hipEvent_t>> run_events;
  std::vector<hipEvent_t> event_pool;
  ~mg_walk() {
    for (auto& r : run_events)
      for (auto e : r) (void)hipEventDestroy(e);
    for (auto e : event_pool) (void)hipEventDestroy(e);
  }
  hipEvent_t take_event() {
    if (event_pool.empty()) {
      hipEvent_t e;
      MG_HIP(hipEventCreate(&e));
      return e;
    }
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  void reset_log() {
    for (auto& r : run_events)
      for (auto e : r) event_pool.push_back(e);
    run_events.clear();
  }
};

namespace {

template <int W>
void launch_numeric_w(hipStream_t st, mg_walk* w, uint32_t k, hipEvent_t ev_begin, hipEvent_t ev_end) {
  const mg_tiled_matrix* at = w->at;
  unsigned long long* ctr = w->counters.get() + (k + 1) * kNumCounters;
  if (w->n_split) {
    zero_strips_kernel<W><<<w->n_split, kBlock, 0, st>>>(w->split_strips.get(), w->x[k + 1].get());
    MG_LAUNCH_CHECK();
  }
  if (ev_begin) MG_HIP(hipEventRecord(ev_begin, st));
  walk_numeric_kernel<W><<<w->n_chunks, kBlock, 0, st>>>(
      w->chunks.get(), at->offset_row.get(), at->col_idx.get(), at->data.get(), w->x[k].get(), w->vmask[k].get(),
      w->x[k + 1].get(), w->nmask[k + 1].get(), w->vmask[k + 1].get(), ctr);
  MG_LAUNCH_CHECK();
  if (ev_end) MG_HIP(hipEventRecord(ev_end, st));
  if (w->n_split) {
    finalize_strips_kernel<W><<<w->n_split, kBlock, 0, st>>>(w->split_strips.get(), w->x[k + 1].get(),
                                                            w->nmask[k + 1].get(), w->vmask[k + 1].get(), ctr);
    MG_LAUNCH_CHECK();
  }
}

void launch_numeric(mg_context* ctx, mg_walk* w, uint32_t k, hipEvent_t ev_begin, hipEvent_t ev_end) {
  switch (w->W) {
    case 1: launch_numeric_w<1>(ctx->stream, w, k, ev_begin, ev_end); break;
    case 2: launch_numeric_w<2>(ctx->stream, w, k, ev_begin, ev_end); break;
    case 4: launch_numeric_w<4>(ctx->stream, w, k, ev_begin, ev_end); break;
    case 8: launch_numeric_w<8>(ctx->stream, w, k, ev_begin, ev_end); break;
    case 16: launch_numeric_w<16>(ctx->stream, w, k, ev_begin, ev_end); break;
    case 32: launch_numeric_w<32>(ctx->stream, w, k, ev_begin, ev_end); break;
    case 64: launch_numeric_w<64>(ctx->stream, w, k, ev_begin, ev_end); break;
    default: fail(MG_ERR_INTERNAL, "walk: unsupported words per vertex " + std::to_string(w->W));
  }
}

}  // namespace

extern "C" int mg_walk_create(mg_context* ctx, const mg_tiled_matrix* at, uint32_t max_sources, uint32_t max_hops,
                              uint32_t flags, mg_walk** out) {
  return guarded([&] {
    MG_REQUIRE(ctx && at && out, "mg_walk_create: NULL argument");
    *out = nullptr;
    MG_REQUIRE(max_sources >= 1 && max_sources <= 4096, "mg_walk_create: max_sources must be in [1, 4096]");
    MG_REQUIRE(max_hops >= 1 && max_hops <= 64, "mg_walk_create: max_hops must be in [1, 64]");
    MG_REQUIRE((flags & ~MG_WALK_PRUNE) == 0, "mg_walk_create: unknown flags");
    DeviceGuard g(ctx->device);
    auto w = std::make_unique<mg_walk>();
    w->device = ctx->device;
    w->at = at;
    w->n_strips = at->n_strips;
    // Words per vertex: the smallest power of two covering max_sources (64 sources per word).
    w->W = 1;
    while (w->W * 64 < max_sources) w->W <<= 1;
    w->max_sources = max_sources;
    w->max_hops = max_hops;
    w->flags = flags;
    const uint64_t n_vertices = static_cast<uint64_t>(at->n_strips) * 64;
    w->x.resize(max_hops + 1);
    w->s.resize(max_hops + 1);
    w->nmask.resize(max_hops + 1);
    w->vmask.resize(max_hops + 1);
    for (uint32_t k = 0; k <= max_hops; ++k) {
      w->x[k].alloc(n_vertices * w->W);
      w->s[k].alloc(at->n_strips);
      w->nmask[k].alloc(at->n_strips);
      w->vmask[k].alloc(at->n_strips);
    }
    // Scheduler: chunks over the A^T tile stream, built once per graph.
    std::vector<uint32_t> off(at->n_strips + 1ull);
    MG_HIP(hipMemcpy(off.data(), at->offset_row.get(), sizeof(uint32_t) * off.size(), hipMemcpyDeviceToHost));
    std::vector<uint32_t> split;
    std::vector<Chunk> chunks = build_chunks(off, at->n_strips, &split);
    w->n_chunks = static_cast<uint32_t>(chunks.size());
    w->chunks.alloc(chunks.size());
    MG_HIP(hipMemcpy(w->chunks.get(), chunks.data(), sizeof(Chunk) * chunks.size(), hipMemcpyHostToDevice));
    w->n_split = static_cast<uint32_t>(split.size());
    if (!split.empty()) {
      w->split_strips.alloc(split.size());
      MG_HIP(hipMemcpy(w->split_strips.get(), split.data(), sizeof(uint32_t) * split.size(), hipMemcpyHostToDevice));
    }
    w->counters.alloc(static_cast<size_t>(max_hops + 1) * kNumCounters);
    w->sources.alloc(max_sources);
    w->staging.alloc(max_sources);
    *out = w.release();
  });
}

extern "C" int mg_walk_set_timing(mg_walk* w, int enable) {
  return guarded([&] {
    MG_REQUIRE(w != nullptr, "mg_walk_set_timing: NULL walk");
    w->timing = enable != 0;
    w->reset_log();
  });
}

extern "C" int mg_walk_run(mg_context* ctx, mg_walk* w, uint32_t n_sources, const uint32_t* sources, uint32_t hops) {
  return guarded([&] {
    MG_REQUIRE(ctx && w, "mg_walk_run: NULL argument");
    MG_REQUIRE(ctx->device == w->device, "mg_walk_run: walk belongs to another device");
    MG_REQUIRE(n_sources >= 1 && n_sources <= w->max_sources, "mg_walk_run: n_sources out of range");
    MG_REQUIRE(hops >= 1 && hops <= w->max_hops, "mg_walk_run: hops out of range");
    MG_REQUIRE(sources != nullptr, "mg_walk_run: sources is NULL");
    const uint64_t n_vertices = static_cast<uint64_t>(w->n_strips) * 64;
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
    std::memcpy(w->staging.get(), sources, sizeof(uint32_t) * n_sources);
    MG_HIP(hipMemcpyAsync(w->sources.get(), w->staging.get(), sizeof(uint32_t) * n_sources, hipMemcpyHostToDevice, st));
    const uint32_t ns = w->n_strips;
    MG_HIP(hipMemsetAsync(w->x[0].get(), 0, w->x[0].bytes(), st));
    MG_HIP(hipMemsetAsync(w->s[0].get(), 0, sizeof(uint64_t) * ns, st));
    MG_HIP(hipMemsetAsync(w->vmask[0].get(), 0, sizeof(uint64_t) * ns, st));
    MG_HIP(hipMemsetAsync(w->counters.get(), 0, w->counters.bytes(), st));
    init_sources_kernel<<<grid_for(n_sources, 256), 256, 0, st>>>(
        w->sources.get(), n_sources, w->W, reinterpret_cast<unsigned long long*>(w->x[0].get()),
        reinterpret_cast<unsigned long long*>(w->s[0].get()), reinterpret_cast<unsigned long long*>(w->vmask[0].get()));
    MG_LAUNCH_CHECK();
    MG_HIP(hipMemcpyAsync(w->nmask[0].get(), w->s[0].get(), sizeof(uint64_t) * ns, hipMemcpyDeviceToDevice, st));
    const bool prune = (w->flags & MG_WALK_PRUNE) != 0;
    std::vector<hipEvent_t> ev;
    if (w->timing) {
      ev.resize(4 * static_cast<size_t>(hops));
      for (auto& e : ev) e = w->take_event();
    }
    for (uint32_t k = 0; k < hops; ++k) {
      if (w->timing) MG_HIP(hipEventRecord(ev[4 * k], st));
      const uint64_t* s_in = prune ? w->nmask[k].get() : w->s[k].get();
      walk_symbolic_kernel<<<grid_for(ns, kWaves, 16384), kBlock, 0, st>>>(
          ns, w->at->offset_row.get(), w->at->col_idx.get(), s_in, w->s[k + 1].get(),
          w->counters.get() + (k + 1) * kNumCounters);
      MG_LAUNCH_CHECK();
      if (w->timing) MG_HIP(hipEventRecord(ev[4 * k + 1], st));
      launch_numeric(ctx, w, k, w->timing ? ev[4 * k + 2] : nullptr, w->timing ? ev[4 * k + 3] : nullptr);
    }
    if (w->timing) w->run_events.push_back(std::move(ev));
  });
}

extern "C" int mg_walk_hop_stats(mg_context* ctx, mg_walk* w, uint32_t hop, mg_hop_stats* out) {
  return guarded([&] {
    MG_REQUIRE(ctx && w && out, "mg_walk_hop_stats: NULL argument");
    MG_REQUIRE(hop >= 1 && hop <= w->hops, "mg_walk_hop_stats: hop out of range");
    DeviceGuard g(ctx->device);
    unsigned long long c[kNumCounters];
    MG_HIP(hipMemcpyAsync(c, w->counters.get() + hop * kNumCounters, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
    MG_HIP(hipStreamSynchronize(ctx->stream));
    std::memset(out, 0, sizeof(*out));
    out->tile_products = c[kTP];
    out->left_tiles = c[kLeft];
    out->right_tiles_used = c[kRightUsed];
    out->out_tiles = c[kOut];
    out->out_tiles_numeric = c[kOutNumeric];
    out->popcount = c[kPop];
    out->edges_visited = c[kEdges];
    const uint64_t rows = (w->n_sources + 63) / 64;
    out->alg_bytes = 520ull * (c[kLeft] + c[kRightUsed] + c[kOut]) + 4ull * (2 * rows + 2);
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
    *words = w->W;
  });
}

extern "C" int mg_walk_frontier_download(mg_context* ctx, const mg_walk* w, uint32_t hop, uint64_t* xt) {
  return guarded([&] {
    MG_REQUIRE(ctx && w && xt, "mg_walk_frontier_download: NULL argument");
    MG_REQUIRE(hop <= w->hops, "mg_walk_frontier_download: hop out of range");
    DeviceGuard g(ctx->device);
    MG_HIP(hipMemcpyAsync(xt, w->x[hop].get(), w->x[hop].bytes(), hipMemcpyDeviceToHost, ctx->stream));
    MG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

extern "C" int mg_walk_frontier_digest(mg_context* ctx, const mg_walk* w, uint32_t hop, uint64_t* digest) {
  return guarded([&] {
    MG_REQUIRE(ctx && w && digest, "mg_walk_frontier_digest: NULL argument");
    MG_REQUIRE(hop <= w->hops, "mg_walk_frontier_digest: hop out of range");
    DeviceGuard g(ctx->device);
    unsigned long long* d = ctx->counters.get();
    MG_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), ctx->stream));
    const uint64_t nv = static_cast<uint64_t>(w->n_strips) * 64;
    frontier_digest_kernel<<<grid_for(nv * w->W, kBlock * 16, 16384), kBlock, 0, ctx->stream>>>(
        w->x[hop].get(), nv, w->W, (w->n_sources + 63) / 64, d);
    MG_LAUNCH_CHECK();
    unsigned long long h = 0;
    MG_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    MG_HIP(hipStreamSynchronize(ctx->stream));
    *digest = h;
  });
}

extern "C" int mg_walk_export(mg_context* ctx, const mg_walk* w, uint32_t hop, mg_tiled_matrix** out) {
  return guarded([&] {
    MG_REQUIRE(ctx && w && out, "mg_walk_export: NULL argument");
    *out = nullptr;
    MG_REQUIRE(hop <= w->hops, "mg_walk_export: hop out of range");
    DeviceGuard g(ctx->device);
    const uint32_t ns = w->n_strips;
    const uint32_t rows = (w->n_sources + 63) / 64;
    std::vector<uint64_t> s(ns);
    MG_HIP(hipMemcpyAsync(s.data(), w->s[hop].get(), sizeof(uint64_t) * ns, hipMemcpyDeviceToHost, ctx->stream));
    MG_HIP(hipStreamSynchronize(ctx->stream));
    // Tile index in row-major (x, y) order over the padded-square grid.
    std::vector<uint32_t> off(ns + 1ull, 0), ri, ci;
    for (uint32_t xr = 0; xr < rows; ++xr) {
      for (uint32_t y = 0; y < ns; ++y)
        if ((s[y] >> xr) & 1) {
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
      export_tiles_kernel<<<grid_for(nt, kWaves, 65536), kBlock, 0, ctx->stream>>>(
          m->row_idx.get(), m->col_idx.get(), nt, w->x[hop].get(), w->W, m->data.get());
      MG_LAUNCH_CHECK();
    }
    MG_HIP(hipStreamSynchronize(ctx->stream));
    *out = m.release();
  });
}

extern "C" int mg_walk_destroy(mg_walk* w) {
  return guarded([&] {
    if (!w) return;
    DeviceGuard g(w->device);
    delete w;
  });
}
