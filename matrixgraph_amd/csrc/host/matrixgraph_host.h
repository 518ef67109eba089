// C++ drop-in surface for the bit-tiled Boolean product path: the reference's class
// names, namespaces, argument meanings and phase order, implemented over the C-ABI of
// include/matrixgraph_amd.h (HIP instead of CUDA: hipStream_t replaces cudaStream_t).
//
// Error behaviour: the reference terminates the process through CUDA_CHECK
// (core/util/cuda_check.cuh:8-13); here every failure throws std::runtime_error carrying
// the C-ABI status and message, so a host application can recover.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "matrixgraph_amd.h"

namespace sics {
namespace matrixgraph {
namespace core {

// ---------------------------------------------------------------------------
// core/common/types.h:23-58, core/common/consts.h:14-17
// ---------------------------------------------------------------------------
namespace common {
typedef uint32_t GraphID;
typedef uint32_t VertexID;
typedef uint32_t EdgeIndex;
typedef uint16_t TileIndex;
static const uint32_t kDefaultTileSize = 64;
static const uint32_t kMaxNChunks = 256;
enum TaskType { kGEMM, kMatrixAnalysis, kPPRQuery, kSubIso, kGARMatch, kCPUSubIso, kWCC, kSSSP, kBFS, kPageRank, kGEMV };
}  // namespace common

// Throws std::runtime_error for a non-zero C-ABI status.
void ThrowIfError(int status, const char* where);

namespace util {
// Host bitmap with rank (core/util/bitmap.h:23-159): bit i = word i>>6, bit i&63;
// PreElementCount(i) = set bits strictly before i.
class Bitmap {
 public:
  Bitmap() = default;
  explicit Bitmap(size_t size) { Init(size); }
  void Init(size_t size) {
    size_ = size;
    data_.assign(size / 64 + 1, 0);
  }
  bool GetBit(size_t i) const { return i < size_ && ((data_[i >> 6] >> (i & 63)) & 1); }
  void SetBit(size_t i) {
    if (i < size_) data_[i >> 6] |= 1ull << (i & 63);
  }
  void ClearBit(size_t i) {
    if (i < size_) data_[i >> 6] &= ~(1ull << (i & 63));
  }
  size_t Count() const;
  size_t PreElementCount(size_t i) const;
  size_t size() const { return size_; }
  uint64_t* data() { return data_.data(); }
  const uint64_t* data() const { return data_.data(); }
  size_t GetBufferSize() const { return sizeof(uint64_t) * data_.size(); }

 private:
  size_t size_ = 0;
  std::vector<uint64_t> data_;
};
}  // namespace util

namespace data_structures {
using common::EdgeIndex;
using common::GraphID;
using common::VertexID;

// core/data_structures/metadata.h:12-37
struct GridGraphMetadata {
  GraphID n_chunks = 0;
  VertexID n_vertices = 0;
  EdgeIndex n_edges = 0;
  VertexID max_vid = 0;
};
struct TiledMatrixMetadata {
  VertexID n_strips = 0;
  VertexID n_nz_tile = 0;
  VertexID tile_size = 0;
  VertexID n_cols = 0;
  VertexID n_rows = 0;
};

// core/data_structures/host_buffer.cuh:12-24 (size in bytes)
template <typename T>
struct Buffer {
  T* data = nullptr;
  size_t size = 0;
  T* GetPtr() const { return data; }
  size_t GetSize() const { return size; }
  size_t GetElementSize() const { return sizeof(T); }
};

// core/data_structures/device_buffer.cuh:15-151: device memory, size in bytes,
// Init(size) zero-fills, Init(Buffer) copies host -> device.
template <typename T>
class DeviceOwnedBuffer {
 public:
  DeviceOwnedBuffer() = default;
  explicit DeviceOwnedBuffer(size_t s) { Init(s); }
  DeviceOwnedBuffer(const Buffer<T>& h, const hipStream_t& stream) { Init(h, stream); }
  DeviceOwnedBuffer(const DeviceOwnedBuffer&) = delete;
  DeviceOwnedBuffer& operator=(const DeviceOwnedBuffer&) = delete;
  DeviceOwnedBuffer(DeviceOwnedBuffer&& r) noexcept : ptr_(r.ptr_), s_(r.s_) { r.ptr_ = nullptr; r.s_ = 0; }
  DeviceOwnedBuffer& operator=(DeviceOwnedBuffer&& r) noexcept {
    if (this != &r) {
      Release();
      ptr_ = r.ptr_; s_ = r.s_; r.ptr_ = nullptr; r.s_ = 0;
    }
    return *this;
  }
  ~DeviceOwnedBuffer() { Release(); }

  void Init(size_t s) { Init(s, nullptr); }
  void Init(size_t s, const hipStream_t& stream) {
    Alloc(s);
    if (s_) Check(stream ? hipMemsetAsync(ptr_, 0, s_, stream) : hipMemset(ptr_, 0, s_), "hipMemset");
  }
  void Init(const Buffer<T>& h) { Init(h, nullptr); }
  void Init(const Buffer<T>& h, const hipStream_t& stream) {
    Alloc(h.size);
    if (s_)
      Check(stream ? hipMemcpyAsync(ptr_, h.data, s_, hipMemcpyHostToDevice, stream)
                   : hipMemcpy(ptr_, h.data, s_, hipMemcpyHostToDevice),
            "hipMemcpy H2D");
  }
  void Device2Host(const hipStream_t& stream, Buffer<T>* h) const {
    h->size = s_;
    if (s_) Check(hipMemcpyAsync(h->data, ptr_, s_, hipMemcpyDeviceToHost, stream), "hipMemcpy D2H");
  }
  void Device2Host(Buffer<T>* h) const {
    h->size = s_;
    if (s_) Check(hipMemcpy(h->data, ptr_, s_, hipMemcpyDeviceToHost), "hipMemcpy D2H");
  }
  T* GetPtr() const { return ptr_; }
  size_t GetSize() const { return s_; }
  size_t GetElementSize() const { return sizeof(T); }

 private:
  static void Check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
  }
  void Alloc(size_t s) {
    Release();
    s_ = s;
    if (s_) Check(hipMalloc(reinterpret_cast<void**>(&ptr_), s_), "hipMalloc");
  }
  void Release() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    s_ = 0;
  }
  T* ptr_ = nullptr;
  size_t s_ = 0;
};

// core/data_structures/unified_buffer.cuh:15-101 is managed memory in the reference;
// here the same interface is plain device memory (no page migration on the hot path).
template <typename T>
using UnifiedOwnedBuffer = DeviceOwnedBuffer<T>;

// core/data_structures/bit_tiled_matrix.cuh:17-72: one block of 64x64 bit tiles on the
// host (tile data, tile-CSR index and the n_strips^2 occupancy bitmap), zero-initialised.
class BitTiledMatrix {
 public:
  BitTiledMatrix() = default;
  void Init(const TiledMatrixMetadata& metadata, util::Bitmap* nz_tile_bm = nullptr);
  void Init(VertexID tile_size, VertexID n_strips, util::Bitmap* bm);
  bool IsNzTile(size_t x, size_t y) const { return nz_tile_bm_.GetBit(x * metadata_.n_strips + y); }
  void Print() const;
  TiledMatrixMetadata GetMetadata() const { return metadata_; }
  VertexID* GetTileOffsetRowPtr() const { return const_cast<VertexID*>(tile_offset_row_.data()); }
  VertexID* GetTileRowIdxPtr() const { return const_cast<VertexID*>(tile_row_idx_.data()); }
  VertexID* GetTileColIdxPtr() const { return const_cast<VertexID*>(tile_col_idx_.data()); }
  util::Bitmap* GetNzTileBitmapPtr() const { return const_cast<util::Bitmap*>(&nz_tile_bm_); }
  uint64_t* GetDataPtrByIdx(uint32_t idx) const { return GetDataPtr() + static_cast<size_t>(kTileWords) * idx; }
  uint64_t* GetDataPtr() const { return const_cast<uint64_t*>(data_.data()); }
  uint64_t GetDataBufferSize() const { return sizeof(uint64_t) * data_.size(); }
  // Rebuilds the occupancy bitmap from the tile index.
  void RebuildNzTileBitmap();

  // Device round trip through the C-ABI (device-resident copy owned by the caller).
  mg_tiled_matrix* ToDevice(mg_context* ctx) const;
  void FromDevice(mg_context* ctx, const mg_tiled_matrix* m);

  static constexpr uint32_t kTileWords = 64;  // max(1, 64*64/64)

 private:
  TiledMatrixMetadata metadata_;
  util::Bitmap nz_tile_bm_;
  std::vector<uint64_t> data_;
  std::vector<VertexID> tile_offset_row_;
  std::vector<VertexID> tile_row_idx_;
  std::vector<VertexID> tile_col_idx_;
};

// core/data_structures/grid_bit_tiled_matrix.cuh:15-39: n_chunks^2 blocks, gid = x*n_chunks + y.
class GridBitTiledMatrix {
 public:
  explicit GridBitTiledMatrix(const GridGraphMetadata& metadata);
  ~GridBitTiledMatrix();
  GridBitTiledMatrix(const GridBitTiledMatrix&) = delete;
  GridBitTiledMatrix& operator=(const GridBitTiledMatrix&) = delete;
  void SetMetadata(const GridGraphMetadata& metadata) { metadata_ = metadata; }
  GridGraphMetadata get_metadata() const { return metadata_; }
  BitTiledMatrix* GetBitTileMatrixPtrByIdx(GraphID idx) const { return tiled_matrix_vec_.at(idx); }
  void Print() const;

 private:
  GridGraphMetadata metadata_;
  std::vector<BitTiledMatrix*> tiled_matrix_vec_;
};
}  // namespace data_structures

namespace io {
// core/io/bit_tiled_matrix_io.cu:30-136 (byte-compatible; restricted YAML for meta.yaml).
class BitTiledMatrixIO {
 public:
  static void Write(const std::string& output_path, const data_structures::BitTiledMatrix& bit_tiled_matrix);
  static void Read(const std::string& input_path, data_structures::BitTiledMatrix* bit_tiled_matrix);
};
// core/io/grid_bit_tiled_matrix_io.cu:22-67
class GridBitTiledMatrixIO {
 public:
  static void Read(const std::string& input_path, data_structures::GridBitTiledMatrix** grid_tiled_matrix);
  static void Write(const std::string& output_path, const data_structures::GridBitTiledMatrix& grid_tiled_matrix);
};
}  // namespace io

namespace task {
namespace kernel {
// core/task/gpu_task/kernel/matrix_operations.cuh:14-124 (the bit-tiled entry points;
// the CSR-tile FillCSRTiles/Walk belong to another path and are not provided).
class MatrixOperationsKernelWrapper {
 public:
  MatrixOperationsKernelWrapper(const MatrixOperationsKernelWrapper&) = delete;
  void operator=(const MatrixOperationsKernelWrapper&) = delete;
  static MatrixOperationsKernelWrapper* GetInstance();

  static void MatrixBitAnd(const hipStream_t& stream, const data_structures::DeviceOwnedBuffer<uint64_t>& matrix_a_buf,
                           const data_structures::DeviceOwnedBuffer<uint64_t>& matrix_b_buf,
                           data_structures::DeviceOwnedBuffer<uint64_t>* matrix_c_buf, uint32_t m, uint32_t k,
                           uint32_t n);
  static void MatrixBitCount(const hipStream_t& stream, const data_structures::DeviceOwnedBuffer<uint64_t>& matrix_buf,
                             data_structures::DeviceOwnedBuffer<uint64_t>* count_buf, uint64_t size);
  // tile_offset_row receives per-row COUNTS (as in the reference); its size fixes n_strips
  // (n_strips + 1 entries). tile_size must be 64.
  static void InitBitTiledMatrixMetadataByLayoutMatrix(
      const hipStream_t& stream, const data_structures::DeviceOwnedBuffer<uint64_t>& layout_matrix,
      data_structures::DeviceOwnedBuffer<uint32_t>* tile_offset_row,
      data_structures::DeviceOwnedBuffer<uint32_t>* tile_row_idx,
      data_structures::DeviceOwnedBuffer<uint32_t>* tile_col_idx, uint32_t tile_size);
  static void FillTiles(const hipStream_t& stream, size_t tile_size, size_t n_strips, size_t n_nz_tile_a,
                        size_t n_nz_tile_b, size_t n_nz_tile_c,
                        const data_structures::UnifiedOwnedBuffer<uint64_t>& layout_matrix_c,
                        const data_structures::UnifiedOwnedBuffer<uint32_t>& tile_offset_row_a,
                        const data_structures::UnifiedOwnedBuffer<uint32_t>& tile_offset_row_b,
                        const data_structures::UnifiedOwnedBuffer<uint32_t>& tile_offset_row_c,
                        const data_structures::UnifiedOwnedBuffer<uint32_t>& tile_row_idx_a,
                        const data_structures::UnifiedOwnedBuffer<uint32_t>& tile_row_idx_b,
                        const data_structures::UnifiedOwnedBuffer<uint32_t>& tile_row_idx_c,
                        const data_structures::UnifiedOwnedBuffer<uint32_t>& tile_col_idx_a,
                        const data_structures::UnifiedOwnedBuffer<uint32_t>& tile_col_idx_b,
                        const data_structures::UnifiedOwnedBuffer<uint32_t>& tile_col_idx_c,
                        const data_structures::UnifiedOwnedBuffer<uint64_t>& data_a,
                        const data_structures::UnifiedOwnedBuffer<uint64_t>& data_b,
                        data_structures::UnifiedOwnedBuffer<uint64_t>* data_c);
  static void Count(const hipStream_t& stream, size_t tile_size, size_t n_strips, size_t n_nz_tile,
                    const data_structures::UnifiedOwnedBuffer<uint64_t>& data,
                    data_structures::UnifiedOwnedBuffer<uint64_t>* count);

 private:
  MatrixOperationsKernelWrapper() = default;
};
}  // namespace kernel

// core/task/gpu_task/task_base.cuh:16-35
class TaskBase {
 public:
  virtual ~TaskBase() = default;
  virtual void Run() = 0;
};

// core/task/gpu_task/ppr_query.cuh:15-60: C = A * B with B read from the transposed grid;
// Run() = symbolic (InitResultMatrix) -> numeric (FillTiles) -> Count, printing the
// same phase timers and "[Count] done! total_count:N". When output_path is non-empty the
// result grid is written there (the reference keeps it in memory only).
class PPRQuery : public TaskBase {
 public:
  PPRQuery(const std::string& input_path, const std::string& input_path_transposed, const std::string& output_path,
           size_t count);
  ~PPRQuery() override;
  void Run() override;
  uint64_t GetTotalCount() const { return total_count_; }
  uint64_t GetTileProducts() const { return tile_products_; }
  const data_structures::GridBitTiledMatrix* GetResult() const { return C_; }

 private:
  void LoadData();
  void InitResultMatrix();
  void FillTiles();
  void Count();
  void WriteResult();

  data_structures::GridBitTiledMatrix* A_ = nullptr;
  data_structures::GridBitTiledMatrix* B_ = nullptr;
  data_structures::GridBitTiledMatrix* C_ = nullptr;
  const std::string input_path_;
  const std::string input_path_transposed_;
  const std::string output_path_;
  [[maybe_unused]] const size_t count_;  // kept for signature parity; unused, as in the reference
  mg_context* ctx_ = nullptr;
  std::vector<mg_tiled_matrix*> dev_a_, dev_b_, dev_c_;
  uint64_t total_count_ = 0;
  uint64_t tile_products_ = 0;
};
}  // namespace task

// core/matrixgraph.cuh / core/matrixgraph.cu:18-89: task dispatch (single device here).
enum SchedulerType { kCHBL, kEvenSplit, kRoundRobin };
class MatrixGraph {
 public:
  explicit MatrixGraph(SchedulerType scheduler_type = kCHBL);
  void Run(common::TaskType task_type, task::TaskBase* task);

 private:
  [[maybe_unused]] SchedulerType scheduler_type_;  // one device: no block placement to schedule
  int n_device_ = 0;
};

}  // namespace core
}  // namespace matrixgraph
}  // namespace sics
