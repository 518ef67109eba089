// MatrixOperationsKernelWrapper, PPRQuery and MatrixGraph of the C++ drop-in surface,
// built on the C-ABI.
#include <algorithm>
#include <chrono>
#include <iostream>

#include "host/matrixgraph_host.h"

namespace sics {
namespace matrixgraph {
namespace core {
namespace task {
namespace kernel {

using data_structures::DeviceOwnedBuffer;
using data_structures::UnifiedOwnedBuffer;

MatrixOperationsKernelWrapper* MatrixOperationsKernelWrapper::GetInstance() {
  static MatrixOperationsKernelWrapper instance;
  return &instance;
}

void MatrixOperationsKernelWrapper::MatrixBitAnd(const hipStream_t& stream, const DeviceOwnedBuffer<uint64_t>& a,
                                                 const DeviceOwnedBuffer<uint64_t>& b, DeviceOwnedBuffer<uint64_t>* c,
                                                 uint32_t m, uint32_t k, uint32_t n) {
  ThrowIfError(mg_matrix_bit_and(stream, a.GetPtr(), b.GetPtr(), c->GetPtr(), m, k, n), "MatrixBitAnd");
}

void MatrixOperationsKernelWrapper::MatrixBitCount(const hipStream_t& stream, const DeviceOwnedBuffer<uint64_t>& bm,
                                                   DeviceOwnedBuffer<uint64_t>* count, uint64_t size) {
  ThrowIfError(mg_matrix_bit_count(stream, bm.GetPtr(), count->GetPtr(), size), "MatrixBitCount");
}

void MatrixOperationsKernelWrapper::InitBitTiledMatrixMetadataByLayoutMatrix(
    const hipStream_t& stream, const DeviceOwnedBuffer<uint64_t>& layout, DeviceOwnedBuffer<uint32_t>* tile_offset_row,
    DeviceOwnedBuffer<uint32_t>* tile_row_idx, DeviceOwnedBuffer<uint32_t>* tile_col_idx, uint32_t tile_size) {
  if (tile_size != common::kDefaultTileSize)
    ThrowIfError(MG_ERR_UNSUPPORTED, "InitBitTiledMatrixMetadataByLayoutMatrix (tile_size != 64)");
  const size_t entries = tile_offset_row->GetSize() / sizeof(uint32_t);
  if (entries < 2) ThrowIfError(MG_ERR_INVALID, "InitBitTiledMatrixMetadataByLayoutMatrix (empty tile_offset_row)");
  ThrowIfError(mg_init_bit_tiled_matrix_metadata_by_layout_matrix(stream, layout.GetPtr(),
                                                                  static_cast<uint32_t>(entries - 1),
                                                                  tile_offset_row->GetPtr(), tile_row_idx->GetPtr(),
                                                                  tile_col_idx->GetPtr()),
               "InitBitTiledMatrixMetadataByLayoutMatrix");
}

void MatrixOperationsKernelWrapper::FillTiles(
    const hipStream_t& stream, size_t tile_size, size_t n_strips, size_t n_nz_tile_a, size_t n_nz_tile_b,
    size_t n_nz_tile_c, const UnifiedOwnedBuffer<uint64_t>& layout_matrix_c,
    const UnifiedOwnedBuffer<uint32_t>& off_a, const UnifiedOwnedBuffer<uint32_t>& off_b,
    const UnifiedOwnedBuffer<uint32_t>& off_c, const UnifiedOwnedBuffer<uint32_t>& row_a,
    const UnifiedOwnedBuffer<uint32_t>& row_b, const UnifiedOwnedBuffer<uint32_t>& row_c,
    const UnifiedOwnedBuffer<uint32_t>& col_a, const UnifiedOwnedBuffer<uint32_t>& col_b,
    const UnifiedOwnedBuffer<uint32_t>& col_c, const UnifiedOwnedBuffer<uint64_t>& data_a,
    const UnifiedOwnedBuffer<uint64_t>& data_b, UnifiedOwnedBuffer<uint64_t>* data_c) {
  ThrowIfError(mg_fill_tiles(stream, tile_size, n_strips, n_nz_tile_a, n_nz_tile_b, n_nz_tile_c,
                             layout_matrix_c.GetPtr(), off_a.GetPtr(), off_b.GetPtr(), off_c.GetPtr(), row_a.GetPtr(),
                             row_b.GetPtr(), row_c.GetPtr(), col_a.GetPtr(), col_b.GetPtr(), col_c.GetPtr(),
                             data_a.GetPtr(), data_b.GetPtr(), data_c->GetPtr()),
               "FillTiles");
}

void MatrixOperationsKernelWrapper::Count(const hipStream_t& stream, size_t tile_size, size_t n_strips,
                                          size_t n_nz_tile, const UnifiedOwnedBuffer<uint64_t>& data,
                                          UnifiedOwnedBuffer<uint64_t>* count) {
  ThrowIfError(mg_count(stream, tile_size, n_strips, n_nz_tile, data.GetPtr(), count->GetPtr()), "Count");
}

}  // namespace kernel

namespace {
double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Row-major sorted union of tile index sets.
void union_index(const std::vector<std::pair<uint32_t, uint32_t>>& add, std::vector<std::pair<uint32_t, uint32_t>>* acc) {
  std::vector<std::pair<uint32_t, uint32_t>> out;
  out.reserve(acc->size() + add.size());
  std::set_union(acc->begin(), acc->end(), add.begin(), add.end(), std::back_inserter(out));
  acc->swap(out);
}
}  // namespace

using data_structures::BitTiledMatrix;
using data_structures::GridBitTiledMatrix;

PPRQuery::PPRQuery(const std::string& input_path, const std::string& input_path_transposed,
                   const std::string& output_path, size_t count)
    : input_path_(input_path), input_path_transposed_(input_path_transposed), output_path_(output_path), count_(count) {
  LoadData();
}

PPRQuery::~PPRQuery() {
  for (auto* v : {&dev_a_, &dev_b_, &dev_c_})
    for (auto* m : *v) mg_tiled_matrix_destroy(m);
  if (ctx_) mg_context_destroy(ctx_);
  delete A_;
  delete B_;
  delete C_;
}

// ppr_query.cu:63-71
void PPRQuery::LoadData() {
  io::GridBitTiledMatrixIO::Read(input_path_, &A_);
  io::GridBitTiledMatrixIO::Read(input_path_transposed_, &B_);
  if (A_->get_metadata().n_chunks != B_->get_metadata().n_chunks)
    ThrowIfError(MG_ERR_INVALID, "PPRQuery: A and A^T grids differ in n_chunks");
  C_ = new GridBitTiledMatrix(A_->get_metadata());
}

// Symbolic phase (ppr_query.cu:73-457): for every output block (i, j), the structural
// tile set is the union over k of the structural sets of A(i,k) * B(k,j); blocks are
// uploaded once and stay device-resident for the numeric phase.
void PPRQuery::InitResultMatrix() {
  if (!ctx_) ThrowIfError(mg_context_create(-1, &ctx_), "PPRQuery: mg_context_create");
  const uint32_t N = A_->get_metadata().n_chunks;
  const size_t nb = static_cast<size_t>(N) * N;
  dev_a_.assign(nb, nullptr);
  dev_b_.assign(nb, nullptr);
  dev_c_.assign(nb, nullptr);
  for (size_t g = 0; g < nb; ++g) {
    if (A_->GetBitTileMatrixPtrByIdx(g)->GetMetadata().n_nz_tile) dev_a_[g] = A_->GetBitTileMatrixPtrByIdx(g)->ToDevice(ctx_);
    if (B_->GetBitTileMatrixPtrByIdx(g)->GetMetadata().n_nz_tile) dev_b_[g] = B_->GetBitTileMatrixPtrByIdx(g)->ToDevice(ctx_);
  }
  tile_products_ = 0;
  for (uint32_t i = 0; i < N; ++i)
    for (uint32_t j = 0; j < N; ++j) {
      std::vector<std::pair<uint32_t, uint32_t>> tiles;
      uint32_t n_strips = 0;
      int contributions = 0;
      mg_tiled_matrix* single = nullptr;
      for (uint32_t k = 0; k < N; ++k) {
        mg_tiled_matrix* a = dev_a_[static_cast<size_t>(i) * N + k];
        mg_tiled_matrix* b = dev_b_[static_cast<size_t>(j) * N + k];
        if (!a || !b) continue;
        mg_tiled_matrix* c = nullptr;
        mg_product_stats st{};
        ThrowIfError(mg_multiply_symbolic(ctx_, a, b, &c, &st), "PPRQuery: symbolic");
        tile_products_ += st.tile_products;
        ThrowIfError(mg_tiled_matrix_metadata(c, &n_strips, nullptr, nullptr), "PPRQuery: metadata");
        if (++contributions == 1) {
          single = c;  // one contributing k (always the case with n_chunks == 1): use it as is
          continue;
        }
        for (mg_tiled_matrix* m : {single, c}) {
          if (!m) continue;
          uint32_t nt = 0;
          ThrowIfError(mg_tiled_matrix_metadata(m, nullptr, &nt, nullptr), "PPRQuery: metadata");
          std::vector<uint32_t> r(nt), cc(nt);
          ThrowIfError(mg_tiled_matrix_download(ctx_, m, nullptr, r.data(), cc.data(), nullptr, nullptr),
                       "PPRQuery: download index");
          std::vector<std::pair<uint32_t, uint32_t>> add(nt);
          for (uint32_t t = 0; t < nt; ++t) add[t] = {r[t], cc[t]};
          union_index(add, &tiles);
          mg_tiled_matrix_destroy(m);
        }
        single = nullptr;
      }
      const size_t gid = static_cast<size_t>(i) * N + j;
      if (contributions == 1) {
        dev_c_[gid] = single;
      } else if (contributions > 1) {
        std::vector<uint32_t> off(n_strips + 1ull, 0), r(tiles.size()), cc(tiles.size());
        for (size_t t = 0; t < tiles.size(); ++t) {
          r[t] = tiles[t].first;
          cc[t] = tiles[t].second;
          off[r[t] + 1]++;
        }
        for (uint32_t x = 0; x < n_strips; ++x) off[x + 1] += off[x];
        ThrowIfError(mg_tiled_matrix_upload(ctx_, n_strips, static_cast<uint32_t>(tiles.size()), 64, off.data(),
                                            r.data(), cc.data(), nullptr, &dev_c_[gid]),
                     "PPRQuery: upload C structure");
      }
    }
}

// Numeric phase (ppr_query.cu:459-780): OR-accumulate A(i,k) * B(k,j) into C(i,j).
void PPRQuery::FillTiles() {
  const uint32_t N = A_->get_metadata().n_chunks;
  for (uint32_t i = 0; i < N; ++i)
    for (uint32_t j = 0; j < N; ++j) {
      mg_tiled_matrix* c = dev_c_[static_cast<size_t>(i) * N + j];
      if (!c) continue;
      for (uint32_t k = 0; k < N; ++k) {
        mg_tiled_matrix* a = dev_a_[static_cast<size_t>(i) * N + k];
        mg_tiled_matrix* b = dev_b_[static_cast<size_t>(j) * N + k];
        if (a && b) ThrowIfError(mg_multiply_numeric(ctx_, a, b, c), "PPRQuery: numeric");
      }
    }
  ThrowIfError(mg_context_synchronize(ctx_), "PPRQuery: synchronize");
}

// ppr_query.cu:874-956
void PPRQuery::Count() {
  total_count_ = 0;
  for (auto* c : dev_c_) {
    if (!c) continue;
    uint64_t n = 0;
    ThrowIfError(mg_tiled_matrix_popcount(ctx_, c, &n), "PPRQuery: popcount");
    total_count_ += n;
  }
  std::cout << "[Count] done! total_count:" << total_count_ << std::endl;
}

void PPRQuery::WriteResult() {
  const uint32_t N = A_->get_metadata().n_chunks;
  uint32_t n_strips = A_->GetBitTileMatrixPtrByIdx(0)->GetMetadata().n_strips;
  for (size_t g = 0; g < static_cast<size_t>(N) * N; ++g) {
    BitTiledMatrix* dst = C_->GetBitTileMatrixPtrByIdx(static_cast<common::GraphID>(g));
    if (dev_c_[g]) {
      dst->FromDevice(ctx_, dev_c_[g]);
    } else {
      data_structures::TiledMatrixMetadata md;
      md.n_strips = n_strips;
      md.tile_size = common::kDefaultTileSize;
      dst->Init(md);
    }
  }
  if (!output_path_.empty()) io::GridBitTiledMatrixIO::Write(output_path_, *C_);
}

// ppr_query.cu:958-989
void PPRQuery::Run() {
  auto t0 = std::chrono::steady_clock::now();
  InitResultMatrix();
  std::cout << "[PPRQuery] Run Step1 InitResultMatrix() elapsed: " << seconds_since(t0) << std::endl;
  auto t1 = std::chrono::steady_clock::now();
  FillTiles();
  std::cout << "[PPRQuery] Run Step2 FillTiles() elapsed:" << seconds_since(t1) << std::endl;
  auto t2 = std::chrono::steady_clock::now();
  Count();
  std::cout << "[PPRQuery] Run Step3 Count() elapsed:" << seconds_since(t2) << std::endl;
  WriteResult();
}

}  // namespace task

MatrixGraph::MatrixGraph(SchedulerType scheduler_type) : scheduler_type_(scheduler_type) {
  if (hipGetDeviceCount(&n_device_) != hipSuccess) n_device_ = 0;
}

void MatrixGraph::Run(common::TaskType task_type, task::TaskBase* task) {
  auto t0 = std::chrono::steady_clock::now();
  switch (task_type) {
    case common::kPPRQuery:
      task->Run();
      break;
    default:
      ThrowIfError(MG_ERR_UNSUPPORTED, "MatrixGraph::Run (only kPPRQuery is provided)");
  }
  std::cout << "[MatrixGraph] Run() elapsed: "
            << std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() << std::endl;
}

}  // namespace core
}  // namespace matrixgraph
}  // namespace sics
