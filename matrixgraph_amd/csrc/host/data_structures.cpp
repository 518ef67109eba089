// Host containers of the C++ drop-in surface: Bitmap, BitTiledMatrix, GridBitTiledMatrix.
#include <algorithm>
#include <cstring>
#include <iostream>

#include "host/matrixgraph_host.h"

namespace sics {
namespace matrixgraph {
namespace core {

void ThrowIfError(int status, const char* where) {
  if (status != MG_OK)
    throw std::runtime_error(std::string(where) + " failed (status " + std::to_string(status) + "): " +
                             mg_last_error());
}

namespace util {

size_t Bitmap::Count() const {
  size_t c = 0;
  for (uint64_t w : data_) c += static_cast<size_t>(__builtin_popcountll(w));
  return c;
}

size_t Bitmap::PreElementCount(size_t i) const {
  if (i > size_) return Count();
  size_t c = 0;
  for (size_t w = 0; w < (i >> 6); ++w) c += static_cast<size_t>(__builtin_popcountll(data_[w]));
  if (i & 63) c += static_cast<size_t>(__builtin_popcountll(data_[i >> 6] & ((1ull << (i & 63)) - 1)));
  return c;
}

}  // namespace util

namespace data_structures {

void BitTiledMatrix::Init(const TiledMatrixMetadata& metadata, util::Bitmap* nz_tile_bm) {
  metadata_.n_strips = metadata.n_strips;
  metadata_.n_nz_tile = metadata.n_nz_tile;
  metadata_.tile_size = metadata.tile_size;
  tile_offset_row_.assign(static_cast<size_t>(metadata_.n_strips) + 1, 0);
  tile_row_idx_.assign(metadata_.n_nz_tile, 0);
  tile_col_idx_.assign(metadata_.n_nz_tile, 0);
  data_.assign(static_cast<size_t>(kTileWords) * metadata_.n_nz_tile, 0);
  if (nz_tile_bm) {
    nz_tile_bm_ = *nz_tile_bm;
  } else {
    nz_tile_bm_.Init(static_cast<size_t>(metadata_.n_strips) * metadata_.n_strips);
  }
}

void BitTiledMatrix::Init(VertexID tile_size, VertexID n_strips, util::Bitmap* bm) {
  TiledMatrixMetadata m;
  m.tile_size = tile_size;
  m.n_strips = n_strips;
  m.n_nz_tile = bm ? static_cast<VertexID>(bm->Count()) : 0;
  Init(m, bm);
}

void BitTiledMatrix::RebuildNzTileBitmap() {
  nz_tile_bm_.Init(static_cast<size_t>(metadata_.n_strips) * metadata_.n_strips);
  for (VertexID t = 0; t < metadata_.n_nz_tile; ++t)
    nz_tile_bm_.SetBit(static_cast<size_t>(tile_row_idx_[t]) * metadata_.n_strips + tile_col_idx_[t]);
}

void BitTiledMatrix::Print() const {
  if (metadata_.n_nz_tile == 0) return;
  std::cout << "[BitTiledMatrix Print]\n * tile_offset_row:";
  for (VertexID i = 0; i <= metadata_.n_strips; ++i) std::cout << " " << tile_offset_row_[i];
  std::cout << "\n";
  for (VertexID t = 0; t < metadata_.n_nz_tile; ++t) {
    std::cout << "offset: " << t << ", x: " << tile_row_idx_[t] << " y: " << tile_col_idx_[t] << "\n";
    const uint64_t* tile = GetDataPtrByIdx(t);
    for (uint32_t r = 0; r < 64; ++r) {
      for (uint32_t c = 0; c < 64; ++c) std::cout << ((tile[r] >> c) & 1) << " ";
      std::cout << "\n";
    }
  }
  std::cout << "[Print] Done!" << std::endl;
}

mg_tiled_matrix* BitTiledMatrix::ToDevice(mg_context* ctx) const {
  mg_tiled_matrix* m = nullptr;
  const VertexID n_strips = metadata_.n_strips ? metadata_.n_strips : 1;
  std::vector<VertexID> zero_off;
  const VertexID* off = tile_offset_row_.data();
  if (metadata_.n_strips == 0) {
    zero_off.assign(2, 0);
    off = zero_off.data();
  }
  ThrowIfError(mg_tiled_matrix_upload(ctx, n_strips, metadata_.n_nz_tile,
                                      metadata_.tile_size ? metadata_.tile_size : common::kDefaultTileSize, off,
                                      tile_row_idx_.data(), tile_col_idx_.data(), data_.data(), &m),
               "BitTiledMatrix::ToDevice");
  return m;
}

void BitTiledMatrix::FromDevice(mg_context* ctx, const mg_tiled_matrix* m) {
  TiledMatrixMetadata md;
  ThrowIfError(mg_tiled_matrix_metadata(m, &md.n_strips, &md.n_nz_tile, &md.tile_size), "BitTiledMatrix::FromDevice");
  Init(md);
  ThrowIfError(mg_tiled_matrix_download(ctx, m, tile_offset_row_.data(), tile_row_idx_.data(), tile_col_idx_.data(),
                                        data_.data(), nz_tile_bm_.data()),
               "BitTiledMatrix::FromDevice");
}

GridBitTiledMatrix::GridBitTiledMatrix(const GridGraphMetadata& metadata) : metadata_(metadata) {
  tiled_matrix_vec_.resize(static_cast<size_t>(metadata_.n_chunks) * metadata_.n_chunks);
  for (auto& p : tiled_matrix_vec_) p = new BitTiledMatrix();
}

GridBitTiledMatrix::~GridBitTiledMatrix() {
  for (auto* p : tiled_matrix_vec_) delete p;
}

void GridBitTiledMatrix::Print() const {
  std::cout << "[GridBitTiledMatrix Print] " << metadata_.n_chunks << "x" << metadata_.n_chunks << std::endl;
  for (GraphID x = 0; x < metadata_.n_chunks; ++x)
    for (GraphID y = 0; y < metadata_.n_chunks; ++y) {
      auto* b = tiled_matrix_vec_[static_cast<size_t>(x) * metadata_.n_chunks + y];
      if (b->GetMetadata().n_nz_tile == 0) continue;
      std::cout << "Block (" << x << "," << y << ")" << std::endl;
      b->Print();
    }
}

}  // namespace data_structures
}  // namespace core
}  // namespace matrixgraph
}  // namespace sics
