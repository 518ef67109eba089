// On-disk bit-tile format (byte-compatible with BitTiledMatrixIO / GridBitTiledMatrixIO,
// core/io/bit_tiled_matrix_io.cu:30-136, core/io/grid_bit_tiled_matrix_io.cu:22-67) and
// the host-only C-ABI IO entry points.
//
// meta.yaml is the two-level block mapping yaml-cpp emits for these files:
//   BitTiledMatrixMetadata:\n  n_strips: N\n  n_nz_tile: N\n  tile_size: N\n
//   GridGraphMetadata:\n  n_vertices: N\n  n_edges: N\n  n_chunks: N\n
// The reader accepts that form (any indentation, comments, blank lines, "---", quoted
// scalars) and nothing else: yaml-cpp is not available here.
#include <algorithm>
#include <cerrno>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>

#include "host/matrixgraph_host.h"
#include "mg_internal.hpp"

namespace sics {
namespace matrixgraph {
namespace core {
namespace io {
namespace {

using mgamd::Failure;
using Section = std::map<std::string, std::string>;

[[noreturn]] void io_fail(const std::string& msg) { throw Failure(MG_ERR_IO, msg); }

std::string trim(const std::string& s) {
  const size_t b = s.find_first_not_of(" \t\r");
  if (b == std::string::npos) return "";
  const size_t e = s.find_last_not_of(" \t\r");
  return s.substr(b, e - b + 1);
}

std::map<std::string, Section> read_yaml(const std::string& path) {
  std::ifstream in(path);
  if (!in) io_fail("cannot open " + path);
  std::map<std::string, Section> doc;
  std::string line, current;
  int lineno = 0;
  while (std::getline(in, line)) {
    ++lineno;
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    if (trim(line).empty() || trim(line) == "---" || trim(line) == "...") continue;
    const bool indented = line[0] == ' ' || line[0] == '\t';
    const std::string t = trim(line);
    const size_t colon = t.find(':');
    if (colon == std::string::npos) io_fail(path + ":" + std::to_string(lineno) + ": expected 'key: value'");
    const std::string key = trim(t.substr(0, colon));
    std::string val = trim(t.substr(colon + 1));
    if (val.size() >= 2 && (val.front() == '"' || val.front() == '\'') && val.back() == val.front())
      val = val.substr(1, val.size() - 2);
    if (!indented) {
      if (!val.empty()) io_fail(path + ":" + std::to_string(lineno) + ": top-level key must open a mapping");
      current = key;
      doc[current];
    } else {
      if (current.empty()) io_fail(path + ":" + std::to_string(lineno) + ": value outside a mapping");
      doc[current][key] = val;
    }
  }
  return doc;
}

uint32_t get_u32(const std::map<std::string, Section>& doc, const std::string& sec, const std::string& key,
                 const std::string& path) {
  auto s = doc.find(sec);
  if (s == doc.end()) io_fail(path + ": missing " + sec);
  auto k = s->second.find(key);
  if (k == s->second.end()) io_fail(path + ": missing " + sec + "." + key);
  errno = 0;
  char* end = nullptr;
  const unsigned long long v = std::strtoull(k->second.c_str(), &end, 10);
  if (errno || end == k->second.c_str() || *end != '\0' || v > 0xFFFFFFFFull)
    io_fail(path + ": " + sec + "." + key + " is not a u32: '" + k->second + "'");
  return static_cast<uint32_t>(v);
}

void write_text(const std::string& path, const std::string& text) {
  std::ofstream out(path, std::ios::binary | std::ios::trunc);
  if (!out) io_fail("cannot write " + path);
  out << text;
  if (!out) io_fail("write failed: " + path);
}

void write_bin(const std::string& path, const void* p, size_t bytes) {
  std::ofstream out(path, std::ios::binary | std::ios::trunc);
  if (!out) io_fail("cannot write " + path);
  if (bytes) out.write(static_cast<const char*>(p), static_cast<std::streamsize>(bytes));
  if (!out) io_fail("write failed: " + path);
}

void read_bin(const std::string& path, void* p, size_t bytes) {
  std::ifstream in(path, std::ios::binary);
  if (!in) io_fail("cannot open " + path);
  in.seekg(0, std::ios::end);
  const auto have = static_cast<size_t>(in.tellg());
  if (have < bytes) io_fail(path + " is truncated: " + std::to_string(have) + " bytes, need " + std::to_string(bytes));
  in.seekg(0, std::ios::beg);
  if (bytes && p) in.read(static_cast<char*>(p), static_cast<std::streamsize>(bytes));
  if (!in) io_fail("read failed: " + path);
}

std::string block_meta(uint32_t n_strips, uint32_t n_nz_tile, uint32_t tile_size) {
  std::ostringstream o;
  o << "BitTiledMatrixMetadata:\n  n_strips: " << n_strips << "\n  n_nz_tile: " << n_nz_tile
    << "\n  tile_size: " << tile_size << "\n";
  return o.str();
}

void make_dirs(const std::string& dir) {
  std::error_code ec;
  std::filesystem::create_directories(dir + "meta_buf", ec);
  if (ec) io_fail("cannot create " + dir + "meta_buf: " + ec.message());
}

}  // namespace

// Block writer shared by the class and the C-ABI.
void WriteBlock(const std::string& dir, uint32_t n_strips, uint32_t n_nz_tile, uint32_t tile_size, const uint32_t* off,
                const uint32_t* row, const uint32_t* col, const uint64_t* data) {
  make_dirs(dir);
  const uint32_t words = std::max(1u, tile_size * tile_size / 64);
  write_bin(dir + "meta_buf/row_idx.bin", row, sizeof(uint32_t) * n_nz_tile);
  write_bin(dir + "meta_buf/col_idx.bin", col, sizeof(uint32_t) * n_nz_tile);
  std::vector<uint32_t> zero;
  if (!off) {
    zero.assign(static_cast<size_t>(n_strips) + 1, 0);
    off = zero.data();
  }
  write_bin(dir + "meta_buf/tile_offset_row.bin", off, sizeof(uint32_t) * (static_cast<size_t>(n_strips) + 1));
  const uint64_t bits = static_cast<uint64_t>(n_strips) * n_strips;
  std::vector<uint64_t> bm(bits / 64 + 1, 0);
  for (uint32_t t = 0; t < n_nz_tile; ++t) {
    const uint64_t b = static_cast<uint64_t>(row[t]) * n_strips + col[t];
    bm[b >> 6] |= 1ull << (b & 63);
  }
  write_bin(dir + "meta_buf/nz_tile_bm.bin", bm.data(), sizeof(uint64_t) * bm.size());
  write_bin(dir + "meta_buf/data.bin", data, sizeof(uint64_t) * words * static_cast<size_t>(n_nz_tile));
  write_text(dir + "meta.yaml", block_meta(n_strips, n_nz_tile, tile_size));
}

void ReadBlockMetadata(const std::string& dir, uint32_t* n_strips, uint32_t* n_nz_tile, uint32_t* tile_size) {
  const auto doc = read_yaml(dir + "meta.yaml");
  *n_strips = get_u32(doc, "BitTiledMatrixMetadata", "n_strips", dir + "meta.yaml");
  *n_nz_tile = get_u32(doc, "BitTiledMatrixMetadata", "n_nz_tile", dir + "meta.yaml");
  *tile_size = get_u32(doc, "BitTiledMatrixMetadata", "tile_size", dir + "meta.yaml");
}

void ReadBlock(const std::string& dir, uint32_t n_strips, uint32_t n_nz_tile, uint32_t tile_size, uint32_t* off,
               uint32_t* row, uint32_t* col, uint64_t* data, uint64_t* bm) {
  const uint32_t words = std::max(1u, tile_size * tile_size / 64);
  if (n_nz_tile == 0) {  // the reference stops after meta.yaml for empty blocks
    if (off) std::memset(off, 0, sizeof(uint32_t) * (static_cast<size_t>(n_strips) + 1));
    if (bm) std::memset(bm, 0, sizeof(uint64_t) * (static_cast<uint64_t>(n_strips) * n_strips / 64 + 1));
    return;
  }
  if (row) read_bin(dir + "meta_buf/row_idx.bin", row, sizeof(uint32_t) * n_nz_tile);
  if (col) read_bin(dir + "meta_buf/col_idx.bin", col, sizeof(uint32_t) * n_nz_tile);
  if (off) read_bin(dir + "meta_buf/tile_offset_row.bin", off, sizeof(uint32_t) * (static_cast<size_t>(n_strips) + 1));
  if (bm)
    read_bin(dir + "meta_buf/nz_tile_bm.bin", bm, sizeof(uint64_t) * (static_cast<uint64_t>(n_strips) * n_strips / 64 + 1));
  if (data) read_bin(dir + "meta_buf/data.bin", data, sizeof(uint64_t) * words * static_cast<size_t>(n_nz_tile));
}

void BitTiledMatrixIO::Write(const std::string& output_path, const data_structures::BitTiledMatrix& m) {
  const auto md = m.GetMetadata();
  std::cout << "[Write BitTiledMatrix] root_dir: " << output_path << " n_nz_tile: " << md.n_nz_tile << std::endl;
  WriteBlock(output_path, md.n_strips, md.n_nz_tile, md.tile_size ? md.tile_size : common::kDefaultTileSize,
             m.GetTileOffsetRowPtr(), m.GetTileRowIdxPtr(), m.GetTileColIdxPtr(), m.GetDataPtr());
}

void BitTiledMatrixIO::Read(const std::string& input_path, data_structures::BitTiledMatrix* m) {
  data_structures::TiledMatrixMetadata md;
  ReadBlockMetadata(input_path, &md.n_strips, &md.n_nz_tile, &md.tile_size);
  m->Init(md);
  ReadBlock(input_path, md.n_strips, md.n_nz_tile, md.tile_size, m->GetTileOffsetRowPtr(), m->GetTileRowIdxPtr(),
            m->GetTileColIdxPtr(), m->GetDataPtr(), m->GetNzTileBitmapPtr()->data());
}

void ReadGridMetadata(const std::string& dir, uint32_t* n_chunks, uint32_t* n_vertices, uint32_t* n_edges) {
  const std::string meta = dir + "meta.yaml";
  const auto doc = read_yaml(meta);
  *n_chunks = get_u32(doc, "GridGraphMetadata", "n_chunks", meta);
  *n_vertices = get_u32(doc, "GridGraphMetadata", "n_vertices", meta);
  *n_edges = get_u32(doc, "GridGraphMetadata", "n_edges", meta);
  if (*n_chunks == 0 || *n_chunks > common::kMaxNChunks) io_fail(meta + ": n_chunks out of range");
}

void WriteGridMetadata(const std::string& dir, uint32_t n_chunks, uint32_t n_vertices, uint32_t n_edges) {
  std::ostringstream o;
  o << "GridGraphMetadata:\n  n_vertices: " << n_vertices << "\n  n_edges: " << n_edges << "\n  n_chunks: " << n_chunks
    << "\n";
  std::error_code ec;
  std::filesystem::create_directories(dir, ec);
  if (ec) io_fail("cannot create " + dir + ": " + ec.message());
  write_text(dir + "meta.yaml", o.str());
}

void GridBitTiledMatrixIO::Read(const std::string& input_path, data_structures::GridBitTiledMatrix** grid) {
  data_structures::GridGraphMetadata md;
  ReadGridMetadata(input_path, &md.n_chunks, &md.n_vertices, &md.n_edges);
  md.max_vid = md.n_vertices ? md.n_vertices - 1 : 0;
  auto* g = new data_structures::GridBitTiledMatrix(md);
  try {
    for (uint64_t gid = 0; gid < static_cast<uint64_t>(md.n_chunks) * md.n_chunks; ++gid)
      BitTiledMatrixIO::Read(input_path + "block" + std::to_string(gid) + "/",
                             g->GetBitTileMatrixPtrByIdx(static_cast<common::GraphID>(gid)));
  } catch (...) {
    delete g;
    throw;
  }
  *grid = g;
}

void GridBitTiledMatrixIO::Write(const std::string& output_path, const data_structures::GridBitTiledMatrix& grid) {
  const auto md = grid.get_metadata();
  for (uint64_t gid = 0; gid < static_cast<uint64_t>(md.n_chunks) * md.n_chunks; ++gid)
    BitTiledMatrixIO::Write(output_path + "block" + std::to_string(gid) + "/",
                            *grid.GetBitTileMatrixPtrByIdx(static_cast<common::GraphID>(gid)));
  WriteGridMetadata(output_path, md.n_chunks, md.n_vertices, md.n_edges);
}

}  // namespace io
}  // namespace core
}  // namespace matrixgraph
}  // namespace sics

namespace mio = sics::matrixgraph::core::io;
using mgamd::guarded;

extern "C" int mg_io_read_block_metadata(const char* dir, uint32_t* n_strips, uint32_t* n_nz_tile,
                                         uint32_t* tile_size) {
  return guarded([&] {
    MG_REQUIRE(dir && n_strips && n_nz_tile && tile_size, "mg_io_read_block_metadata: NULL argument");
    mio::ReadBlockMetadata(dir, n_strips, n_nz_tile, tile_size);
  });
}

extern "C" int mg_io_read_block(const char* dir, uint32_t* tile_offset_row, uint32_t* tile_row_idx,
                                uint32_t* tile_col_idx, uint64_t* data, uint64_t* nz_tile_bm) {
  return guarded([&] {
    MG_REQUIRE(dir != nullptr, "mg_io_read_block: NULL dir");
    uint32_t ns, nt, ts;
    mio::ReadBlockMetadata(dir, &ns, &nt, &ts);
    mio::ReadBlock(dir, ns, nt, ts, tile_offset_row, tile_row_idx, tile_col_idx, data, nz_tile_bm);
  });
}

extern "C" int mg_io_write_block(const char* dir, uint32_t n_strips, uint32_t n_nz_tile, uint32_t tile_size,
                                 const uint32_t* tile_offset_row, const uint32_t* tile_row_idx,
                                 const uint32_t* tile_col_idx, const uint64_t* data) {
  return guarded([&] {
    MG_REQUIRE(dir != nullptr, "mg_io_write_block: NULL dir");
    MG_REQUIRE(n_nz_tile == 0 || (tile_row_idx && tile_col_idx && data), "mg_io_write_block: NULL arrays");
    mio::WriteBlock(dir, n_strips, n_nz_tile, tile_size, tile_offset_row, tile_row_idx, tile_col_idx, data);
  });
}

extern "C" int mg_io_read_grid_metadata(const char* dir, uint32_t* n_chunks, uint32_t* n_vertices, uint32_t* n_edges) {
  return guarded([&] {
    MG_REQUIRE(dir && n_chunks && n_vertices && n_edges, "mg_io_read_grid_metadata: NULL argument");
    mio::ReadGridMetadata(dir, n_chunks, n_vertices, n_edges);
  });
}

extern "C" int mg_io_write_grid_metadata(const char* dir, uint32_t n_chunks, uint32_t n_vertices, uint32_t n_edges) {
  return guarded([&] {
    MG_REQUIRE(dir != nullptr, "mg_io_write_grid_metadata: NULL dir");
    mio::WriteGridMetadata(dir, n_chunks, n_vertices, n_edges);
  });
}
