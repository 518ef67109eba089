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

// ---------------------------------------------------------------------------
// CSR-tiled blocks (CSRTiledMatrixIO, core/io/csr_tiled_matrix_io.cu:33-172): every
// non-empty tile is a small CSR graph (ImmutableCSR) stored as one blob in data.bin at
// csr_offset[i]. Blob layout (ImmutableCSR::ParseBasePtr, core/data_structures/
// immutable_csr.cu:79-101), all u32 (EdgeIndex is u32, core/common/types.h:30):
//   globalid[nv] indegree[nv] outdegree[nv] in_offset[nv+1] out_offset[nv+1]
//   incoming[n_in] outgoing[n_out] edges_globalid_by_localid[max_vid+1]
// Vertex v of the tile is source globalid[v]; its out-edges are tile-local targets y
// (out_offset[v] .. out_offset[v+1]) whose global id is edges_globalid_by_localid[y].
// The per-tile sizes come from meta.yaml's subgraphs list (SubGraphMetadata,
// core/data_structures/metadata.h:39-51), in tile order.
// ---------------------------------------------------------------------------
namespace {

struct SubGraphMeta {
  uint32_t gid = 0, num_vertices = 0, num_incoming_edges = 0, num_outgoing_edges = 0, max_vid = 0, min_vid = 0;
};

int indent_of(const std::string& line) {
  int n = 0;
  while (n < static_cast<int>(line.size()) && (line[n] == ' ' || line[n] == '\t')) ++n;
  return n;
}

uint32_t parse_u32(const std::string& v, const std::string& what) {
  errno = 0;
  char* end = nullptr;
  const unsigned long long x = std::strtoull(v.c_str(), &end, 10);
  if (v.empty() || errno || end == v.c_str() || *end != '\0' || x > 0xFFFFFFFFull)
    io_fail(what + " is not a u32: '" + v + "'");
  return static_cast<uint32_t>(x);
}

std::string unquote(std::string v) {
  v = trim(v);
  if (v.size() >= 2 && (v.front() == '"' || v.front() == '\'') && v.back() == v.front()) v = v.substr(1, v.size() - 2);
  return v;
}

void set_sub_field(SubGraphMeta* m, const std::string& key, const std::string& val, const std::string& where) {
  const uint32_t x = parse_u32(unquote(val), where + " " + key);
  if (key == "gid") m->gid = x;
  else if (key == "num_vertices") m->num_vertices = x;
  else if (key == "num_incoming_edges") m->num_incoming_edges = x;
  else if (key == "num_outgoing_edges") m->num_outgoing_edges = x;
  else if (key == "max_vid") m->max_vid = x;
  else if (key == "min_vid") m->min_vid = x;
  else io_fail(where + ": unknown subgraph key '" + key + "'");
}

// Fields of a flow mapping "{gid: 0, num_vertices: 3, ...}".
void parse_flow_sub(const std::string& text, SubGraphMeta* m, const std::string& where) {
  std::string t = trim(text);
  if (t.size() < 2 || t.front() != '{' || t.back() != '}') io_fail(where + ": expected a subgraph mapping");
  std::stringstream ss(t.substr(1, t.size() - 2));
  std::string kv;
  while (std::getline(ss, kv, ',')) {
    if (trim(kv).empty()) continue;
    const size_t c = kv.find(':');
    if (c == std::string::npos) io_fail(where + ": expected 'key: value' in a subgraph");
    set_sub_field(m, trim(kv.substr(0, c)), kv.substr(c + 1), where);
  }
}

// meta.yaml of a CSR-tiled block: the yaml-cpp emission of
//   CSRTiledMatrixMetadata: {n_strips, n_nz_tile, tile_size, subgraphs: [SubGraphMetadata...]}
// with the list in block style (items "- key: v" followed by indented "key: v" lines, the
// dash at or deeper than the key's indentation), items in flow style, or "[]".
void read_csr_tiled_meta(const std::string& path, uint32_t* n_strips, uint32_t* n_nz_tile, uint32_t* tile_size,
                         std::vector<SubGraphMeta>* subs) {
  std::ifstream in(path);
  if (!in) io_fail("cannot open " + path);
  Section top;
  bool in_section = false, in_list = false, have_list = false;
  int list_indent = -1;
  std::string line;
  int lineno = 0;
  subs->clear();
  while (std::getline(in, line)) {
    ++lineno;
    const std::string where = path + ":" + std::to_string(lineno);
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    const std::string t = trim(line);
    if (t.empty() || t == "---" || t == "...") continue;
    const int ind = indent_of(line);
    if (ind == 0) {
      if (t != "CSRTiledMatrixMetadata:") io_fail(where + ": expected 'CSRTiledMatrixMetadata:'");
      in_section = true;
      in_list = false;
      continue;
    }
    if (!in_section) io_fail(where + ": value outside CSRTiledMatrixMetadata");
    if (in_list) {
      if (t[0] == '-') {
        if (ind < list_indent) io_fail(where + ": list item less indented than 'subgraphs'");
        subs->emplace_back();
        const std::string rest = trim(t.substr(1));
        if (rest.empty()) continue;
        if (rest[0] == '{') {
          parse_flow_sub(rest, &subs->back(), where);
        } else {
          const size_t c = rest.find(':');
          if (c == std::string::npos) io_fail(where + ": expected 'key: value' after '-'");
          set_sub_field(&subs->back(), trim(rest.substr(0, c)), rest.substr(c + 1), where);
        }
        continue;
      }
      if (ind > list_indent) {
        if (subs->empty()) io_fail(where + ": subgraph field before the first '-'");
        const size_t c = t.find(':');
        if (c == std::string::npos) io_fail(where + ": expected 'key: value'");
        set_sub_field(&subs->back(), trim(t.substr(0, c)), t.substr(c + 1), where);
        continue;
      }
      in_list = false;  // back to the section's own keys
    }
    const size_t c = t.find(':');
    if (c == std::string::npos) io_fail(where + ": expected 'key: value'");
    const std::string key = trim(t.substr(0, c));
    const std::string val = unquote(t.substr(c + 1));
    if (key == "subgraphs") {
      have_list = true;
      if (val == "[]") continue;
      if (!val.empty()) {
        if (val.front() != '[' || val.back() != ']') io_fail(where + ": subgraphs must be a list");
        // flow list of flow mappings: [{...}, {...}]
        std::string body = val.substr(1, val.size() - 2);
        size_t pos = 0;
        while ((pos = body.find('{', pos)) != std::string::npos) {
          const size_t close = body.find('}', pos);
          if (close == std::string::npos) io_fail(where + ": unterminated subgraph mapping");
          subs->emplace_back();
          parse_flow_sub(body.substr(pos, close - pos + 1), &subs->back(), where);
          pos = close + 1;
        }
        continue;
      }
      in_list = true;
      list_indent = ind;
      continue;
    }
    top[key] = val;
  }
  std::map<std::string, Section> doc{{"CSRTiledMatrixMetadata", top}};
  *n_strips = get_u32(doc, "CSRTiledMatrixMetadata", "n_strips", path);
  *n_nz_tile = get_u32(doc, "CSRTiledMatrixMetadata", "n_nz_tile", path);
  *tile_size = get_u32(doc, "CSRTiledMatrixMetadata", "tile_size", path);
  if (*n_nz_tile != 0 && !have_list) io_fail(path + ": missing CSRTiledMatrixMetadata.subgraphs");
  if (subs->size() != *n_nz_tile)
    io_fail(path + ": " + std::to_string(subs->size()) + " subgraphs listed for n_nz_tile " + std::to_string(*n_nz_tile));
}

uint64_t blob_words(const SubGraphMeta& m) {
  const uint64_t nv = m.num_vertices;
  return 3 * nv + 2 * (nv + 1) + m.num_incoming_edges + m.num_outgoing_edges + static_cast<uint64_t>(m.max_vid) + 1;
}

std::vector<uint8_t> read_all(const std::string& path, uint64_t bytes) {
  std::vector<uint8_t> buf(bytes);
  read_bin(path, buf.data(), bytes);
  return buf;
}

}  // namespace

void ReadCSRTiledBlockInfo(const std::string& dir, uint32_t* n_strips, uint32_t* n_nz_tile, uint32_t* tile_size,
                           uint64_t* n_edges) {
  std::vector<SubGraphMeta> subs;
  read_csr_tiled_meta(dir + "meta.yaml", n_strips, n_nz_tile, tile_size, &subs);
  uint64_t e = 0;
  for (const auto& m : subs) e += m.num_outgoing_edges;
  *n_edges = e;
}

void ReadCSRTiledBlock(const std::string& dir, uint32_t* src, uint32_t* dst, uint64_t capacity, uint32_t* tile_row,
                       uint32_t* tile_col) {
  uint32_t ns, nt, ts;
  std::vector<SubGraphMeta> subs;
  read_csr_tiled_meta(dir + "meta.yaml", &ns, &nt, &ts, &subs);
  uint64_t total = 0;
  for (const auto& m : subs) total += m.num_outgoing_edges;
  if (total > capacity) io_fail(dir + ": " + std::to_string(total) + " edges exceed the output capacity");
  if (nt == 0) return;
  std::vector<uint32_t> row(nt), col(nt);
  read_bin(dir + "meta_buf/row_idx.bin", row.data(), sizeof(uint32_t) * nt);
  read_bin(dir + "meta_buf/col_idx.bin", col.data(), sizeof(uint32_t) * nt);
  std::vector<uint64_t> off(static_cast<size_t>(nt) + 1);
  read_bin(dir + "meta_buf/csr_offset.bin", off.data(), sizeof(uint64_t) * off.size());
  for (uint32_t i = 0; i < nt; ++i)
    if (off[i + 1] < off[i]) io_fail(dir + "meta_buf/csr_offset.bin: offsets decrease at tile " + std::to_string(i));
  if (off[0] != 0) io_fail(dir + "meta_buf/csr_offset.bin: first offset is not 0");
  const std::vector<uint8_t> data = read_all(dir + "meta_buf/data.bin", off[nt]);
  uint64_t out = 0;
  for (uint32_t i = 0; i < nt; ++i) {
    const SubGraphMeta& m = subs[i];
    const std::string tile = dir + " tile " + std::to_string(i);
    if (off[i] % 4) io_fail(tile + ": blob offset not 4-byte aligned");
    if (blob_words(m) * 4 > off[i + 1] - off[i]) io_fail(tile + ": blob shorter than its subgraph metadata says");
    const uint32_t* b = reinterpret_cast<const uint32_t*>(data.data() + off[i]);
    const uint64_t nv = m.num_vertices;
    const uint32_t* globalid = b;
    const uint32_t* out_offset = b + 3 * nv + (nv + 1);
    const uint32_t* outgoing = out_offset + (nv + 1) + m.num_incoming_edges;
    const uint32_t* egid = outgoing + m.num_outgoing_edges;
    if (out_offset[0] != 0 || out_offset[nv] != m.num_outgoing_edges)
      io_fail(tile + ": out_offset does not span num_outgoing_edges");
    for (uint64_t v = 0; v < nv; ++v) {
      if (out_offset[v + 1] < out_offset[v]) io_fail(tile + ": out_offset decreases");
      for (uint32_t e = out_offset[v]; e < out_offset[v + 1]; ++e) {
        const uint32_t y = outgoing[e];
        if (y > m.max_vid) io_fail(tile + ": target " + std::to_string(y) + " beyond max_vid");
        if (src) src[out] = globalid[v];
        if (dst) dst[out] = egid[y];
        if (tile_row) tile_row[out] = row[i];
        if (tile_col) tile_col[out] = col[i];
        ++out;
      }
    }
  }
}

// Edgelist2CSRTiledMatrix (core/util/format_converter.cuh:417-755) followed by
// CSRTiledMatrixIO::Write. Edge (u, v) goes to tile ((u mod S) / T, (v mod S) / T) with
// tile-local source x = (u mod S) mod T and target y = (v mod S) mod T; a tile's vertices are
// its distinct x in ascending order; tiles are listed by ascending id x_tile * n_strips +
// y_tile (the reference's order with one host thread; with more its order interleaves by
// thread, which readers do not rely on); a vertex's targets are ascending (the reference keeps
// its thread-interleaved arrival order). Incoming-edge arrays are empty, as the reference's.
void WriteCSRTiledBlock(const std::string& dir, uint32_t block_scope, uint32_t tile_size, const uint32_t* src,
                        const uint32_t* dst, uint64_t n_edges) {
  if (block_scope == 0 || tile_size == 0) io_fail("CSR-tiled block: block_scope and tile_size must be positive");
  const uint64_t n_strips64 = (static_cast<uint64_t>(block_scope) + tile_size - 1) / tile_size;
  if (n_strips64 * n_strips64 > 0xFFFFFFFFull) io_fail("CSR-tiled block: n_strips^2 exceeds the u32 tile id");
  const uint32_t n_strips = static_cast<uint32_t>(n_strips64);
  struct Key {
    uint32_t tile, x, y, e;
  };
  std::vector<Key> keys(n_edges);
  for (uint64_t e = 0; e < n_edges; ++e) {
    const uint32_t ls = src[e] % block_scope, ld = dst[e] % block_scope;
    keys[e] = {(ls / tile_size) * n_strips + ld / tile_size, ls % tile_size, ld % tile_size, static_cast<uint32_t>(e)};
  }
  std::sort(keys.begin(), keys.end(), [](const Key& a, const Key& b) {
    if (a.tile != b.tile) return a.tile < b.tile;
    if (a.x != b.x) return a.x < b.x;
    if (a.y != b.y) return a.y < b.y;
    return a.e < b.e;
  });
  std::vector<SubGraphMeta> subs;
  std::vector<uint32_t> row, col, blob;
  std::vector<uint64_t> off{0};
  for (size_t i = 0; i < keys.size();) {
    size_t j = i;
    while (j < keys.size() && keys[j].tile == keys[i].tile) ++j;
    std::vector<uint32_t> gids, outdeg, targets;
    std::vector<uint32_t> egid(static_cast<size_t>(tile_size) + 1, 0);
    std::vector<int64_t> y_owner(static_cast<size_t>(tile_size) + 1, -1);
    for (size_t k = i; k < j;) {
      size_t l = k;
      while (l < j && keys[l].x == keys[k].x) ++l;
      const uint32_t u = src[keys[k].e];
      for (size_t q = k; q < l; ++q) {
        if (src[keys[q].e] != u) io_fail("CSR-tiled block: edges of one tile row come from two sources (more than one grid block)");
        const uint32_t y = keys[q].y, v = dst[keys[q].e];
        if (y_owner[y] >= 0 && egid[y] != v) io_fail("CSR-tiled block: two targets share a tile column (more than one grid block)");
        egid[y] = v;
        y_owner[y] = 1;
        targets.push_back(y);
      }
      gids.push_back(u);
      outdeg.push_back(static_cast<uint32_t>(l - k));
      k = l;
    }
    SubGraphMeta m;
    m.gid = keys[i].tile;
    m.num_vertices = static_cast<uint32_t>(gids.size());
    m.num_outgoing_edges = static_cast<uint32_t>(targets.size());
    m.max_vid = tile_size;
    subs.push_back(m);
    row.push_back(m.gid / n_strips);
    col.push_back(m.gid % n_strips);
    const size_t nv = gids.size();
    blob.insert(blob.end(), gids.begin(), gids.end());           // globalid
    blob.insert(blob.end(), nv, 0u);                              // indegree
    blob.insert(blob.end(), outdeg.begin(), outdeg.end());       // outdegree
    blob.insert(blob.end(), nv + 1, 0u);                          // in_offset
    uint32_t acc = 0;                                             // out_offset
    blob.push_back(0);
    for (uint32_t d : outdeg) blob.push_back(acc += d);
    blob.insert(blob.end(), targets.begin(), targets.end());     // outgoing (no incoming)
    blob.insert(blob.end(), egid.begin(), egid.end());           // edges_globalid_by_localid
    off.push_back(static_cast<uint64_t>(blob.size()) * 4);
    i = j;
  }
  const uint32_t n_nz_tile = static_cast<uint32_t>(subs.size());
  std::error_code ec;
  std::filesystem::create_directories(dir + "meta_buf", ec);
  if (ec) io_fail("cannot create " + dir + "meta_buf: " + ec.message());
  std::ostringstream o;
  o << "CSRTiledMatrixMetadata:\n  n_strips: " << n_strips << "\n  n_nz_tile: " << n_nz_tile
    << "\n  tile_size: " << tile_size << "\n  subgraphs:";
  if (subs.empty()) o << " []";
  o << "\n";
  for (const auto& m : subs)
    o << "    - gid: " << m.gid << "\n      num_vertices: " << m.num_vertices
      << "\n      num_incoming_edges: " << m.num_incoming_edges << "\n      num_outgoing_edges: "
      << m.num_outgoing_edges << "\n      max_vid: " << m.max_vid << "\n      min_vid: " << m.min_vid << "\n";
  if (n_nz_tile) {  // the reference writes the buffers only for non-empty blocks
    std::vector<uint32_t> per_row(static_cast<size_t>(n_strips) + 1, 0);
    for (uint32_t r : row) ++per_row[r + 1];
    for (uint32_t r = 0; r < n_strips; ++r) per_row[r + 1] += per_row[r];
    const uint64_t bits = static_cast<uint64_t>(n_strips) * n_strips;
    std::vector<uint64_t> bm(bits / 64 + 1, 0);
    for (uint32_t t = 0; t < n_nz_tile; ++t) {
      const uint64_t b = static_cast<uint64_t>(row[t]) * n_strips + col[t];
      bm[b >> 6] |= 1ull << (b & 63);
    }
    write_bin(dir + "meta_buf/row_idx.bin", row.data(), sizeof(uint32_t) * n_nz_tile);
    write_bin(dir + "meta_buf/col_idx.bin", col.data(), sizeof(uint32_t) * n_nz_tile);
    write_bin(dir + "meta_buf/tile_offset_row.bin", per_row.data(), sizeof(uint32_t) * per_row.size());
    write_bin(dir + "meta_buf/nz_tile_bm.bin", bm.data(), sizeof(uint64_t) * bm.size());
    write_bin(dir + "meta_buf/csr_offset.bin", off.data(), sizeof(uint64_t) * off.size());
    write_bin(dir + "meta_buf/data.bin", blob.data(), sizeof(uint32_t) * blob.size());
  }
  write_text(dir + "meta.yaml", o.str());
}

// Grid root of a CSR-tiled grid (GridCSRTiledMatrixIO, core/io/grid_csr_tiled_matrix_io.cu;
// written by tools/graph_converter/converter/to_grid_csr_tiled_matrix.cuh:54-85): meta.yaml
// GridGraphMetadata {n_vertices, n_edges, n_chunks, max_vid} and label.bin (u32 per vertex).
void ReadCSRTiledGridMetadata(const std::string& dir, uint32_t* n_chunks, uint32_t* n_vertices, uint32_t* n_edges,
                              uint32_t* max_vid) {
  const std::string meta = dir + "meta.yaml";
  const auto doc = read_yaml(meta);
  *n_chunks = get_u32(doc, "GridGraphMetadata", "n_chunks", meta);
  *n_vertices = get_u32(doc, "GridGraphMetadata", "n_vertices", meta);
  *n_edges = get_u32(doc, "GridGraphMetadata", "n_edges", meta);
  *max_vid = get_u32(doc, "GridGraphMetadata", "max_vid", meta);
  if (*n_chunks == 0 || *n_chunks > common::kMaxNChunks) io_fail(meta + ": n_chunks out of range");
  if (!std::filesystem::exists(dir + "label.bin")) io_fail("cannot open " + dir + "label.bin");
}

void WriteCSRTiledGridMetadata(const std::string& dir, uint32_t n_chunks, uint32_t n_vertices, uint32_t n_edges,
                               uint32_t max_vid) {
  std::ostringstream o;
  o << "GridGraphMetadata:\n  n_vertices: " << n_vertices << "\n  n_edges: " << n_edges << "\n  n_chunks: " << n_chunks
    << "\n  max_vid: " << max_vid << "\n";
  std::error_code ec;
  std::filesystem::create_directories(dir, ec);
  if (ec) io_fail("cannot create " + dir + ": " + ec.message());
  write_text(dir + "meta.yaml", o.str());
  const std::vector<uint32_t> labels(n_vertices, 0);  // the converter writes all-zero labels
  write_bin(dir + "label.bin", labels.data(), sizeof(uint32_t) * labels.size());
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

extern "C" int mg_io_read_csr_tiled_block_info(const char* dir, uint32_t* n_strips, uint32_t* n_nz_tile,
                                               uint32_t* tile_size, uint64_t* n_edges) {
  return guarded([&] {
    MG_REQUIRE(dir && n_strips && n_nz_tile && tile_size && n_edges, "mg_io_read_csr_tiled_block_info: NULL argument");
    mio::ReadCSRTiledBlockInfo(dir, n_strips, n_nz_tile, tile_size, n_edges);
  });
}

extern "C" int mg_io_read_csr_tiled_block(const char* dir, uint32_t* src, uint32_t* dst, uint64_t capacity,
                                          uint32_t* tile_row, uint32_t* tile_col) {
  return guarded([&] {
    MG_REQUIRE(dir != nullptr, "mg_io_read_csr_tiled_block: NULL dir");
    mio::ReadCSRTiledBlock(dir, src, dst, capacity, tile_row, tile_col);
  });
}

extern "C" int mg_io_write_csr_tiled_block(const char* dir, uint32_t block_scope, uint32_t tile_size,
                                           const uint32_t* src, const uint32_t* dst, uint64_t n_edges) {
  return guarded([&] {
    MG_REQUIRE(dir != nullptr, "mg_io_write_csr_tiled_block: NULL dir");
    MG_REQUIRE(n_edges == 0 || (src && dst), "mg_io_write_csr_tiled_block: NULL arrays");
    MG_REQUIRE(n_edges <= 0xFFFFFFFFull, "mg_io_write_csr_tiled_block: more than 2^32-1 edges");
    mio::WriteCSRTiledBlock(dir, block_scope, tile_size, src, dst, n_edges);
  });
}

extern "C" int mg_io_read_csr_tiled_grid_metadata(const char* dir, uint32_t* n_chunks, uint32_t* n_vertices,
                                                  uint32_t* n_edges, uint32_t* max_vid) {
  return guarded([&] {
    MG_REQUIRE(dir && n_chunks && n_vertices && n_edges && max_vid, "mg_io_read_csr_tiled_grid_metadata: NULL argument");
    mio::ReadCSRTiledGridMetadata(dir, n_chunks, n_vertices, n_edges, max_vid);
  });
}

extern "C" int mg_io_write_csr_tiled_grid_metadata(const char* dir, uint32_t n_chunks, uint32_t n_vertices,
                                                   uint32_t n_edges, uint32_t max_vid) {
  return guarded([&] {
    MG_REQUIRE(dir != nullptr, "mg_io_write_csr_tiled_grid_metadata: NULL dir");
    mio::WriteCSRTiledGridMetadata(dir, n_chunks, n_vertices, n_edges, max_vid);
  });
}
