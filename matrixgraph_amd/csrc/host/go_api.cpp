// Foreign-function surface for the bit path in the style of the reference's Go API
// (core/go_api/matrixgraph_go_api.h:8-55, core/go_api/matrixgraph_go_api.cu:23-104): host
// pointers only, int status (0 = success, 1 = error), the error message on stderr, and the
// device taken from the environment (MATRIXGRAPH_HIP_DEVICE, else 0) on every call. Each call
// creates its context, uploads, runs on the device, downloads and releases everything.
#include <cstdio>
#include <vector>

#include "matrixgraph_amd.h"

namespace {

struct Scoped {
  mg_context* ctx = nullptr;
  mg_edge_list* el = nullptr;
  mg_tiled_matrix* at = nullptr;
  mg_walk* w = nullptr;
  void* buf[3] = {nullptr, nullptr, nullptr};
  ~Scoped() {
    if (w) mg_walk_destroy(w);
    if (at) mg_tiled_matrix_destroy(at);
    if (el) mg_edges_destroy(el);
    for (void* b : buf)
      if (b && ctx) mg_buffer_free(ctx, b);
    if (ctx) mg_context_destroy(ctx);
  }
};

int report(const char* fn, const char* what) {
  std::fprintf(stderr, "%s: %s\n", fn, what);
  return 1;
}

}  // namespace

extern "C" int matrixgraph_bit_matmult(const uint64_t* A, const uint64_t* Bt, uint64_t* C, int m, int k, int n) {
  const char* fn = "matrixgraph_bit_matmult";
  if (!A || !Bt || !C || m <= 0 || k <= 0 || n <= 0) return report(fn, "NULL matrix or non-positive size");
  const size_t wa = (static_cast<size_t>(m) * k + 63) / 64;
  const size_t wb = (static_cast<size_t>(n) * k + 63) / 64;
  const size_t wc = (static_cast<size_t>(m) * n + 63) / 64;
  Scoped s;
  void* stream = nullptr;
  if (mg_context_create(-1, &s.ctx) || mg_context_stream(s.ctx, &stream) ||
      mg_buffer_alloc(s.ctx, wa * 8, &s.buf[0]) || mg_buffer_alloc(s.ctx, wb * 8, &s.buf[1]) ||
      mg_buffer_alloc(s.ctx, wc * 8, &s.buf[2]) || mg_buffer_upload(s.ctx, s.buf[0], A, wa * 8) ||
      mg_buffer_upload(s.ctx, s.buf[1], Bt, wb * 8) ||
      mg_matrix_bit_and(stream, static_cast<const uint64_t*>(s.buf[0]), static_cast<const uint64_t*>(s.buf[1]),
                        static_cast<uint64_t*>(s.buf[2]), static_cast<uint32_t>(m), static_cast<uint32_t>(k),
                        static_cast<uint32_t>(n)) ||
      mg_buffer_download(s.ctx, C, s.buf[2], wc * 8))
    return report(fn, mg_last_error());
  return 0;
}

extern "C" int matrixgraph_bit_walk(const uint32_t* src, const uint32_t* dst, int n_edges, int n_vertices,
                                    const uint32_t* sources, int n_sources, int hops, uint64_t* x_out,
                                    uint64_t* popcount) {
  const char* fn = "matrixgraph_bit_walk";
  if (n_edges < 0 || n_vertices <= 0 || n_sources <= 0 || n_sources > 4096 || hops <= 0)
    return report(fn, "size out of range (n_vertices > 0, 0 < n_sources <= 4096, hops > 0)");
  if ((n_edges > 0 && (!src || !dst)) || !sources || !x_out) return report(fn, "NULL array");
  Scoped s;
  uint32_t words = 0;
  if (mg_context_create(-1, &s.ctx) ||
      mg_edges_upload(s.ctx, static_cast<uint32_t>(n_vertices), static_cast<uint64_t>(n_edges), src, dst, &s.el) ||
      mg_edges_to_tiled_matrix(s.ctx, s.el, 1, &s.at) ||
      mg_walk_create(s.ctx, s.at, static_cast<uint32_t>(n_sources), static_cast<uint32_t>(hops), 0, &s.w) ||
      mg_walk_run(s.ctx, s.w, static_cast<uint32_t>(n_sources), sources, static_cast<uint32_t>(hops)) ||
      mg_walk_words_per_vertex(s.w, &words))
    return report(fn, mg_last_error());
  // Frontier rows are `words` wide and cover the padded strips; the caller's rows are
  // ceil(n_sources / 64) words for the n_vertices vertices.
  const size_t padded = (static_cast<size_t>(n_vertices) + 63) / 64 * 64;
  std::vector<uint64_t> x(padded * words);
  if (mg_walk_frontier_download(s.ctx, s.w, static_cast<uint32_t>(hops), x.data())) return report(fn, mg_last_error());
  const size_t used = (static_cast<size_t>(n_sources) + 63) / 64;
  uint64_t pc = 0;
  for (size_t v = 0; v < static_cast<size_t>(n_vertices); ++v)
    for (size_t j = 0; j < used; ++j) {
      const uint64_t val = x[v * words + j];
      x_out[v * used + j] = val;
      pc += static_cast<uint64_t>(__builtin_popcountll(val));
    }
  if (popcount) *popcount = pc;
  return 0;
}
