/*
 * matrixgraph_amd.h — C-ABI of the MI355X (gfx950) bit-tiled Boolean walk engine.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes
 * (MG_OK == 0). No entry point calls exit(); on failure the status is non-zero
 * and mg_last_error() returns a message (thread-local, valid until the next
 * failing call on the same thread).
 *
 * Tile layout (identical to the reference's BitTiledMatrix, tile_size 64):
 *   tile_offset_row[n_strips+1]  u32 row pointer of the tile-CSR, [0] == 0
 *   tile_row_idx[n_nz_tile]      u32 tile row of each tile
 *   tile_col_idx[n_nz_tile]      u32 tile column of each tile
 *   data[64 * n_nz_tile]         u64; tile t = words [64t, 64t+64);
 *                                word i = row i, bit j (LSB = column 0) = column j
 *   tiles listed in ascending (row, col) order.
 * Reference: core/data_structures/bit_tiled_matrix.cuh:17-72,
 *            core/data_structures/bit_tiled_matrix.cu:33-71,113-123,
 *            core/util/format_converter.cuh:292-415.
 *
 * Products follow the reference convention: the right operand is passed
 * TRANSPOSED (Bt = B^T), C = A * B over the Boolean semiring
 * (core/task/gpu_task/ppr_query.cu:63-71, :153-210).
 *
 * Threading: a context is bound to one HIP device and one stream; it is not
 * thread-safe (one context per host thread). Calls on a context are ordered
 * on its stream; functions that return host values synchronise that stream.
 */
#ifndef MATRIXGRAPH_AMD_H_
#define MATRIXGRAPH_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MG_OK = 0,
  MG_ERR_INVALID = 1,     /* bad argument, shape or handle */
  MG_ERR_HIP = 2,         /* HIP runtime error */
  MG_ERR_OOM = 3,         /* device or host allocation failed */
  MG_ERR_UNSUPPORTED = 4, /* valid in the reference but not supported here (e.g. tile_size != 64) */
  MG_ERR_IO = 5,          /* file missing, truncated or malformed */
  MG_ERR_INTERNAL = 6
};

#define MG_TILE_SIZE 64u

typedef struct mg_context mg_context;
typedef struct mg_tiled_matrix mg_tiled_matrix;
typedef struct mg_edge_list mg_edge_list;
typedef struct mg_walk mg_walk;

/* ---------------------------------------------------------------------------
 * Library and context
 * ------------------------------------------------------------------------- */

const char* mg_version(void);
const char* mg_last_error(void);

/* device < 0: take MATRIXGRAPH_HIP_DEVICE from the environment, else 0
 * (replaces set_device_from_env, core/go_api/matrixgraph_go_api.cu:23-36). */
int mg_context_create(int device, mg_context** out);
int mg_context_destroy(mg_context* ctx);
/* stream: a hipStream_t, or NULL for the context's own stream. Switching to another stream
 * first waits for the work queued on the previous one. */
int mg_context_set_stream(mg_context* ctx, void* stream);
int mg_context_synchronize(mg_context* ctx);
/* Device time on the context's stream between the two calls (HIP events; _stop waits for the
 * stream). Measurement helpers for the C-ABI's callers; no counterpart in the reference. */
int mg_context_timer_start(mg_context* ctx);
int mg_context_timer_stop(mg_context* ctx, double* ms);
int mg_context_device(const mg_context* ctx, int* device, int* n_compute_units);
/* The context's hipStream_t (for layer-1 calls ordered with the context's work). */
int mg_context_stream(const mg_context* ctx, void** stream);

/* Device buffers for C callers of layer 1, with DeviceOwnedBuffer semantics
 * (core/data_structures/device_buffer.cuh:15-151): mg_buffer_alloc zero-fills (Init(size)),
 * upload/download copy host <-> device on the context's stream and synchronise. */
int mg_buffer_alloc(mg_context* ctx, size_t bytes, void** dptr);
int mg_buffer_free(mg_context* ctx, void* dptr);
int mg_buffer_upload(mg_context* ctx, void* dptr, const void* host, size_t bytes);
int mg_buffer_download(mg_context* ctx, void* host, const void* dptr, size_t bytes);
/* Device-to-device copy on the context's stream (asynchronous). */
int mg_buffer_copy(mg_context* ctx, void* dst, const void* src, size_t bytes);

/* ---------------------------------------------------------------------------
 * Layer 1 — the reference's kernel-wrapper entry points, one for one.
 * Replaces sics::matrixgraph::core::task::kernel::MatrixOperationsKernelWrapper
 * (core/task/gpu_task/kernel/matrix_operations.cuh:14-124). All buffers are
 * DEVICE pointers owned by the caller; outputs that accumulate must be
 * zeroed by the caller exactly as with the reference (Init(size) zero-fills,
 * core/data_structures/device_buffer.cuh:92-106). stream is a hipStream_t
 * (NULL = default stream). Launches are asynchronous.
 * ------------------------------------------------------------------------- */

/* Boolean product of bit-packed layout matrices, inner-product form:
 *   c[i][j] |= OR_t a[i][t] & b[j][t]   (a: m x k, b: n x k, c: m x n, row-major bit-packed,
 *   element (i,j) of an r x s matrix at bit i*s+j).
 * Replaces MatrixBitAnd (matrix_operations.cuh:29-34, matrix_operations.cu:256-280,623-644). */
int mg_matrix_bit_and(void* stream, const uint64_t* matrix_a, const uint64_t* matrix_b,
                      uint64_t* matrix_c, uint32_t m, uint32_t k, uint32_t n);

/* *count += popcount of bits [0, size) of matrix.
 * Replaces MatrixBitCount (matrix_operations.cuh:36-40, matrix_operations.cu:282-296,646-661). */
int mg_matrix_bit_count(void* stream, const uint64_t* matrix, uint64_t* count, uint64_t size);

/* From an n_strips x n_strips layout bitmap: tile_count_row[x] = number of set bits in
 * row x (per-row COUNTS, as the reference writes into its "tile_offset_row" buffer; the
 * caller forms the prefix sum), tile_row_idx/tile_col_idx[rank] = (x, y) for the
 * rank-th set bit in row-major order.
 * Replaces InitBitTiledMatrixMetadataByLayoutMatrix (matrix_operations.cuh:42-48,
 * matrix_operations.cu:298-316,663-686); n_strips is passed explicitly (the reference
 * passes tile_size and loops over it). */
int mg_init_bit_tiled_matrix_metadata_by_layout_matrix(void* stream, const uint64_t* layout_matrix,
                                                       uint32_t n_strips, uint32_t* tile_count_row,
                                                       uint32_t* tile_row_idx, uint32_t* tile_col_idx);

/* Numeric pass: for every tile i of C (x, y) = (row_idx_c[i], col_idx_c[i]),
 *   data_c[i] |= OR over t with A(x,t) and Bt(y,t) present of  A(x,t) (x) Bt(y,t)
 * where (P = L (x) Rt)[r] bit c = ((L[r] & Rt[c]) != 0). OR-accumulates into data_c.
 * tile_size must be 64. layout_matrix_c and tile_row_idx_a/b may be NULL (unused).
 * Replaces FillTiles (matrix_operations.cuh:50-65, matrix_operations.cu:318-366,688-738). */
int mg_fill_tiles(void* stream, size_t tile_size, size_t n_strips, size_t n_nz_tile_a,
                  size_t n_nz_tile_b, size_t n_nz_tile_c, const uint64_t* layout_matrix_c,
                  const uint32_t* tile_offset_row_a, const uint32_t* tile_offset_row_b,
                  const uint32_t* tile_offset_row_c, const uint32_t* tile_row_idx_a,
                  const uint32_t* tile_row_idx_b, const uint32_t* tile_row_idx_c,
                  const uint32_t* tile_col_idx_a, const uint32_t* tile_col_idx_b,
                  const uint32_t* tile_col_idx_c, const uint64_t* data_a, const uint64_t* data_b,
                  uint64_t* data_c);

/* *count += popcount of the 64 * n_nz_tile words of data.
 * Replaces Count (matrix_operations.cuh:115-118, matrix_operations.cu:494-503,867-884). */
int mg_count(void* stream, size_t tile_size, size_t n_strips, size_t n_nz_tile, const uint64_t* data,
             uint64_t* count);

/* ---------------------------------------------------------------------------
 * Layer 2 — device-resident bit-tiled matrices (one block of a grid).
 * Replaces the host-pinned BitTiledMatrix (bit_tiled_matrix.cuh:17-72) for GPU work.
 * ------------------------------------------------------------------------- */

/* data == NULL uploads the index with all tiles zero (a structural result to be filled). */
int mg_tiled_matrix_upload(mg_context* ctx, uint32_t n_strips, uint32_t n_nz_tile, uint32_t tile_size,
                           const uint32_t* tile_offset_row, const uint32_t* tile_row_idx,
                           const uint32_t* tile_col_idx, const uint64_t* data, mg_tiled_matrix** out);
/* ---------------------------------------------------------------------------
 * On-disk format (host only, no GPU): the block layout of BitTiledMatrixIO
 * (core/io/bit_tiled_matrix_io.cu:30-136): dir/meta.yaml +
 * dir/meta_buf/{row_idx,col_idx,tile_offset_row,nz_tile_bm,data}.bin, native
 * little-endian, no header. dir must end with '/', as in the reference (file names are
 * appended). Missing, truncated or malformed files give MG_ERR_IO.
 * ------------------------------------------------------------------------- */
int mg_io_read_block_metadata(const char* dir, uint32_t* n_strips, uint32_t* n_nz_tile, uint32_t* tile_size);
/* Arrays sized from mg_io_read_block_metadata; NULL arrays are skipped. nz_tile_bm gets
 * n_strips*n_strips/64 + 1 words. A block with n_nz_tile 0 yields a zero row pointer. */
int mg_io_read_block(const char* dir, uint32_t* tile_offset_row, uint32_t* tile_row_idx, uint32_t* tile_col_idx,
                     uint64_t* data, uint64_t* nz_tile_bm);
/* Writes meta.yaml and the five .bin files (nz_tile_bm derived from the index). Empty
 * blocks are written too (the reference dereferences a NULL bitmap there). */
int mg_io_write_block(const char* dir, uint32_t n_strips, uint32_t n_nz_tile, uint32_t tile_size,
                      const uint32_t* tile_offset_row, const uint32_t* tile_row_idx, const uint32_t* tile_col_idx,
                      const uint64_t* data);
/* Grid root meta.yaml (GridGraphMetadata, core/io/grid_bit_tiled_matrix_io.cu:22-67);
 * blocks live in dir/block{x*n_chunks+y}/. */
int mg_io_read_grid_metadata(const char* dir, uint32_t* n_chunks, uint32_t* n_vertices, uint32_t* n_edges);
int mg_io_write_grid_metadata(const char* dir, uint32_t n_chunks, uint32_t n_vertices, uint32_t n_edges);
/* CSR-tiled blocks (host only), the input of the reference's enumerating walk:
 * CSRTiledMatrixIO (core/io/csr_tiled_matrix_io.cu:33-172) with per-tile ImmutableCSR
 * blobs (core/data_structures/immutable_csr.cu:79-101) as written by
 * Edgelist2CSRTiledMatrix (core/util/format_converter.cuh:417-755):
 * dir/meta.yaml (CSRTiledMatrixMetadata with its subgraphs list) +
 * dir/meta_buf/{row_idx,col_idx,tile_offset_row,nz_tile_bm,csr_offset,data}.bin.
 * The reader decodes every tile's edges back to global (src, dst) ids; it relies on the
 * subgraphs list and csr_offset, not on tile order. n_edges = sum of num_outgoing_edges. */
int mg_io_read_csr_tiled_block_info(const char* dir, uint32_t* n_strips, uint32_t* n_nz_tile, uint32_t* tile_size,
                                    uint64_t* n_edges);
/* Fills up to capacity edges (MG_ERR_IO if the block holds more); tile_row/tile_col (the
 * tile of each edge) may be NULL, as may src/dst. */
int mg_io_read_csr_tiled_block(const char* dir, uint32_t* src, uint32_t* dst, uint64_t capacity, uint32_t* tile_row,
                               uint32_t* tile_col);
/* One block's edges (global ids, all in one grid block) -> dir in the same layout:
 * n_strips = ceil(block_scope / tile_size), tile_size being the tile's extent in vertices
 * (the reference converter turns a tile_size argument below 128 into
 * ceil(block_scope / tile_size) before this point). */
int mg_io_write_csr_tiled_block(const char* dir, uint32_t block_scope, uint32_t tile_size, const uint32_t* src,
                                const uint32_t* dst, uint64_t n_edges);
/* Grid root of a CSR-tiled grid (core/io/grid_csr_tiled_matrix_io.cu; written by
 * tools/graph_converter/converter/to_grid_csr_tiled_matrix.cuh:54-85): meta.yaml
 * GridGraphMetadata {n_vertices, n_edges, n_chunks, max_vid} plus label.bin (u32 per
 * vertex; the writer stores zeros, as the converter does). */
int mg_io_read_csr_tiled_grid_metadata(const char* dir, uint32_t* n_chunks, uint32_t* n_vertices, uint32_t* n_edges,
                                       uint32_t* max_vid);
int mg_io_write_csr_tiled_grid_metadata(const char* dir, uint32_t n_chunks, uint32_t n_vertices, uint32_t n_edges,
                                        uint32_t max_vid);
int mg_tiled_matrix_metadata(const mg_tiled_matrix* m, uint32_t* n_strips, uint32_t* n_nz_tile,
                             uint32_t* tile_size);
/* Copies any non-NULL array back to the host. nz_tile_bm receives
 * n_strips*n_strips/64 + 1 words (bit x*n_strips+y set iff tile (x,y) present),
 * the layout of meta_buf/nz_tile_bm.bin (core/io/bit_tiled_matrix_io.cu:62-64). */
int mg_tiled_matrix_download(mg_context* ctx, const mg_tiled_matrix* m, uint32_t* tile_offset_row,
                             uint32_t* tile_row_idx, uint32_t* tile_col_idx, uint64_t* data,
                             uint64_t* nz_tile_bm);
/* Device addresses of the four arrays (borrowed; valid while m lives). */
int mg_tiled_matrix_device_arrays(const mg_tiled_matrix* m, const uint32_t** tile_offset_row,
                                  const uint32_t** tile_row_idx, const uint32_t** tile_col_idx,
                                  const uint64_t** data);
/* Order-independent 64-bit digest of (index, data): equal matrices give equal digests. */
int mg_tiled_matrix_digest(mg_context* ctx, const mg_tiled_matrix* m, uint64_t* digest);
/* Total set bits (the reference's total_count, ppr_query.cu:874-956). */
int mg_tiled_matrix_popcount(mg_context* ctx, const mg_tiled_matrix* m, uint64_t* count);
int mg_tiled_matrix_destroy(mg_tiled_matrix* m);
/* Result egress: every set bit (row r, column c) of tile (x, y) as the edge
 * (x*64 + r, y*64 + c), sorted, on n_strips*64 vertices (the tile -> edge-list expansion
 * the reference leaves to its callers, core/task/gpu_task/ppr_query.cu:874-956 counts only).
 * For an exported X_k the edge (s, v) says source s reaches vertex v. */
int mg_tiled_matrix_to_edges(mg_context* ctx, const mg_tiled_matrix* m, mg_edge_list** out);

/* Two-hop walk in enumerating form (the result of the reference's CSR-tiled GEMM walk,
 * core/task/gpu_task/gemm.cu:452-794 and walk_kernel, matrix_operations.cu:505-621, whose
 * output copy is commented out): the distinct pairs (s, v) with s -> m -> v, m != s, m != v
 * (self-loops are never a step) and v != s, as a sorted edge list on el's vertices. Computed
 * on the bit tiles: A' = A without self-loops, C = A' * A', C without its diagonal. */
int mg_edges_two_hop(mg_context* ctx, const mg_edge_list* el, mg_edge_list** out);

typedef struct mg_product_stats {
  uint64_t tile_products;  /* (r,k,c) with A(r,k) and B(k,c) both present */
  uint64_t out_tiles;      /* structural tiles of C */
  uint64_t right_tiles_used;
  uint64_t alg_bytes;      /* 520*(T_A + T_B,used + T_C) + 4*(2*n_strips + 2) */
} mg_product_stats;

/* Symbolic pass: allocates C with the structural tile set of A*B (tile (x,y) present iff
 * some t has A(x,t) and Bt(y,t) present), tiles zeroed. Both operands share n_strips.
 * Replaces PPRQuery::InitResultMatrix (ppr_query.cu:73-457): MatrixBitAnd + MatrixBitCount +
 * InitBitTiledMatrixMetadataByLayoutMatrix + host prefix sum. Synchronises once (to read the
 * tile count); C's index and zero fill complete asynchronously on the context's stream, and
 * the temporaries are stream-ordered pool allocations. stats may be NULL. */
int mg_multiply_symbolic(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt,
                         mg_tiled_matrix** c, mg_product_stats* stats);
/* Keeps m's per-tile payloads (16 bytes) and set-bit counts (4 bytes) with the handle, so the
 * products that take m as an operand (mg_multiply_numeric*, mg_multiply_pruned*) copy them
 * instead of reading m's tiles again: a walk by products, X_{k+1} = X_k * A, takes A^T at every
 * hop. No counterpart in the reference (its products re-read every tile). The kept copy is
 * dropped by the next library call that writes m's tiles (a numeric pass into m); after
 * writing m's arrays through the pointers of mg_tiled_matrix_device_arrays, call it again.
 * A call while the kept copy is current does nothing; otherwise it synchronises the context's
 * stream. */
int mg_tiled_matrix_prepare_operand(mg_context* ctx, mg_tiled_matrix* m);
/* Numeric pass into a C produced by mg_multiply_symbolic (OR-accumulates).
 * Replaces PPRQuery::FillTilesUnifiedMemory (ppr_query.cu:459-780). Asynchronous on the
 * context's stream; keeps 16 bytes per operand tile in the context's scratch, which grows in
 * stream order (hipMallocAsync/hipFreeAsync on that stream: queued work that reads the old
 * buffer finishes before its memory is reused). */
int mg_multiply_numeric(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt,
                        mg_tiled_matrix* c);
/* The numeric pass with a per-tile-pair choice of engine: a pair whose two tiles both have
 * at least min_bits set bits is multiplied on the matrix cores; every other pair on the
 * vector ALUs. The dense tiles of each operand are expanded once to one byte per bit (4 KiB
 * per tile, at most 32 GiB per operand), the vector pass records the dense pairs of every C
 * tile, and a second kernel sums each C tile's dense pairs in int32 (int8 0/1 products) and
 * turns counts > 0 into bits once. min_bits = 0 sends every pair to the matrix cores,
 * min_bits > 4096 none. *dense_products (may be NULL) receives the number of matrix-core tile
 * products. Same result as mg_multiply_numeric. Synchronises.
 * No counterpart in the reference (its product is VALU only). */
int mg_multiply_numeric_dense(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt,
                              mg_tiled_matrix* c, uint32_t min_bits, uint64_t* dense_products);
/* Pruned product: C = A*B holding only its numerically non-empty tiles (the structural tile
 * set of mg_multiply_symbolic less the tiles whose product is zero), row-major; every stored
 * tile is bit-identical to the same tile of mg_multiply_symbolic + mg_multiply_numeric. No
 * counterpart in the reference, which allocates every structural tile
 * (ppr_query.cu:288-314). capacity chooses how the output is sized:
 *   MG_PRUNED_EXACT (0): a flag pass finds the non-empty tiles, a second pass writes each at
 *     its final position (two numeric passes, memory = the output);
 *   MG_PRUNED_STRUCTURAL: one pass into scratch for every structural tile, then sorted;
 *   any other value (e.g. a cost model's prediction): one pass into that many slots, then
 *     sorted; if more tiles are non-empty the pass is run again with the exact count (fallback).
 * Synchronises. stats may be NULL. */
#define MG_PRUNED_EXACT 0ull
#define MG_PRUNED_STRUCTURAL (~0ull)
typedef struct mg_pruned_stats {
  uint64_t structural_tiles; /* structural tiles of C */
  uint64_t numeric_tiles;    /* stored (non-empty) tiles */
  uint64_t tile_products;    /* structural (r,k,c) triples */
  uint64_t capacity;         /* output slots of the first pass */
  uint64_t scratch_bytes;    /* largest pass scratch (0 for MG_PRUNED_EXACT) */
  uint64_t output_bytes;     /* tile data of the result */
  uint32_t passes;           /* numeric passes run */
  uint32_t fallback;         /* 1: the capacity was too small and the pass ran again */
} mg_pruned_stats;
int mg_multiply_pruned(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt, uint64_t capacity,
                       mg_tiled_matrix** c, mg_pruned_stats* stats);
/* Features of every row x of C = A*B known after the symbolic pass, MG_ROW_FEATURES floats per
 * row into features[x * MG_ROW_FEATURES ...]: log1p of the structural tiles / 16, of the tile
 * products / 20, of A's tiles in row x / 16, of their set bits / 20, of the set bits of the B
 * rows they reach / 24; log1p of the mean set bits per A tile / 8 and per reached B tile / 8;
 * structural tiles / n_strips. structural (may be NULL) receives the structural tiles per
 * row, *tile_products (may be NULL) their total. No counterpart in the reference. */
#define MG_ROW_FEATURES 8
int mg_multiply_row_features(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt,
                             float* features, uint32_t* structural, uint64_t* tile_products);
/* Allocation-free numeric pass: same result as mg_multiply_numeric, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). No counterpart in the reference. Makes no hipMalloc/hipFree call: the
 * operand payloads go to a static per-device array of MG_NOALLOC_SCRATCH_BYTES (held from
 * library load until the process ends), so A and Bt may hold at most MG_NOALLOC_MAX_TILES
 * tiles together; larger operands give MG_ERR_INVALID. Asynchronous on the context's stream.
 * Every call on one device shares that array: calls that overlap on one device (other streams
 * or contexts) race. */
#define MG_NOALLOC_SCRATCH_BYTES (32u << 20)
#define MG_NOALLOC_MAX_TILES ((MG_NOALLOC_SCRATCH_BYTES / 16u) - 1u)
int mg_multiply_numeric_noalloc(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt,
                                mg_tiled_matrix* c);

/* ---------------------------------------------------------------------------
 * Edge lists and the edge-list -> bit-tile converter (input production).
 * Replaces Edges + Edgelist2BitTiledMatrix (core/data_structures/edgelist.h:22-233,
 * core/util/format_converter.cuh:292-415) and Edges::Transpose (edgelist.cu:367-379).
 * Device edge lists are kept sorted by (src, dst) and duplicate-free.
 * ------------------------------------------------------------------------- */

/* R-MAT, 2^scale vertices, edge_factor * 2^scale draws; quadrant probabilities a, b, c
 * (d = 1-a-b-c); duplicates and self-loops removed. Counter-based SplitMix64 streams. */
int mg_edges_rmat(mg_context* ctx, uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                  uint64_t seed, mg_edge_list** out);
/* Directed G(n, p), n = 2^log2_n, p = p_num / 2^p_log2_den exactly; no self-loops. */
int mg_edges_er(mg_context* ctx, uint32_t log2_n, uint64_t p_num, uint32_t p_log2_den, uint64_t seed,
                mg_edge_list** out);
/* n_comm planted directed communities on n_vertices vertices: community c covers
 * [c*stride, c*stride+comm_size); each ordered pair u != v inside one is an edge with
 * p = p_num / 2^p_log2_den exactly. */
int mg_edges_communities(mg_context* ctx, uint32_t n_vertices, uint32_t n_comm, uint32_t comm_size,
                         uint32_t stride, uint64_t p_num, uint32_t p_log2_den, uint64_t seed,
                         mg_edge_list** out);
/* Union (set semantics) of two edge lists on the same vertex count. */
int mg_edges_union(mg_context* ctx, const mg_edge_list* a, const mg_edge_list* b, mg_edge_list** out);
/* Host edges -> device list, sorted and de-duplicated; self-loops KEPT (the reference's
 * converter keeps them). */
int mg_edges_upload(mg_context* ctx, uint32_t n_vertices, uint64_t n_edges, const uint32_t* src,
                    const uint32_t* dst, mg_edge_list** out);
int mg_edges_info(const mg_edge_list* el, uint32_t* n_vertices, uint64_t* n_edges);
int mg_edges_download(mg_context* ctx, const mg_edge_list* el, uint32_t* src, uint32_t* dst);
int mg_edges_digest(mg_context* ctx, const mg_edge_list* el, uint64_t* digest);
/* transpose != 0 builds A^T (tile (y,x) from edge (x,y)), as Edges::Transpose then convert.
 * n_strips = ceil(n_vertices / 64), block scope = n_vertices (n_chunks = 1). */
int mg_edges_to_tiled_matrix(mg_context* ctx, const mg_edge_list* el, int transpose,
                             mg_tiled_matrix** out);
/* Grid-cut partition (replaces GridCutPartitioner::RunPartitioner,
 * tools/graph_partitioner/partitioner/grid_cut.cu:38-206, with the block conversion of
 * tools/graph_converter/converter/to_bit_tiled_matrix.cuh:34-74): block (bx, by) of the
 * n_chunks x n_chunks grid holds the edges (u, v) with u / S == bx and v / S == by, in local
 * ids (u - bx*S, v - by*S), as an edge list on S vertices, S = ceil(n_vertices / n_chunks) for
 * both the partition and the conversion (the reference's partitioner uses
 * ceil((max_vid+1)/n) and its converter num_vertices/n, grid_cut.cu:88-89 and
 * to_bit_tiled_matrix.cuh:50, which place tiles wrongly unless n divides the vertex count).
 * mg_edges_to_tiled_matrix(block, 0) is then block (bx, by) of the grid of A, and
 * mg_edges_to_tiled_matrix(block of (by, bx), 1) is block (bx, by) of the grid of A^T. */
int mg_edges_grid_block(mg_context* ctx, const mg_edge_list* el, uint32_t n_chunks, uint32_t bx, uint32_t by,
                        mg_edge_list** out);
int mg_edges_destroy(mg_edge_list* el);

/* count distinct vertices in [0, n_vertices) by Floyd's sampling over SplitMix64
 * (pure host function, no GPU). Order of out[] = row order of X_0. */
int mg_sample_sources(uint32_t n_vertices, uint32_t count, uint64_t seed, uint32_t* out);

/* ---------------------------------------------------------------------------
 * Walks: X_{k+1} = X_k * A with the right operand supplied as A^T, X_0 one-hot rows
 * (exact-length walks; X_k is not OR-ed into X_{k+1}). Generalises the
 * PPRQuery pipeline (ppr_query.cu:958-989) to a hop loop with X_k device-resident.
 *
 * Device layout of X_k: vertex-major ("strip-interleaved") — vertex v's source set is
 * W = words_per_vertex consecutive u64 (bit s%64 of word s/64 = source s). A strip of
 * 64 vertices is 64*W words: the 64 bit tiles X_k(x, strip) for x < W, word-interleaved.
 * mg_walk_export returns X_k in the reference tile layout and orientation.
 * ------------------------------------------------------------------------- */

enum {
  MG_WALK_PRUNE = 1u,          /* structural tile set of X_{k+1} built from numerically
                                  non-empty X_k tiles instead of structural ones */
  MG_WALK_SWEEP_FIRST_HOP = 2u, /* never push the first hop: every hop sweeps all strips
                                   (no column index of A^T is built) */
  MG_WALK_VERTEX_ORDER = 4u     /* keep the frontiers' strips in vertex order; by default the
                                   engine stores and walks them in a fixed permuted order
                                   (results, digests, downloads and exports are the same) */
};
enum {
  MG_HOP_SWEEP = 0u, /* the hop swept every strip of A^T (walk_numeric kernels) */
  MG_HOP_PUSH = 1u,  /* first hop pushed from the source strips' A^T columns */
  MG_HOP_JOBS = 2u   /* the hop read A^T's job list by row segments (mg_walk_set_hop_kernel) */
};

typedef struct mg_hop_stats {
  uint64_t tile_products;     /* structural (r,k,c) triples of this hop */
  uint64_t left_tiles;        /* T_L: structural tiles of X_{k} */
  uint64_t right_tiles_used;  /* T_R,used: tiles of A in rows referenced by X_k */
  uint64_t out_tiles;         /* T_C: structural tiles of X_{k+1} */
  uint64_t out_tiles_numeric; /* numerically non-empty tiles of X_{k+1} */
  uint64_t popcount;          /* set bits of X_{k+1} */
  uint64_t alg_bytes;         /* 520*(T_L+T_R,used+T_C) + 4*(2*rows+2), rows = tile rows of X */
  uint64_t edges_visited;     /* set bits of A read by the numeric pass */
  double symbolic_ms;         /* 0: the structural pass runs inside the numeric kernel */
  double numeric_ms;          /* hop kernels (structural + numeric) of the last logged run, else 0 */
  uint32_t kernel;            /* MG_HOP_SWEEP, MG_HOP_PUSH or MG_HOP_JOBS */
  uint32_t segments;          /* MG_HOP_JOBS: row segments of the job-list kernel, else 0 */
  /* Single walks only (batch query statistics leave them 0): the same counts over
   * numerically non-empty tiles. */
  uint64_t left_tiles_numeric;       /* numerically non-empty tiles of X_k */
  uint64_t right_tiles_used_numeric; /* tiles of A in rows referenced by those tiles */
  uint64_t alg_bytes_numeric;        /* 520*(left_numeric + right_used_numeric + out_tiles_numeric)
                                        + 4*(2*rows+2) */
} mg_hop_stats;

/* Allocates X_0..X_max_hops (W = ceil(max_sources/64) rounded up to a power of two,
 * max_sources <= 4096) plus the strip scheduler, the per-tile row-occupancy masks of at
 * and (unless MG_WALK_SWEEP_FIRST_HOP) a column index of at (16 bytes per tile). at must
 * outlive w and keep its contents (the masks are taken here).
 * The first hop of a run is pushed from the source strips when their columns of at hold at
 * most half of its tiles; results and statistics are the same either way. */
int mg_walk_create(mg_context* ctx, const mg_tiled_matrix* at, uint32_t max_sources,
                   uint32_t max_hops, uint32_t flags, mg_walk** out);
/* Hop kernel of the 4096-source (64-word) levels from hop 2 on. segments = 0: the
 * tile-decoding sweep (every hop decodes the A^T tiles' payloads into gathers of whole
 * 512-byte source rows); 1, 2, 4 or 8: the job-list kernel, which reads a list of at's set
 * bits built once (4 bytes per set bit plus 4 per tile and per strip, on first use) and
 * splits each source row into that many segments, one per workgroup. With 1 segment the
 * kernel gathers whole rows, once per (strip, source) pair, with eight waves per workgroup.
 * The default is 1 (environment MG_WALK_SEGMENTS overrides it at creation). Walks of fewer
 * than 2049 sources, and graphs above 2^26 vertices or 2^32 - 65536 set bits, keep the
 * sweep. Results and statistics are the same for every choice. Synchronises the context's
 * stream. */
int mg_walk_set_hop_kernel(mg_context* ctx, mg_walk* w, uint32_t segments);
/* The segments in effect (0: sweep). */
int mg_walk_hop_kernel(const mg_walk* w, uint32_t* segments);
/* Runs hops 1..hops from the given distinct sources (async; stats after sync). */
int mg_walk_run(mg_context* ctx, mg_walk* w, uint32_t n_sources, const uint32_t* sources, uint32_t hops);
/* enable != 0 starts (and clears) a log of HIP events around the hop kernel of every hop
 * of every later run; enable == 0 stops logging (the log stays readable). */
int mg_walk_set_timing(mg_walk* w, int enable);
/* Sums over the logged runs that reached `hop`: symbolic-pass (0, fused) and hop-kernel
 * device milliseconds, and the number of such runs. Synchronises. */
int mg_walk_kernel_times(mg_context* ctx, mg_walk* w, uint32_t hop, double* symbolic_ms, double* numeric_ms,
                         uint32_t* n_runs);
/* hop in [1, hops]; synchronises. */
int mg_walk_hop_stats(mg_context* ctx, mg_walk* w, uint32_t hop, mg_hop_stats* out);
int mg_walk_words_per_vertex(const mg_walk* w, uint32_t* words);
/* hop in [0, hops]; xt receives n_strips*64*W words (vertex-major). */
int mg_walk_frontier_download(mg_context* ctx, const mg_walk* w, uint32_t hop, uint64_t* xt);
/* Order-independent digest of X_hop restricted to the first n_sources source bits. */
int mg_walk_frontier_digest(mg_context* ctx, const mg_walk* w, uint32_t hop, uint64_t* digest);
/* X_hop in the reference layout: n_strips x n_strips tile grid (padded square), tile rows
 * 0..ceil(n_sources/64)-1 used, structural tile set (or numeric with MG_WALK_PRUNE). */
int mg_walk_export(mg_context* ctx, const mg_walk* w, uint32_t hop, mg_tiled_matrix** out);
int mg_walk_destroy(mg_walk* w);

/* ---------------------------------------------------------------------------
 * Host-pointer surface for foreign callers (cgo and the like), in the style of the
 * reference's Go API (core/go_api/matrixgraph_go_api.h:8-55): host memory only, 0 on
 * success and 1 on error with the message on stderr, device from MATRIXGRAPH_HIP_DEVICE
 * (else 0); each call sets up, runs and releases its own context.
 * ------------------------------------------------------------------------- */

/* C = A * B over the Boolean semiring with B passed transposed (the bit path's convention):
 * A m x k, Bt n x k, C m x n, each packed row-major as one bit string (element (i, j) of an
 * r x s matrix at bit i*s + j, ceil(r*s/64) words). C is overwritten. Bit-matrix counterpart
 * of matrixgraph_matmult (matrixgraph_go_api.h:11, matrixgraph_go_api.cu:42-104). */
int matrixgraph_bit_matmult(const uint64_t* A, const uint64_t* Bt, uint64_t* C, int m, int k, int n);

/* X_hops of an exact-length walk from n_sources distinct sources (<= 4096) over the directed
 * graph (src[i] -> dst[i]) on n_vertices vertices: x_out receives n_vertices rows of
 * ceil(n_sources/64) words (bit s % 64 of word s / 64 of row v set iff vertex v is reached
 * from sources[s] by a walk of exactly `hops` edges); popcount (may be NULL) gets its set bits. */
int matrixgraph_bit_walk(const uint32_t* src, const uint32_t* dst, int n_edges, int n_vertices,
                         const uint32_t* sources, int n_sources, int hops, uint64_t* x_out, uint64_t* popcount);

/* ---------------------------------------------------------------------------
 * Batched walks: many queries of uneven source-set size and hop count over several
 * graphs in one launch sequence, captured once as a HIP graph and replayed by
 * mg_walk_batch_run. Queries are packed at 64-source word boundaries into 4096-source
 * panels per graph; every hop level of every panel is one set of kernel launches, and
 * a query's words stop propagating after its own hop count, so per-query results are
 * bit-identical to running the query alone with mg_walk_run.
 * ------------------------------------------------------------------------- */
typedef struct mg_walk_batch mg_walk_batch;
typedef struct mg_walk_query {
  uint32_t graph;          /* index into the graphs array of mg_walk_batch_create */
  uint32_t n_sources;      /* 1..4096 distinct source vertices */
  const uint32_t* sources; /* read during mg_walk_batch_create only */
  uint32_t hops;           /* 1..64 */
} mg_walk_query;
enum {
  MG_PACK_BY_HOPS = 0,      /* per graph, descending hop count, first fit into panels */
  MG_PACK_ONE_PER_PANEL = 1 /* every query in its own panel (no sharing) */
};
typedef struct mg_batch_options {
  uint32_t packing;     /* MG_PACK_* */
  uint32_t chunk_tiles; /* scheduler chunk size in A^T tiles (0 = default 2048) */
  uint32_t use_graph;   /* non-zero: capture once, replay as a HIP graph */
  uint32_t flags;       /* MG_BATCH_* */
} mg_batch_options;
enum {
  MG_BATCH_SWEEP_FIRST_HOP = 1u, /* never push a panel's first hop: every level-0 chunk is swept
                                    (and logged); no column index of A^T is built */
  MG_BATCH_VERTEX_ORDER = 2u     /* as MG_WALK_VERTEX_ORDER, for every graph of the batch */
};

/* graphs_at[g] is graph g supplied transposed (A^T); they must outlive the batch and keep
 * their contents (row-occupancy masks are taken here).
 * opt may be NULL (packing by hops, default chunks, HIP graph). */
int mg_walk_batch_create(mg_context* ctx, const mg_tiled_matrix* const* graphs_at, uint32_t n_graphs,
                         const mg_walk_query* queries, uint32_t n_queries, const mg_batch_options* opt,
                         mg_walk_batch** out);
/* Runs every query (asynchronous; the first call captures the graph). */
int mg_walk_batch_run(mg_context* ctx, mg_walk_batch* b);
int mg_walk_batch_info(const mg_walk_batch* b, uint32_t* n_panels, uint32_t* max_hops, uint64_t* n_chunks);
/* Panels whose first hop is pushed from their source strips' A^T columns (decided per panel
 * at creation, when those columns hold at most half of the graph's tiles); the level-0
 * chunk records of mg_walk_batch_level_chunks cover the other panels only. */
int mg_walk_batch_pushed_panels(const mg_walk_batch* b, uint32_t* n_panels);
/* Counters of the panel holding query q at hop (its tile products cover all queries of
 * that panel). */
int mg_walk_batch_query_stats(mg_context* ctx, const mg_walk_batch* b, uint32_t q, uint32_t hop, mg_hop_stats* out);
/* Sums over all panels and hops (structural tile products, B_alg, visited A^T bits). */
int mg_walk_batch_totals(mg_context* ctx, const mg_walk_batch* b, uint64_t* tile_products, uint64_t* alg_bytes,
                         uint64_t* edges_visited);
/* Query q's X_hops: digest equal to mg_walk_frontier_digest of the same query run alone;
 * download gives n_vertices x ceil(n_sources/64) words (vertex-major); export gives the
 * reference tile layout as mg_walk_export does. */
int mg_walk_batch_query_digest(mg_context* ctx, const mg_walk_batch* b, uint32_t q, uint64_t* digest);
int mg_walk_batch_query_download(mg_context* ctx, const mg_walk_batch* b, uint32_t q, uint64_t* out);
int mg_walk_batch_query_export(mg_context* ctx, const mg_walk_batch* b, uint32_t q, mg_tiled_matrix** out);
int mg_walk_batch_destroy(mg_walk_batch* b);

/* Scheduler control for cost models (no counterpart in the reference). A hop level's work is
 * a list of chunks (strip ranges of one panel's A^T tile stream) launched in list order.
 * Logging, when enabled, records every chunk's outcome on each later run. */
typedef struct mg_chunk_record {
  uint32_t panel;                   /* panel the chunk belongs to */
  uint32_t strip_begin, strip_end;  /* target strips */
  uint32_t tile_begin, tile_end;    /* A^T tiles (positions in the tile stream) */
  uint32_t bits;                    /* set bits of those tiles */
  uint32_t jobs;                    /* logged: gathers done (bits with a non-empty source) */
  uint32_t out_tiles;               /* logged: numerically non-empty output tiles (unsplit strips) */
} mg_chunk_record;
int mg_walk_batch_set_logging(mg_walk_batch* b, int enable);
/* Hop kernel of the batch's levels after the first (see mg_walk_set_hop_kernel); levels
 * run while chunk logging is on use the sweep, which logs. Drops the captured graph. */
int mg_walk_batch_set_hop_kernel(mg_context* ctx, mg_walk_batch* b, uint32_t segments);
/* level in [0, max_hops): the X_level -> X_level+1 launch. out == NULL only sets *n.
 * Synchronises. */
int mg_walk_batch_level_chunks(mg_context* ctx, const mg_walk_batch* b, uint32_t level, uint64_t* n,
                               mg_chunk_record* out);
/* Panel p: its graph, hop count, words holding sources, and alive[k] (k < hops): the words
 * that propagate from X_k to X_k+1. */
int mg_walk_batch_panel_info(const mg_walk_batch* b, uint32_t p, uint32_t* graph, uint32_t* hops,
                             uint32_t* words_used, uint64_t* alive);
/* Reorders level's chunks: the i-th launched chunk becomes the current chunk order[i]
 * (order is a permutation of 0..n-1). Results do not depend on the order. */
int mg_walk_batch_set_level_order(mg_walk_batch* b, uint32_t level, const uint32_t* order, uint64_t n);

/* ---------------------------------------------------------------------------
 * Learned cost model: MLP n_in -> 64 -> 64 -> n_out (ReLU), BF16 weights rounded to
 * nearest-even from FP32 masters, FP32 accumulation, hand-written forward / backward /
 * Adam kernels with deterministic gradient reduction. Supported shapes: n_in in
 * {4, 8, 16}, n_out in {1, 2}. Sample arrays (x: n x n_in, t and y: n x n_out, FP32,
 * row-major) are DEVICE pointers. Parameter layout: W1[64][n_in], b1[64], W2[64][64],
 * b2[64], W3[n_out][64], b3[n_out]. No counterpart in the reference (new subsystem).
 * ------------------------------------------------------------------------- */
typedef struct mg_mlp mg_mlp;
int mg_mlp_create(mg_context* ctx, uint32_t n_in, uint32_t n_out, uint64_t seed, mg_mlp** out);
int mg_mlp_num_params(const mg_mlp* p, uint32_t* n_params);
/* master: FP32 masters (n_params); bf16: the BF16 bit patterns used by the kernels. */
int mg_mlp_get_params(mg_context* ctx, const mg_mlp* p, float* master, uint16_t* bf16);
int mg_mlp_set_params(mg_context* ctx, mg_mlp* p, const float* master);
int mg_mlp_forward(mg_context* ctx, const mg_mlp* p, const float* x, uint32_t n, float* y);
/* mg_multiply_pruned sized by a cost model: after the structural pass, the model (n_in =
 * MG_ROW_FEATURES; output 0 = predicted fraction of a row's structural tiles that are
 * non-empty) runs on the rows' features on the device, and the first numeric pass gets
 * sum over rows of min(structural, ceil(structural * prediction * (1 + margin))) slots; a
 * prediction below the non-empty count makes the pass run again at the exact count (the
 * exact fallback), so the result is always that of mg_multiply_pruned. row_predicted (host,
 * n_strips floats, may be NULL) receives structural * prediction per row. No counterpart in the
 * reference. */
int mg_multiply_pruned_model(mg_context* ctx, const mg_tiled_matrix* a, const mg_tiled_matrix* bt,
                             const mg_mlp* model, float margin, float* row_predicted, mg_tiled_matrix** c,
                             mg_pruned_stats* stats);
/* loss = mean over samples and outputs of (y - t)^2; grad (host, n_params) of that loss
 * with respect to the BF16 weights. Either output may be NULL. Synchronises. */
int mg_mlp_loss_grad(mg_context* ctx, mg_mlp* p, const float* x, const float* t, uint32_t n, float* loss, float* grad);
/* Adam (lr, 0.9, 0.999, 1e-8) over mini-batches of `batch` samples in array order. */
int mg_mlp_train(mg_context* ctx, mg_mlp* p, const float* x, const float* t, uint32_t n, uint32_t epochs,
                 uint32_t batch, float lr, float* final_loss);
int mg_mlp_destroy(mg_mlp* p);

#ifdef __cplusplus
}
#endif

#endif /* MATRIXGRAPH_AMD_H_ */
