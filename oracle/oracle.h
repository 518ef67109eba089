/*
 * CPU oracle for the bit-tiled Boolean walk path. TEST INFRASTRUCTURE ONLY:
 * used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker; never linked into or called by the product library.
 *
 * Every function restates the reference's intended semantics (Appendix-C style:
 * the Boolean-semiring product over 64x64 bit tiles) or a published algorithm,
 * and cites the reference file:line it follows. Pins (what checks each function):
 *   or_bool_matmul_bruteforce  O1: textbook triple loop; checked against closed forms
 *                              (identity, zero, permutations, cycles) in tests.
 *   or_bitset_matmul           O2: checked against O1 on random tiny inputs.
 *   or_frontier_step           O3: checked against O2 (k-fold products) and cycle/path
 *                              closed forms.
 *   or_tile_symbolic/numeric   O4: checked against O1/O2 through dense expansion.
 *   or_walk_tile_hop           CPU tile algorithm (the timed CPU baseline); checked
 *                              against O3.
 *   generators / converter     checked by invariants (no self-loops, sortedness,
 *                              uniqueness, edge count window), tile round trip
 *                              edges -> tiles -> edges, and the toy-graph fixture.
 */
#ifndef MG_ORACLE_H_
#define MG_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* SplitMix64 (Steele, Lea, Flood 2014): or_sm64_at(key, i) = output i+1 of the
 * generator seeded with key. */
uint64_t or_sm64_mix(uint64_t z);
uint64_t or_sm64_at(uint64_t key, uint64_t i);
uint64_t or_stream_key(uint64_t seed, uint64_t tag);
uint64_t or_digest_term(uint64_t index, uint64_t value);

/* Generators: sorted by (src, dst), unique, self-loops removed. Return the edge
 * count, or -1 if cap is too small. */
int64_t or_rmat_edges(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
                      uint32_t* src, uint32_t* dst, uint64_t cap);
int64_t or_er_edges(uint32_t log2_n, uint64_t p_num, uint32_t p_log2_den, uint64_t seed, uint32_t* src,
                    uint32_t* dst, uint64_t cap);
int64_t or_community_edges(uint32_t n_vertices, uint32_t n_comm, uint32_t comm_size, uint32_t stride,
                           uint64_t p_num, uint32_t p_log2_den, uint64_t seed, uint32_t* src, uint32_t* dst,
                           uint64_t cap);
/* Sort + de-duplicate an edge list in place (keeps self-loops); returns new count. */
uint64_t or_sort_unique_edges(uint64_t n, uint32_t* src, uint32_t* dst);
uint64_t or_edges_digest(uint64_t n, const uint32_t* src, const uint32_t* dst);
void or_sample_sources(uint32_t n_vertices, uint32_t count, uint64_t seed, uint32_t* out);

/* Edge list -> bit tiles (core/util/format_converter.cuh:292-415, n_chunks = 1). */
uint64_t or_count_tiles(uint32_t n_vertices, uint64_t n_edges, const uint32_t* src, const uint32_t* dst,
                        int transpose);
void or_edges_to_tiles(uint32_t n_vertices, uint64_t n_edges, const uint32_t* src, const uint32_t* dst,
                       int transpose, uint32_t* off, uint32_t* row, uint32_t* col, uint64_t* data);
uint64_t or_tiles_digest(uint32_t n_strips, uint32_t n_nz_tile, const uint32_t* row, const uint32_t* col,
                         const uint64_t* data);

/* O1: c[i*n+j] = OR_k a[i*n+k] & b[k*n+j] over uint8 0/1 matrices. */
void or_bool_matmul_bruteforce(uint32_t n, const uint8_t* a, const uint8_t* b, uint8_t* c);
/* O2: bitset rows. a: m rows of k bits (ka = ceil(k/64) words per row); b: k rows of n bits
 * (nb = ceil(n/64) words per row); c: m rows of n bits; c = a*b. */
void or_bitset_matmul(uint32_t m, uint32_t k, uint32_t n, const uint64_t* a, const uint64_t* b, uint64_t* c,
                      int threads);
/* O3: vertex-major frontier step y[v] = OR_{u in in(v)} x[u] (W words per vertex),
 * in-CSR in_off[n+1], in_src[]. */
void or_frontier_step(uint32_t n, const uint64_t* in_off, const uint32_t* in_src, uint32_t W, const uint64_t* x,
                      uint64_t* y, int threads);
uint64_t or_frontier_digest(uint64_t n_vertices, uint32_t W, uint32_t n_words, const uint64_t* x);
/* In-neighbour CSR of an edge list (counting sort on dst; in_src ascending by src when
 * the input is sorted by (src, dst)). */
void or_in_csr(uint32_t n_vertices, uint64_t n_edges, const uint32_t* src, const uint32_t* dst, uint64_t* in_off,
               uint32_t* in_src);

/* O4: tile algorithm, inner-product form of PPRQuery (ppr_query.cu:73-457, 459-780;
 * matrix_operations.cu:256-316, 318-366 with intended semantics). Symbolic returns the
 * number of structural C tiles (or -1 if cap too small); tile_products receives
 * sum_x sum_{t in A.row(x)} |B.row(t)|. */
int64_t or_tile_symbolic(uint32_t n_strips, const uint32_t* a_off, const uint32_t* a_col, const uint32_t* bt_off,
                         const uint32_t* bt_col, uint32_t* c_off, uint32_t* c_row, uint32_t* c_col, uint64_t cap,
                         uint64_t* tile_products);
void or_tile_numeric(uint32_t n_strips, const uint32_t* a_off, const uint32_t* a_col, const uint64_t* a_data,
                     const uint32_t* bt_off, const uint32_t* bt_col, const uint64_t* bt_data, uint32_t n_c,
                     const uint32_t* c_row, const uint32_t* c_col, uint64_t* c_data, int threads);

/* CPU tile algorithm for one walk hop on strips y = strip_begin + i*stride < strip_end
 * of A^T: x_out strip y = OR over tiles (y,t), set bits (r,t') of x_in[t*64+t'] into
 * row r. Returns the number of set bits of A^T visited. OpenMP over strips. */
uint64_t or_walk_tile_hop(uint32_t n_strips, const uint32_t* at_off, const uint32_t* at_col, const uint64_t* at_data,
                          uint32_t W, const uint64_t* x_in, uint64_t* x_out, uint32_t strip_begin,
                          uint32_t strip_end, uint32_t stride, int threads);
/* Structural walk symbolic: s_out[y] = OR_{t in At.row(y)} s_in[t]; returns tile products. */
uint64_t or_walk_symbolic(uint32_t n_strips, const uint32_t* at_off, const uint32_t* at_col, const uint64_t* s_in,
                          uint64_t* s_out);

/* MLP oracle (n_in -> 64 -> 64 -> n_out, ReLU): FP64 mean-squared-error loss and its
 * gradient for the given weights (layout W1[64][n_in], b1, W2[64][64], b2, W3[n_out][64],
 * b3), and the BF16 round-to-nearest-even conversion defined from the IEEE bit layout.
 * Pinned in tests by central finite differences and by bit-pattern cases. */
void or_mlp_loss_grad(uint32_t n_in, uint32_t n_out, const double* w, const float* x, const float* t, uint32_t n,
                      double* loss, double* grad);
uint16_t or_f32_to_bf16(float f);
float or_bf16_to_f32(uint16_t h);

int or_max_threads(void);

#ifdef __cplusplus
}
#endif

#endif /* MG_ORACLE_H_ */
