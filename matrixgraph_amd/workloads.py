"""Synthetic workloads of the walk benchmarks (graph and query specifications).

Graph sets (R-MAT with (a,b,c,d) = (0.57,0.19,0.19,0.05), edge factor 16, duplicates and
self-loops removed; Erdos-Renyi G(n, p) with n = 2^18, p = 32/n = 2^-13 exactly):

  batch graphs, seeds s0..s0+7: R-MAT scales 16, 17, 18, 19, 20 (seeds s0..s0+4) and three
  ER graphs (seeds s0+5..s0+7).

Queries: each picks a graph uniformly, a source-set size round(exp(U[0, ln 4096])) clipped
to [1, 4096] (log-uniform), a hop count uniform in {1..6}, and distinct sources drawn by
mg_sample_sources from a per-query seed. All draws come from numpy's PCG64 seeded with the
query seed, so a (graph seed, query seed) pair fixes the workload.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

RMAT_ABC = (0.57, 0.19, 0.19)


@dataclass
class GraphSpec:
    kind: str  # "rmat" | "er"
    scale: int
    seed: int

    @property
    def n_vertices(self) -> int:
        return 1 << self.scale

    def describe(self) -> str:
        if self.kind == "rmat":
            return f"R-MAT s{self.scale} ef16 seed {self.seed}"
        return f"ER n=2^{self.scale} p=2^-13 seed {self.seed}"


@dataclass
class QuerySpec:
    graph: int
    sources: np.ndarray
    hops: int


def batch_graph_specs(seed0: int = 10):
    specs = [GraphSpec("rmat", s, seed0 + i) for i, s in enumerate(range(16, 21))]
    specs += [GraphSpec("er", 18, seed0 + 5 + i) for i in range(3)]
    return specs


def build_graph(ctx, spec: GraphSpec):
    """Device A^T (right operand of the walk) for a graph spec, plus the edge count."""
    import matrixgraph_amd as M

    if spec.kind == "rmat":
        el = M.EdgeList.rmat(ctx, spec.scale, 16, *RMAT_ABC, spec.seed)
    else:
        el = M.EdgeList.er(ctx, spec.scale, 1, spec.scale - 5, spec.seed)  # p = 32 / 2^scale
    _, m = el.info()
    at = el.to_tiled_matrix(transpose=True)
    el.close()
    return at, m


def oracle_edges(spec: GraphSpec):
    """The same graph from the CPU oracle (test use only)."""
    import oracle as O

    if spec.kind == "rmat":
        return O.rmat_edges(spec.scale, 16, *RMAT_ABC, spec.seed)
    return O.er_edges(spec.scale, 1, spec.scale - 5, spec.seed)


def batch_queries(graph_specs, n_queries: int = 512, seed: int = 1, max_sources: int = 4096, max_hops: int = 6):
    import matrixgraph_amd as M

    rng = np.random.default_rng(seed)
    out = []
    for q in range(n_queries):
        g = int(rng.integers(0, len(graph_specs)))
        size = int(np.clip(np.rint(np.exp(rng.uniform(0.0, np.log(max_sources)))), 1, max_sources))
        hops = int(rng.integers(1, max_hops + 1))
        qseed = int(rng.integers(0, 2**63 - 1))
        src = M.sample_sources(graph_specs[g].n_vertices, size, qseed)
        out.append(QuerySpec(g, src, hops))
    return out


# ----------------------------------------------------------------------------
# Density-mixed union graph and walks by tile products
# ----------------------------------------------------------------------------
# Planted communities: community c is vertices [c*stride, c*stride + size), tile-aligned and
# spread over the vertex range; intra-community edge probability p = 1288490189 / 2^32
# (0.3 to 10 digits; the generator draws exact dyadic Bernoulli bits).
COMMUNITY_P_NUM = 1288490189
COMMUNITY_P_LOG2 = 32


def union_graph(ctx, scale: int, seed: int, n_comm: int, comm_size: int, stride: int):
    """R-MAT (edge factor 16) union planted dense communities; returns the edge list."""
    import matrixgraph_amd as M

    rm = M.EdgeList.rmat(ctx, scale, 16, *RMAT_ABC, seed)
    cm = M.EdgeList.communities(ctx, 1 << scale, n_comm, comm_size, stride, COMMUNITY_P_NUM, COMMUNITY_P_LOG2, seed)
    u = rm.union(cm)
    rm.close()
    cm.close()
    return u


def union_graph_oracle(scale: int, seed: int, n_comm: int, comm_size: int, stride: int):
    """The same union graph from the CPU oracle (test use only)."""
    import oracle as O

    s1, d1 = O.rmat_edges(scale, 16, *RMAT_ABC, seed)
    s2, d2 = O.community_edges(1 << scale, n_comm, comm_size, stride, COMMUNITY_P_NUM, COMMUNITY_P_LOG2, seed)
    key = np.unique(np.concatenate([s1.astype(np.uint64) << np.uint64(32) | d1, s2.astype(np.uint64) << np.uint64(32) | d2]))
    return (key >> np.uint64(32)).astype(np.uint32), (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def sources_matrix(ctx, n_vertices: int, sources):
    """X_0 as a bit-tiled matrix in the reference layout: row s (0..len-1) has the single bit
    at column sources[s] (padded-square n x n grid, tile rows 0..ceil(len/64)-1 used)."""
    import matrixgraph_amd as M

    src = np.ascontiguousarray(sources, np.uint32)
    el = M.EdgeList.upload(ctx, n_vertices, np.arange(len(src), dtype=np.uint32), src)
    x = el.to_tiled_matrix(transpose=False)
    el.close()
    return x


def product_walk(ctx, at, n_vertices: int, sources, hops: int, min_bits):
    """X_{k+1} = X_k * A by tile products (symbolic pass, then the numeric pass with
    matrix-core products for tile pairs whose tiles both have >= min_bits bits; min_bits None:
    mg_multiply_numeric, compact tiles multiplied from their entry lists).
    Returns the list X_0..X_hops and per-hop stats (tile products, matrix-core products,
    symbolic and numeric milliseconds)."""
    import time

    import matrixgraph_amd as M

    at.prepare_operand()  # A^T's payloads, kept for every hop's product
    xs = [sources_matrix(ctx, n_vertices, sources)]
    stats = []
    for _ in range(hops):
        ctx.synchronize()
        t0 = time.perf_counter()
        c, st = M.multiply_symbolic(ctx, xs[-1], at)
        ctx.synchronize()
        t1 = time.perf_counter()
        if min_bits is None:
            M.multiply_numeric(ctx, xs[-1], at, c)
            ctx.synchronize()
            dense = 0
        else:
            dense = M.multiply_numeric_dense(ctx, xs[-1], at, c, min_bits)
        t2 = time.perf_counter()
        stats.append(dict(st, dense_products=dense, symbolic_ms=(t1 - t0) * 1e3, numeric_ms=(t2 - t1) * 1e3))
        xs.append(c)
    return xs, stats
