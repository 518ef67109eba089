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
