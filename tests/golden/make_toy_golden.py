"""Writes tests/golden/toy_graph.json from the reference's toy graph.

Input: data/toy_graph/test_graph.csv of the reference (25 "src,dst" lines). It is read
from /root/reference when present; the committed JSON keeps the edges, so tests never
need the reference.

Expected outputs are computed here by pure-Python brute force (independent of the C
oracle and of the GPU code):
  * vertex ids compressed by first appearance over the flattened (src, dst) sequence,
    as Edges::ReadFromCSV does by default (core/data_structures/edgelist.cu:200-290);
  * A (self-loops and all edges kept: Edgelist2BitTiledMatrix keeps them,
    core/util/format_converter.cuh:292-415), n_strips = ceil(16/64) = 1;
  * A^k for k = 2..4 over the Boolean semiring (C = A * A with the right operand read
    from the transposed grid, core/task/gpu_task/ppr_query.cu:63-71), as sorted (i, j)
    pairs, plus total_count = popcount (ppr_query.cu:874-956);
  * the single 64x64 tile of A as 64 u64 words (word i = row i, bit j = column j).
"""
from __future__ import annotations

import json
from pathlib import Path

REF_CSV = Path("/root/reference/data/toy_graph/test_graph.csv")
OUT = Path(__file__).resolve().parent / "toy_graph.json"


def read_edges():
    if REF_CSV.exists():
        rows = [ln.strip() for ln in REF_CSV.read_text().splitlines() if ln.strip()]
        return [tuple(int(v) for v in ln.split(",")) for ln in rows]
    return [tuple(e) for e in json.loads(OUT.read_text())["raw_edges"]]


def compress(edges):
    seen = {}
    for u, v in edges:
        for x in (u, v):
            if x not in seen:
                seen[x] = len(seen)
    return seen, [(seen[u], seen[v]) for u, v in edges]


def matmul(a, b, n):
    return [[int(any(a[i][k] and b[k][j] for k in range(n))) for j in range(n)] for i in range(n)]


def main():
    raw = read_edges()
    vid, edges = compress(raw)
    n = len(vid)
    a = [[0] * n for _ in range(n)]
    for u, v in edges:
        a[u][v] = 1
    powers = {1: a}
    for k in range(2, 5):
        powers[k] = matmul(powers[k - 1], a, n)
    tile = [0] * 64
    for u, v in edges:
        tile[u] |= 1 << v
    out = {
        "raw_edges": [list(e) for e in raw],
        "vid_map": {str(k): v for k, v in vid.items()},
        "n_vertices": n,
        "edges": [list(e) for e in edges],
        "tile_words_A": [str(w) for w in tile],
        "powers": {str(k): [[i, j] for i in range(n) for j in range(n) if m[i][j]] for k, m in powers.items()},
        "total_count": {str(k): sum(map(sum, m)) for k, m in powers.items()},
    }
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {OUT}: n={n}, |E|={len(edges)}, counts={out['total_count']}")


if __name__ == "__main__":
    main()
