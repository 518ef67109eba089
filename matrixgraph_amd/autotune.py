"""Autotuner for batched walks.

Searches the launch configuration of mg_walk_batch by timing the real workload on the
device: the scheduler chunk size (A^T tiles per workgroup), the hop kernel of the levels
after the first (the tile-decoding sweep, or the job-list kernel with 1, 2 or 4 row segments,
mg_walk_batch_set_hop_kernel) and the query packing. Each candidate is built, run once to
capture its HIP graph, then timed over `reps` replays; the fastest wins.
Results are cached per workload signature (graph sizes, query sizes and hop counts) in
memory, and optionally in a JSON file.
"""
from __future__ import annotations

import hashlib
import json
import time
from pathlib import Path

DEFAULT_CANDIDATES = [
    {"packing": 0, "chunk_tiles": ct, "segments": seg}
    for ct in (512, 1024, 2048, 4096) for seg in (0, 1, 2, 4)
]
# MG_PACK_ONE_PER_PANEL is not a default candidate: every query then holds 4096-source
# frontiers for each of its levels, which does not fit for hundreds of queries on large graphs.

_CACHE: dict = {}


def signature(graphs, queries) -> str:
    h = hashlib.sha256()
    for g in graphs:
        h.update(f"{g.n_strips}:{g.n_nz_tile};".encode())
    for q in queries:
        h.update(f"{q.graph}:{len(q.sources)}:{q.hops};".encode())
    return h.hexdigest()[:16]


def tune(ctx, graphs, queries, candidates=None, reps=3, cache_file: Path | None = None, log=None):
    """Returns (best_options, table) where table lists {options, ms} per candidate."""
    import matrixgraph_amd as M

    sig = signature(graphs, queries)
    if cache_file and cache_file.exists():
        disk = json.loads(cache_file.read_text())
        _CACHE.update(disk)
    if sig in _CACHE:
        return _CACHE[sig]["best"], _CACHE[sig]["table"]
    table = []
    for opt in candidates or DEFAULT_CANDIDATES:
        b = M.WalkBatch(ctx, graphs, queries, packing=opt["packing"], chunk_tiles=opt["chunk_tiles"], use_graph=True)
        b.set_hop_kernel(opt.get("segments", 0))
        b.run()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            b.run()
        ctx.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / reps
        panels = b.info()[0]
        b.close()
        table.append({"options": opt, "ms": round(ms, 3), "panels": panels})
        if log:
            log(f"autotune {opt}: {ms:.2f} ms ({panels} panels)")
    best = min(table, key=lambda r: r["ms"])["options"]
    _CACHE[sig] = {"best": best, "table": table}
    if cache_file:
        cache_file.write_text(json.dumps(_CACHE, indent=1))
    return best, table
