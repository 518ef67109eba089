"""Learned cost model for scheduling batched walks.

The work of one hop level of a batch is a list of scheduler chunks (strip ranges of one
panel's A^T tile stream), launched in list order; the launch is balanced best when the
expensive chunks start first (longest-processing-time order). A chunk's cost is dominated
by its gathers (jobs: A^T bits whose source vertex is non-empty at that level), which depend
on how far the panel's queries have spread — unknown until the level runs. An MLP
(n_in=8 -> 64 -> 64 -> 2, BF16 weights, FP32 accumulation, trained on the GPU by the
library's own kernels) predicts, per chunk and level, the jobs and the output-tile occupancy
from features known before the run:

  f0 log1p(tiles)/10        f1 log1p(set bits)/12     f2 log1p(strips)/4
  f3 alive words/64         f4 level/6                f5 log(avg degree)/4
  f6 log2(vertices)/24      f7 set bits per tile/4

Targets: log1p(jobs)/12 and the fraction of the chunk's 64 x strips output tiles that are
numerically non-empty. Training samples are logged from the method's own runs
(mg_walk_batch_set_logging). Predictions only order the launches, so results are
bit-identical whatever the model says. The exact outcomes of the first run decide:
a level whose predicted total jobs are off by more than `fallback_ratio` is re-ordered by the
exact logged costs (fallback), and both the per-chunk misprediction rate and the fallback
frequency are reported.
"""
from __future__ import annotations

import time

import numpy as np

N_FEATURES = 8
N_TARGETS = 2
# Relative chunk cost: jobs dominate; the scan of a tile and the write-out of a strip are
# fixed costs (job equivalents, from the measured rates of the hop kernel).
COST_PER_JOB = 1.0
COST_PER_TILE = 0.35
COST_PER_STRIP = 24.0


def graph_stats(graphs, edge_counts):
    out = []
    for g, m in zip(graphs, edge_counts):
        nv = g.n_strips * 64
        out.append({"n_vertices": nv, "edges": int(m), "avg_degree": max(int(m), 1) / nv})
    return out


def level_features(batch, level: int, gstats, panel_cache=None):
    """(records, X) for hop level `level` of a batch: X is float32 (n_chunks, 8)."""
    rec = batch.level_chunks(level)
    if panel_cache is None:
        panel_cache = {}
    n = len(rec)
    X = np.zeros((n, N_FEATURES), np.float32)
    if n == 0:
        return rec, X
    panels = rec["panel"]
    alive_words = np.zeros(n, np.float32)
    avg_deg = np.zeros(n, np.float32)
    log_nv = np.zeros(n, np.float32)
    for p in np.unique(panels):
        if p not in panel_cache:
            panel_cache[p] = batch.panel_info(int(p))
        info = panel_cache[p]
        sel = panels == p
        a = int(info["alive"][level]) if level < info["hops"] else 0
        alive_words[sel] = bin(a).count("1")
        gs = gstats[info["graph"]]
        avg_deg[sel] = gs["avg_degree"]
        log_nv[sel] = np.log2(gs["n_vertices"])
    tiles = (rec["tile_end"] - rec["tile_begin"]).astype(np.float32)
    bits = rec["bits"].astype(np.float32)
    strips = (rec["strip_end"] - rec["strip_begin"]).astype(np.float32)
    X[:, 0] = np.log1p(tiles) / 10
    X[:, 1] = np.log1p(bits) / 12
    X[:, 2] = np.log1p(strips) / 4
    X[:, 3] = alive_words / 64
    X[:, 4] = level / 6
    X[:, 5] = np.log(np.maximum(avg_deg, 1e-3)) / 4
    X[:, 6] = log_nv / 24
    X[:, 7] = bits / np.maximum(tiles, 1) / 4
    return rec, X


def level_targets(rec):
    strips = (rec["strip_end"] - rec["strip_begin"]).astype(np.float32)
    T = np.zeros((len(rec), N_TARGETS), np.float32)
    T[:, 0] = np.log1p(rec["jobs"].astype(np.float32)) / 12
    T[:, 1] = np.minimum(rec["out_tiles"].astype(np.float32) / np.maximum(strips * 64, 1), 1.0)
    return T


def chunk_cost(rec, jobs):
    tiles = (rec["tile_end"] - rec["tile_begin"]).astype(np.float64)
    strips = (rec["strip_end"] - rec["strip_begin"]).astype(np.float64)
    return COST_PER_JOB * np.asarray(jobs, np.float64) + COST_PER_TILE * tiles + COST_PER_STRIP * strips


def collect(ctx, graphs, gstats, queries_list, chunk_tiles=0, log=None):
    """Runs each query batch once with logging; returns (X, T) over all chunk-levels."""
    import matrixgraph_amd as M

    xs, ts = [], []
    for qs in queries_list:
        b = M.WalkBatch(ctx, graphs, qs, chunk_tiles=chunk_tiles, use_graph=False)
        b.set_logging(True)
        b.run()
        ctx.synchronize()
        cache = {}
        for k in range(b.info()[1]):
            rec, X = level_features(b, k, gstats, cache)
            xs.append(X)
            ts.append(level_targets(rec))
        b.close()
        if log:
            log(f"collect: {sum(len(x) for x in xs)} samples")
    return np.concatenate(xs), np.concatenate(ts)


def train(ctx, X, T, epochs=8, batch=4096, lr=2e-3, seed=1, log=None):
    import matrixgraph_amd as M

    mlp = M.MLP(ctx, N_FEATURES, N_TARGETS, seed=seed)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(X))
    X, T = X[perm], T[perm]
    l0, _ = mlp.loss_grad(X[:65536], T[:65536])
    t0 = time.perf_counter()
    l1 = mlp.train(X, T, epochs=epochs, batch=batch, lr=lr)
    if log:
        log(f"train: {len(X)} samples, {epochs} epochs, loss {l0:.5f} -> {l1:.5f} in {time.perf_counter() - t0:.1f}s")
    return mlp, {"samples": int(len(X)), "epochs": epochs, "loss_initial": float(l0), "loss_final": float(l1)}


def predict(mlp, X):
    """(jobs, occupancy) predictions for feature rows X."""
    if len(X) == 0:
        return np.zeros(0), np.zeros(0)
    Y = mlp.forward(X)
    jobs = np.expm1(np.clip(Y[:, 0], 0, 2.0) * 12)
    occ = np.clip(Y[:, 1], 0, 1)
    return jobs, occ


def order_by_cost(cost, groups=None, buckets_per_octave: int = 2):
    """Longest-first launch order over cost buckets (buckets_per_octave per factor of 2); within a
    bucket the current order is kept. With `groups` (one id per chunk: the chunks of one graph's
    tile range in different panels, which read the same A^T data), whole groups are ordered by
    their summed cost and a group's chunks stay adjacent in their current order, so the panels'
    reads of the same job list and tiles follow each other and share L2."""
    cost = np.asarray(cost, np.float64)
    if groups is None:
        key = np.floor(np.log2(cost + 1.0) * buckets_per_octave)
        return np.argsort(-key, kind="stable").astype(np.uint32)
    n = len(cost)
    _, inv = np.unique(np.asarray(groups), return_inverse=True)
    gcost = np.bincount(inv, weights=cost)
    first = np.full(len(gcost), n, np.int64)
    np.minimum.at(first, inv, np.arange(n))
    gkey = np.floor(np.log2(gcost + 1.0) * buckets_per_octave)
    rank = np.empty(len(gcost), np.int64)
    rank[np.lexsort((first, -gkey))] = np.arange(len(gcost))
    return np.lexsort((np.arange(n), rank[inv])).astype(np.uint32)


def chunk_groups(batch, rec, panel_cache):
    """Group id per chunk record: (graph of its panel, first A^T tile of the chunk)."""
    g = np.zeros(len(rec), np.int64)
    for p in np.unique(rec["panel"]):
        if p not in panel_cache:
            panel_cache[p] = batch.panel_info(int(p))
        g[rec["panel"] == p] = panel_cache[p]["graph"]
    return (g << 33) | rec["tile_begin"].astype(np.int64)


def schedule_learned(ctx, batch, mlp, gstats, fallback_ratio=2.0, mispredict_ratio=2.0):
    """Orders every level by predicted cost, runs once with logging to check the predictions,
    re-orders mispredicted levels by their exact costs, and returns the metrics."""
    levels = batch.info()[1]
    cache = {}
    preds = []
    for k in range(levels):
        rec, X = level_features(batch, k, gstats, cache)
        jobs, occ = predict(mlp, X)
        o = order_by_cost(chunk_cost(rec, jobs), chunk_groups(batch, rec, cache))
        batch.set_level_order(k, o)
        preds.append((jobs[o], occ[o]))  # aligned with the new launch order
    batch.set_logging(True)
    batch.run()
    ctx.synchronize()
    n_chunks = n_bad = n_occ_bad = 0
    fallbacks = 0
    for k in range(levels):
        rec = batch.level_chunks(k)
        jobs_pred, occ_pred = preds[k]
        actual = rec["jobs"].astype(np.float64)
        ratio = (jobs_pred + 1.0) / (actual + 1.0)
        n_chunks += len(rec)
        n_bad += int(np.sum((ratio > mispredict_ratio) | (ratio < 1.0 / mispredict_ratio)))
        strips = (rec["strip_end"] - rec["strip_begin"]).astype(np.float64)
        occ_act = rec["out_tiles"] / np.maximum(strips * 64, 1)
        n_occ_bad += int(np.sum(np.abs(occ_pred - occ_act) > 0.25))
        tot_ratio = (jobs_pred.sum() + 1.0) / (actual.sum() + 1.0)
        if tot_ratio > fallback_ratio or tot_ratio < 1.0 / fallback_ratio:
            batch.set_level_order(k, order_by_cost(chunk_cost(rec, actual), chunk_groups(batch, rec, cache)))
            fallbacks += 1
    batch.set_logging(False)
    return {
        "chunk_levels": n_chunks,
        "jobs_misprediction_rate": n_bad / max(n_chunks, 1),
        "occupancy_misprediction_rate": n_occ_bad / max(n_chunks, 1),
        "levels": levels,
        "fallback_levels": fallbacks,
        "fallback_frequency": fallbacks / max(levels, 1),
        "misprediction_definition": f"predicted/actual jobs outside [1/{mispredict_ratio}, {mispredict_ratio}] "
                                    "(+1 smoothing); occupancy off by more than 0.25",
    }


def schedule_exact(ctx, batch):
    """Orders every level by the exact costs of a logged run (the fallback used for every level)."""
    levels = batch.info()[1]
    cache = {}
    batch.set_logging(True)
    batch.run()
    ctx.synchronize()
    for k in range(levels):
        rec = batch.level_chunks(k)
        batch.set_level_order(k, order_by_cost(chunk_cost(rec, rec["jobs"]), chunk_groups(batch, rec, cache)))
    batch.set_logging(False)


def time_batch(ctx, batch, reps=3):
    batch.run()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.run()
    ctx.synchronize()
    return (time.perf_counter() - t0) / reps


# ----------------------------------------------------------------------------
# Pre-allocation of pruned products from the predicted output-tile occupancy
# ----------------------------------------------------------------------------
# A second use of the same network shape (n_in = ROW_FEATURES -> 64 -> 64 -> 2): per row x
# of a product C = A * B, from the features the symbolic pass gives (mg_multiply_row_features),
# predict the fraction of the row's structural tiles that are numerically non-empty
# (target 0) and log1p(non-empty tiles) / 16 (target 1). The predicted non-empty count sizes
# the output of mg_multiply_pruned, which then needs one numeric pass instead of the exact
# two (flag, then write); a prediction below the true count is caught by the pass's counter
# and the pass runs again at the exact count (the fallback), so the result never changes.

ROW_TARGETS = 2
CAPACITY_MARGIN = 0.25  # predicted count x (1 + margin), rounded up per row


def row_outcomes(c_pruned, structural):
    """Per-row non-empty tiles of a pruned product and the training targets."""
    numeric = np.diff(c_pruned.download_offsets()).astype(np.float64)
    frac = numeric / np.maximum(structural.astype(np.float64), 1.0)
    T = np.zeros((len(numeric), ROW_TARGETS), np.float32)
    T[:, 0] = np.clip(frac, 0.0, 1.0)
    T[:, 1] = np.log1p(numeric) / 16
    return numeric, T


def product_samples(ctx, a, bt):
    """(X, T, tile_products, C) for one product: rows with structural tiles only; C is the
    pruned product (exact mode), returned for chaining walks."""
    import matrixgraph_amd as M

    feat, cnt, tp = M.multiply_row_features(ctx, a, bt)
    c, _ = M.multiply_pruned(ctx, a, bt, M.PRUNED_EXACT)
    _, T = row_outcomes(c, cnt)
    rows = cnt > 0
    return feat[rows], T[rows], tp, c


def predict_capacity(mlp, feat, structural, margin=CAPACITY_MARGIN):
    """Output slots for a pruned product: per row round_up(structural * predicted fraction *
    (1 + margin)), at most the structural count; returns (capacity, per-row predicted counts)."""
    rows = structural > 0
    pred = np.zeros(len(structural), np.float64)
    if rows.any():
        frac = np.clip(mlp.forward(feat[rows])[:, 0].astype(np.float64), 0.0, 1.0)
        pred[rows] = frac * structural[rows]
    per_row = np.minimum(np.ceil(pred * (1.0 + margin)), structural.astype(np.float64))
    return int(per_row.sum()), pred


def collect_products(ctx, graphs, queries, square_graphs=(), log=None, groups=None):
    """Training samples from pruned product walks X_{k+1} = X_k * A of the queries (each hop's
    pruned X feeds the next) and from A * A of `square_graphs` ((A, A^T) pairs). With a list
    `groups`, the product each row came from is appended per row (for margin calibration)."""
    from . import workloads as WL

    xs, ts = [], []
    pairs = 0
    n_products = 0

    def add(X, T, tp):
        nonlocal pairs, n_products
        xs.append(X)
        ts.append(T)
        pairs += tp
        if groups is not None:
            groups.extend([n_products] * len(X))
        n_products += 1

    for q in queries:
        at = graphs[q.graph]
        x = WL.sources_matrix(ctx, at.n_strips * 64, q.sources)
        for _ in range(q.hops):
            X, T, tp, c = product_samples(ctx, x, at)
            add(X, T, tp)
            x.close()
            x = c
        x.close()
    for a, at in square_graphs:
        X, T, tp, c = product_samples(ctx, a, at)
        add(X, T, tp)
        c.close()
    if log:
        log(f"collect products: {sum(len(x) for x in xs)} row samples from {n_products} products covering "
            f"{pairs} tile pairs")
    return np.concatenate(xs), np.concatenate(ts), pairs


def calibrate_margin(mlp, X, T, groups, quantile=0.95, floor=0.05):
    """The capacity margin under which `quantile` of the training products would have fitted:
    per product, predicted non-empty tiles (rows' structural x predicted fraction) against the
    actual ones, both from the logged rows (structural tiles = expm1(16 f0))."""
    groups = np.asarray(groups)
    structural = np.expm1(X[:, 0].astype(np.float64) * 16)
    actual = np.expm1(T[:, 1].astype(np.float64) * 16)
    pred = structural * np.clip(mlp.forward(X)[:, 0].astype(np.float64), 0.0, 1.0)
    n = groups.max() + 1 if len(groups) else 0
    pred_p = np.bincount(groups, weights=pred, minlength=n)
    act_p = np.bincount(groups, weights=actual, minlength=n)
    need = (act_p + 1.0) / (pred_p + 1.0) - 1.0
    return float(max(floor, np.quantile(need, quantile))) if n else floor


def pruned_walk(ctx, at, sources, hops, mode, mlp=None, margin=CAPACITY_MARGIN):
    """Product walk with pruned frontiers; mode 'learned' (capacity from the model),
    'exact' (two numeric passes) or 'structural' (one pass, structural-size scratch).
    Returns (X_hops digest, stats dict)."""
    import time

    import matrixgraph_amd as M
    from . import workloads as WL

    at.prepare_operand()  # A^T's payloads, kept for every hop's product
    x = WL.sources_matrix(ctx, at.n_strips * 64, sources)
    st = {"products": 0, "fallbacks": 0, "passes": 0, "scratch_bytes": 0, "output_bytes": 0, "seconds": 0.0,
          "rows": 0, "rows_mispredicted": 0, "predicted_tiles": 0, "numeric_tiles": 0, "capacity_tiles": 0,
          "structural_tiles": 0, "allocated_tiles": 0}
    for _ in range(hops):
        ctx.synchronize()
        t0 = time.perf_counter()
        if mode == "learned":
            c, ps, pred = M.multiply_pruned_model(ctx, x, at, mlp, margin)
        else:
            pred = None
            c, ps = M.multiply_pruned(ctx, x, at, M.PRUNED_EXACT if mode == "exact" else M.PRUNED_STRUCTURAL)
        ctx.synchronize()
        st["seconds"] += time.perf_counter() - t0
        st["products"] += 1
        st["fallbacks"] += ps["fallback"]
        st["passes"] += ps["passes"]
        st["scratch_bytes"] = max(st["scratch_bytes"], ps["scratch_bytes"])
        st["output_bytes"] = max(st["output_bytes"], ps["output_bytes"])
        st["numeric_tiles"] += ps["numeric_tiles"]
        st["capacity_tiles"] += ps["capacity"]
        st["structural_tiles"] += ps["structural_tiles"]
        # Tile slots allocated for the result: the first pass's slots (plus the retry's), or
        # the exact count.
        st["allocated_tiles"] += (ps["numeric_tiles"] if mode == "exact"
                                  else ps["capacity"] + (ps["numeric_tiles"] if ps["fallback"] else 0))
        if pred is not None:
            actual = np.diff(c.download_offsets()).astype(np.float64)
            rows = (pred > 0) | (actual > 0)
            ratio = (pred[rows] + 1.0) / (actual[rows] + 1.0)
            st["rows"] += int(rows.sum())
            st["rows_mispredicted"] += int(np.sum((ratio > 2.0) | (ratio < 0.5)))
            st["predicted_tiles"] += int(pred.sum())
        x.close()
        x = c
    digest = x.digest()
    x.close()
    return digest, st
