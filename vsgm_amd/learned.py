"""Learned planner: an MLP cost model that picks a pattern's matching order.

The heuristic planner (Plan.optimize, a restatement of src/query/plan.h:61-580)
stays the fallback and the reference for counts: every order gives the same
count, so a learned plan can only change the time.

Records. For a (pattern, view, order) the plain enumerator runs the plan of
that order (Plan.from_order) over the view and its device time is logged next
to the heuristic order's time on the same view. The label is
log(time(order) / time(heuristic order)); the model learns relative cost, so
patterns and views of very different sizes share one scale.

Features (IN_DIM = 102, all in [-1, 1] or small):
  28  the pattern's adjacency permuted into the matching order (upper
      triangle of the 8 x 8 matrix, zero padded)
  28  the symmetry-breaking conditions permuted the same way: +1 where
      position i must take a smaller id than position j, -1 for larger, 0 for
      none (orders that are images of each other under an automorphism
      differ only here, and their costs differ)
   8  backward-neighbour count of each position / 7
  32  the view's degree histogram: fraction of view vertices with degree in
      [2^b, 2^(b+1)) for b = 0..31 (isolated vertices excluded)
   4  log2 of vertex count, list entries, sources and maximum degree, / 40
   2  pattern size / 8, plan hop / 4

Choice. All connected orders (up to `max_orders`, the heuristic's always
included) are scored on the GPU in one batch; the lowest predicted relative
cost wins only if it predicts at least `margin` (log units) of gain over the
heuristic order, otherwise the heuristic plan is kept. `calibrate_margin`
sets the margin from record groups held out of training: the smallest
candidate margin under which no held-out decision the model would take is
slower than the heuristic order (the realised gain of the decisions taken
breaks ties toward the larger one).
"""
import itertools
import math
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import api as V
from .mlp import CostModel

IN_DIM = 102
MAXQ = 8
_UPPER = np.triu_indices(MAXQ, 1)
LESS_THAN, LARGER_THAN = 0, 1  # condition codes of Query.info()


def connected_orders(adj: np.ndarray, limit: Optional[int] = None, seed: int = 0) -> List[Tuple[int, ...]]:
    """Matching orders in which every vertex after the first has an earlier
    neighbour (the orders Plan.from_order accepts)."""
    n = adj.shape[0]
    out = [p for p in itertools.permutations(range(n)) if all(any(adj[p[i], p[j]] for j in range(i)) for i in range(1, n))]
    if limit is not None and len(out) > limit:
        rng = np.random.default_rng(seed)
        out = [out[i] for i in sorted(rng.choice(len(out), size=limit, replace=False))]
    return out


def view_stats(row_ptrs: np.ndarray, n_sources: int) -> np.ndarray:
    """The 36 view features: degree histogram and log sizes."""
    deg = np.diff(row_ptrs.astype(np.int64))
    nz = deg[deg > 0]
    hist = np.zeros(32, np.float64)
    if nz.size:
        b = np.floor(np.log2(nz)).astype(np.int64)
        hist = np.bincount(np.clip(b, 0, 31), minlength=32).astype(np.float64) / nz.size
    sizes = np.array([math.log2(len(deg) + 1), math.log2(int(row_ptrs[-1]) + 1), math.log2(n_sources + 1),
                      math.log2(int(deg.max()) + 1 if deg.size else 1)]) / 40.0
    return np.concatenate([hist, sizes]).astype(np.float32)


def featurize(adj: np.ndarray, cond: np.ndarray, order: Sequence[int], hop: int, vstats: np.ndarray) -> np.ndarray:
    n = adj.shape[0]
    o = list(order)
    perm = np.zeros((MAXQ, MAXQ), np.float32)
    perm[:n, :n] = adj[np.ix_(o, o)]
    pc = np.zeros((MAXQ, MAXQ), np.float32)
    c = cond[np.ix_(o, o)]
    pc[:n, :n] = np.where(c == LESS_THAN, 1.0, np.where(c == LARGER_THAN, -1.0, 0.0))
    back = np.zeros(MAXQ, np.float32)
    for i in range(1, n):
        back[i] = perm[i, :i].sum() / 7.0
    return np.concatenate([perm[_UPPER], pc[_UPPER], back, vstats, [n / 8.0, hop / 4.0]]).astype(np.float32)


def _time_plans(ex, repeats: int) -> List[float]:
    best = None
    for _ in range(repeats):
        ex.reset()
        ex.match()
        ms = ex.match_ms
        best = ms if best is None else [min(a, b) for a, b in zip(best, ms)]
    return best


def collect(graphs, queries: Sequence["V.Query"], bins: int = 2, max_orders: int = 24, repeats: int = 2, seed: int = 0,
            log=None) -> Dict[str, np.ndarray]:
    """Training records from the plain enumerator. `graphs` are Graph objects;
    each is split into `bins` view bins (seeded uniform assignment) and every
    query runs in every sampled order over every bin. Returns x, y (log time
    ratio to the heuristic order), the raw times and group ids."""
    xs, ys, ms_all, groups = [], [], [], []
    gid = 0
    rng = np.random.default_rng(seed)
    for gi, g in enumerate(graphs):
        assign = rng.integers(0, bins, size=g.vertex_count).astype(np.int32)
        for q in queries:
            n, _, adj, cond = q.info()
            heur = V.Plan.optimize(q)
            orders = connected_orders(adj, max_orders, seed=seed + gid)
            if tuple(heur.order) not in orders:
                orders.append(tuple(heur.order))
            plans = [V.Plan.from_order(q, o) for o in orders]
            hop = max(p.hop for p in plans)
            rd = min(p.root_degree for p in plans)
            vbm = V.ViewBinManager.from_assignment(g, hop, bins, assign)
            holder = V.ViewBinHolder(g.vertex_count, g.entries * 4 + 4)
            ex = V.PipelineExecutor(0, g.vertex_count, plans, max_sources=vbm.max_sources)
            totals = np.zeros(len(plans), np.uint64)
            for b in range(vbm.n_bins):
                vbm.materialize(b, holder, rd)
                rp, _, src = holder.csr(g.vertex_count)
                ex.transfer(holder)
                ms = _time_plans(ex, repeats)
                # per bin the count depends on which pattern vertex is the
                # root (bins partition the roots); summed over bins it is the
                # pattern's count for every order
                totals += np.array(ex.total_counts, np.uint64)
                vs = view_stats(rp, len(src))
                h = orders.index(tuple(heur.order))
                for o, p, t in zip(orders, plans, ms):
                    xs.append(featurize(adj, cond, o, p.hop, vs))
                    ys.append(math.log((t + 1e-3) / (ms[h] + 1e-3)))
                    ms_all.append(t)
                    groups.append(gid)
                gid += 1
            if len(set(totals.tolist())) != 1:
                raise RuntimeError(f"orders disagree on the count: {totals.tolist()}")
            if log:
                log(f"graph {gi} query n={n}: {len(orders)} orders x {vbm.n_bins} bins")
    return {"x": np.array(xs, np.float32), "y": np.array(ys, np.float32), "ms": np.array(ms_all, np.float64),
            "group": np.array(groups, np.int64)}


def train(data: Dict[str, np.ndarray], epochs: int = 60, batch: int = 128, lr: float = 2e-3, seed: int = 0, ctx=None):
    """Fit a CostModel on collected records; returns (model, per-epoch loss)."""
    m = CostModel(IN_DIM, seed=seed, ctx=ctx)
    losses = m.train(data["x"], data["y"], epochs=epochs, batch=batch, lr=lr, seed=seed)
    return m, losses


def split_groups(data: Dict[str, np.ndarray], holdout: float = 0.2, seed: int = 0):
    """Records split by group (one pattern over one view): (train, held out)."""
    groups = np.unique(data["group"])
    rng = np.random.default_rng(seed)
    held = set(rng.choice(groups, size=max(1, int(round(len(groups) * holdout))), replace=False).tolist())
    mask = np.array([g in held for g in data["group"]])
    return {k: v[~mask] for k, v in data.items()}, {k: v[mask] for k, v in data.items()}


MARGINS = (0.0, 0.05, 0.1, 0.2, 0.3, 0.5, 0.75, 1.0, 1.5, 2.0, math.inf)


def decisions(pred: np.ndarray, data: Dict[str, np.ndarray]):
    """Per group: (predicted gain of its best-predicted order over the
    heuristic order, realised gain of that order), both in log units. The
    heuristic order is the record whose label is 0 (its own time ratio)."""
    out = []
    for g in np.unique(data["group"]):
        idx = np.nonzero(data["group"] == g)[0]
        h = idx[np.argmin(np.abs(data["y"][idx]))]
        b = idx[np.argmin(pred[idx])]
        out.append((float(pred[h] - pred[b]), float(-data["y"][b])))
    return out


def calibrate_margin(model: CostModel, held: Dict[str, np.ndarray], margins: Sequence[float] = MARGINS):
    """The margin taken from held-out groups: the smallest one under which no
    decision taken (predicted gain above it) is slower than the heuristic
    order, preferring more realised gain; infinity (always the heuristic
    plan) when every margin lets a regression through. Returns (margin,
    per-margin table of decisions taken, regressions and realised gain)."""
    dec = decisions(model.forward(held["x"]), held)
    table = []
    for m in margins:
        taken = [(p, r) for p, r in dec if p > m]
        table.append({"margin": m, "taken": len(taken), "regressions": sum(r < 0 for _, r in taken),
                      "realised_gain": sum(r for _, r in taken)})
    safe = [t for t in table if t["regressions"] == 0]
    best = max(safe, key=lambda t: (t["realised_gain"], -t["margin"]))
    return best["margin"], table


def sample_roots(vertex_count: int, seed: int = 0, fraction: int = 64, minimum: int = 1024) -> np.ndarray:
    """A seeded uniform sample of root vertices (1/fraction of them, at
    least `minimum`), sorted."""
    n = min(vertex_count, max(minimum, vertex_count // fraction))
    return np.sort(np.random.default_rng(seed).permutation(vertex_count)[:n]).astype(np.uint32)


def check_on_roots(graph: "V.Graph", learned_plan: "V.Plan", heuristic_plan: "V.Plan", roots: np.ndarray,
                   options: Optional[Dict[str, int]] = None, min_gain: float = 0.05):
    """Times both plans over the same roots of the workload graph (the whole
    graph as the view, the roots as sources) and keeps the learned plan only
    when it is at least `min_gain` faster there. Each plan's time over
    uniformly drawn roots estimates its whole time; its count over them is
    not comparable between orders (each order roots an embedding at a
    different pattern vertex), only whole-graph totals are. Returns (keep,
    learned ms, heuristic ms)."""
    ms = []
    for plan in (learned_plan, heuristic_plan):
        ex = V.PipelineExecutor(0, graph.vertex_count, [plan])
        for k, v in (options or {}).items():
            ex.set_option(k, v)
        ex.transfer_graph(graph, roots)
        ex.match()
        ms.append(ex.match_ms[0])
        del ex
    return ms[0] < (1.0 - min_gain) * ms[1], ms[0], ms[1]


class LearnedPlanner:
    """Picks matching orders with a trained CostModel; falls back to the
    heuristic plan when no order is predicted to beat it by `margin`."""

    def __init__(self, model: CostModel, max_orders: int = 120, margin: float = 0.1):
        self.model, self.max_orders, self.margin = model, max_orders, margin

    def plan(self, query: "V.Query", vstats: np.ndarray):
        n, _, adj, cond = query.info()
        heur = V.Plan.optimize(query)
        orders = connected_orders(adj, self.max_orders)
        h = tuple(heur.order)
        if h not in orders:
            orders.append(h)
        hops = [heur.hop if o == h else V.Plan.from_order(query, o).hop for o in orders]
        x = np.stack([featurize(adj, cond, o, hp, vstats) for o, hp in zip(orders, hops)])
        pred = self.model.forward(x)
        best = int(np.argmin(pred))
        hi = orders.index(h)
        if orders[best] == h or pred[best] > pred[hi] - self.margin:
            return heur, {"learned": False, "pred_gain": float(pred[hi] - pred[best]), "best_order": list(orders[best])}
        return V.Plan.from_order(query, orders[best]), {"learned": True, "pred_gain": float(pred[hi] - pred[best]),
                                                        "order": list(orders[best])}
