"""GPU: the learned planner end to end — records from the plain enumerator,
training with the HIP kernels, and plans whose counts equal the heuristic
plan's. Prints what the trained network gains over the heuristic order on
held-out (pattern, view) groups, measured, and checks that its ranking of
orders beats a model without the network (all orders equal)."""
import math

import numpy as np
import pytest

from vsgm_amd.shapes import all_shapes

pytestmark = pytest.mark.gpu


def _spearman(a, b):
    ra = np.argsort(np.argsort(a)).astype(np.float64)
    rb = np.argsort(np.argsort(b)).astype(np.float64)
    ra -= ra.mean()
    rb -= rb.mean()
    den = math.sqrt((ra * ra).sum() * (rb * rb).sum())
    return float((ra * rb).sum() / den) if den else 0.0


def test_learned_planner(vsgm, oracle, gpu):
    from vsgm_amd import learned as L

    shapes = [s for s in all_shapes((3, 4, 5))]
    rng = np.random.default_rng(0)
    pick = sorted(rng.choice(len(shapes), size=16, replace=False))
    queries = [vsgm.Query.from_matrix(shapes[i][2]) for i in pick] + [vsgm.Query.preset(4), vsgm.Query.preset(1)]
    graphs = [vsgm.Graph.rmat(10, 16, seed=s) for s in (1, 2)]
    print(f"[gpu] collecting records: {len(queries)} patterns x {len(graphs)} graphs", flush=True)
    data = L.collect(graphs, queries, bins=2, max_orders=10, repeats=1, seed=1)
    n = len(data["y"])
    assert n > 300 and data["x"].shape == (n, L.IN_DIM)
    groups = np.unique(data["group"])
    test_groups = set(groups[::4].tolist())
    tr = np.array([g not in test_groups for g in data["group"]])
    train = {k: v[tr] for k, v in data.items()}
    model, losses = L.train(train, epochs=80, batch=64, lr=2e-3, seed=0)
    print(f"[gpu] trained on {tr.sum()} records: loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert losses[-1] < 0.7 * losses[0]

    # held-out ranking quality vs the network-free baseline (constant cost: rank correlation 0)
    rhos = []
    gains = []
    for gid in sorted(test_groups):
        sel = data["group"] == gid
        if sel.sum() < 3 or np.ptp(data["ms"][sel]) == 0:
            continue
        pred = model.forward(data["x"][sel])
        rhos.append(_spearman(pred, data["y"][sel]))
        ms = data["ms"][sel]
        heur = ms[int(np.argmin(np.abs(data["y"][sel])))]  # the heuristic order's record has label 0
        gains.append(heur / ms[int(np.argmin(pred))])
    rho = float(np.mean(rhos))
    geo = float(np.exp(np.mean(np.log(gains))))
    print(f"[gpu] held-out: mean Spearman {rho:.3f} over {len(rhos)} groups; chosen-order speedup over the "
          f"heuristic order (geometric mean, measured) {geo:.3f}", flush=True)
    assert rho > 0.2

    # learned plans count exactly what the heuristic plans count
    planner = L.LearnedPlanner(model, margin=0.05)
    g = vsgm.Graph.rmat(9, 16, seed=3)
    rp, cl = g.csr()
    vs = L.view_stats(rp, g.vertex_count)
    learned, heur = [], []
    for q in queries[:10] + [vsgm.Query.clique(4), vsgm.Query.preset(2)]:
        p, info = planner.plan(q, vs)
        learned.append(p)
        heur.append(vsgm.Plan.optimize(q))
    ex = vsgm.PipelineExecutor(0, g.vertex_count, learned + heur)
    ex.transfer_graph(g)
    ex.match()
    c = ex.total_counts
    k = len(learned)
    assert c[:k] == c[k:]
    print("[gpu] learned plans match the heuristic counts; oracle check", flush=True)
    assert c[k:] == [oracle.count_plan(rp, cl, p.serialize()) for p in heur]


def test_root_sample_check(vsgm, gpu):
    """The root-sample check runs both plans over the same sampled roots and
    keeps the learned plan only when it beats the heuristic one by the given
    margin: with a margin no plan can meet (99 %) it refuses, with one every
    plan meets (-1000 %) it keeps, and both plans count the same over the
    whole graph. Prints the two times of a 4-clique order against the
    heuristic order on a degree-ordered R-MAT graph."""
    from vsgm_amd import learned as L

    g, _ = vsgm.Graph.rmat(14, 16, seed=31).relabel_by_degree()
    q = vsgm.Query.clique(4)
    heur, other = vsgm.Plan.optimize(q), vsgm.Plan.from_order(q, [2, 3, 1, 0])
    roots = L.sample_roots(g.vertex_count, seed=5, fraction=4)
    assert len(roots) == g.vertex_count // 4 and np.all(np.diff(roots.astype(np.int64)) > 0)
    print(f"[gpu] root-sample check, R-MAT s14 degree ids, 4-clique, {len(roots)} roots", flush=True)
    keep, ms_o, ms_h = L.check_on_roots(g, other, heur, roots, min_gain=0.99)
    assert not keep and ms_o > 0 and ms_h > 0
    keep, _, _ = L.check_on_roots(g, other, heur, roots, min_gain=-10.0)
    assert keep
    print(f"[gpu] order (2,3,1,0) {ms_o:.3f} ms, heuristic (0,1,2,3) {ms_h:.3f} ms", flush=True)
    totals = []
    for plan in (heur, other):
        ex = vsgm.PipelineExecutor(0, g.vertex_count, [plan])
        ex.transfer_graph(g)
        ex.match()
        totals.append(ex.total_counts[0])
    assert totals[0] == totals[1]
