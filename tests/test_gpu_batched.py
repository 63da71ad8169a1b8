"""GPU: batched mode — several view bins resident at once, every plan over
every bin, per-plan launch tuning, and the recorded run captured into one
device graph and replayed. Counts must equal the one-at-a-time runs and the
oracle on every replay."""
import numpy as np
import pytest

from vsgm_amd.shapes import all_shapes

pytestmark = pytest.mark.gpu


def _bins(vsgm, g, plans, k, seed):
    hop = max(p.hop for p in plans)
    rd = min(p.root_degree for p in plans)
    assign = np.random.default_rng(seed).integers(0, k, size=g.vertex_count).astype(np.int32)
    vbm = vsgm.ViewBinManager.from_assignment(g, hop, k, assign)
    holders = []
    for b in range(vbm.n_bins):
        h = vsgm.ViewBinHolder(g.vertex_count, g.entries * 4 + 4)
        vbm.materialize(b, h, rd)
        holders.append(h)
    return vbm, holders


def _oracle(oracle, g, plans):
    rp, cl = g.csr()
    return [oracle.count_plan(rp, cl, p.serialize()) for p in plans]


@pytest.fixture(scope="module")
def setup(vsgm, oracle, gpu):
    g = vsgm.Graph.rmat(8, 8, seed=31)
    rng = np.random.default_rng(5)
    shapes = all_shapes()
    pick = rng.choice(len(shapes), size=24, replace=False)
    qs = [vsgm.Query.from_matrix(shapes[i][2]) for i in sorted(pick)] + [vsgm.Query.clique(4), vsgm.Query.preset(1)]
    plans = [vsgm.Plan.optimize(q) for q in qs]
    print("[gpu] oracle counts for the batched tests", flush=True)
    want = _oracle(oracle, g, plans)
    return g, plans, want


def test_resident_views_match_all(vsgm, setup):
    g, plans, want = setup
    vbm, holders = _bins(vsgm, g, plans, 6, 1)
    ex = vsgm.PipelineExecutor(0, g.vertex_count, plans, max_sources=vbm.max_sources)
    for h in holders:
        ex.add_view(h)
    print(f"[gpu] match_all over {len(holders)} resident views, {len(plans)} plans", flush=True)
    ex.match_all()
    assert ex.total_counts == want
    # a subset of plans, and one bin at a time through transfer, agree
    ex.reset()
    ex.match_all([0, 3, len(plans) - 1])
    got = ex.total_counts
    assert [got[0], got[3], got[-1]] == [want[0], want[3], want[-1]] and got[1] == 0
    one = vsgm.PipelineExecutor(0, g.vertex_count, plans, max_sources=vbm.max_sources)
    for h in holders:
        one.transfer(h)
        one.match()
    assert one.total_counts == want


def test_plan_tuning_keeps_counts(vsgm, setup):
    g, plans, want = setup
    ex = vsgm.PipelineExecutor(0, g.vertex_count, plans)
    for i in range(len(plans)):
        ex.set_plan_tuning(i, isect_kernel=i % 3, max_grid=(256, 2048, 8192)[i % 3], batch_bytes=(0, 1 << 16)[i % 2])
    assert ex.plan_tuning(4) == {"isect_kernel": 1, "max_grid": 2048, "batch_bytes": 0, "cold_weight": 16, "ticket_chunks": 64}
    ex.transfer_graph(g)
    print("[gpu] per-plan tuning", flush=True)
    ex.match()
    assert ex.total_counts == want


def test_capture_and_replay(vsgm, setup):
    g, plans, want = setup
    vbm, holders = _bins(vsgm, g, plans, 4, 2)
    ex = vsgm.PipelineExecutor(0, g.vertex_count, plans, max_sources=vbm.max_sources)
    for h in holders:
        ex.add_view(h)
    ex.set_plan_tuning(1, isect_kernel=2, max_grid=1024, batch_bytes=1 << 17)  # several batches per level
    print("[gpu] capture", flush=True)
    nodes = ex.capture()
    assert nodes > len(plans) * len(holders)
    assert ex.total_counts == want  # the recorded run
    for rep in range(3):
        ex.reset()
        ms = ex.replay()
        print(f"[gpu] replay {rep}: {ms:.2f} ms", flush=True)
        assert ms > 0
        assert ex.total_counts == want, rep
        # eager work between replays reuses the pool the graph was captured with
        ex.reset()
        ex.match_all([0, 2])
        c = ex.total_counts
        assert c[0] == want[0] and c[2] == want[2]
    # a capture of a subset of plans
    ex.reset()
    ex.capture([5, 6])
    ex.reset()
    ex.replay()
    c = ex.total_counts
    assert c[5] == want[5] and c[6] == want[6] and sum(c) == want[5] + want[6]


def test_autotuner_installs_a_config_and_keeps_counts(vsgm, setup):
    from vsgm_amd import autotune

    g, plans, want = setup
    ex = vsgm.PipelineExecutor(0, g.vertex_count, plans)
    ex.transfer_graph(g)
    cands = {"isect_kernel": [4, 0, 2], "max_grid": [1024, 8192], "batch_bytes": [0, 1 << 16], "ticket_chunks": [16, 64]}
    print("[gpu] autotune 3 plans", flush=True)
    res = autotune.tune(ex, plans=[0, len(plans) - 2, len(plans) - 1], repeats=1, candidates=cands)
    for p, r in res.items():
        assert r["trials"] == 6  # the default plus every other value of each knob
        assert r["ms"] <= r["default_ms"]
        assert ex.plan_tuning(p) == r["config"]
        assert all(r["config"][k] in v for k, v in cands.items())
    ex.match()
    assert ex.total_counts == want
