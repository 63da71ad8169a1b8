"""GPU: prefix-shared matching (plans of one call run as a prefix tree, a
level op shared by several plans runs once per batch for all of them) counts
exactly what one-plan-at-a-time matching counts: graphlet queries of 3..5
vertices (the -km workload), presets and cliques, duplicated plans, small
batches, several view bins, and a captured replay."""
import os
import subprocess
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _plans(vsgm):
    qs = [q for k in (3, 4, 5) for q in vsgm.Query.graphlets(k)]
    qs += [vsgm.Query.preset(k) for k in range(17)] + [vsgm.Query.clique(k) for k in (3, 4, 5)]
    plans = [vsgm.Plan.optimize(q) for q in qs]
    return plans + plans[:3]  # identical plans are counted once and copied


def _run(vsgm, g, plans, share, **opts):
    ex = vsgm.PipelineExecutor(0, g.vertex_count, plans)
    ex.set_option("share_prefixes", int(share))
    for k, v in opts.items():
        ex.set_option(k, v)
    ex.transfer_graph(g)
    ex.match()
    return ex.total_counts, ex.shared_stats()


def test_shared_counts_equal_unshared_and_oracle(vsgm, oracle, gpu):
    g = vsgm.Graph.rmat(9, 16, seed=23)
    plans = _plans(vsgm)
    rp, cl = g.csr()
    want = [oracle.count_plan(rp, cl, p.serialize()) for p in plans]
    print(f"[gpu] prefix sharing, {len(plans)} plans on R-MAT s9", flush=True)
    plain, st0 = _run(vsgm, g, plans, False)
    shared, st1 = _run(vsgm, g, plans, True)
    assert plain == want
    assert shared == want
    assert st0["saved_runs"] == 0 and st1["saved_runs"] > 0
    # many batches per level (the 3- and 4-vertex graphlets and the cliques:
    # every level shared at least once, without thousands of tiny batches)
    few = [vsgm.Plan.optimize(q) for k in (3, 4) for q in vsgm.Query.graphlets(k)]
    few += [vsgm.Plan.optimize(vsgm.Query.clique(k)) for k in (3, 4, 5)]
    few += few[:2]
    small, _ = _run(vsgm, g, few, True, batch_bytes=1 << 16)
    assert small == [oracle.count_plan(rp, cl, p.serialize()) for p in few]


def test_shared_over_bins_and_capture(vsgm, oracle, gpu):
    g = vsgm.Graph.rmat(10, 16, seed=24)
    plans = [vsgm.Plan.optimize(q) for k in (4, 5) for q in vsgm.Query.graphlets(k)]
    hop = max(p.hop for p in plans)
    rd = min(p.root_degree for p in plans)
    rp, cl = g.csr()
    want = [oracle.count_plan(rp, cl, p.serialize()) for p in plans]
    assign = np.random.default_rng(2).integers(0, 3, size=g.vertex_count)
    ex = vsgm.PipelineExecutor(0, g.vertex_count, plans)
    ex.set_option("share_prefixes", 1)
    ex.set_graph(g)
    for b in range(3):
        ex.add_device_view(np.nonzero(assign == b)[0].astype(np.uint32), hop, rd)
    print("[gpu] prefix sharing over 3 device-built bins, then captured", flush=True)
    ex.match_all()
    assert ex.total_counts == want
    ex.reset()
    ex.capture()
    assert ex.total_counts == want  # the recorded run
    for _ in range(2):
        ex.reset()
        ex.replay()
        assert ex.total_counts == want


def test_pm_pipeline_graphlets_shared(vsgm, oracle, gpu, tmp_path):
    """pm_pipeline -km 4 with -sp 1 (prefix tree) prints the same per-query
    totals as with -sp 0, equal to the oracle's."""
    g = vsgm.Graph.rmat(10, 16, seed=25)
    path = str(tmp_path / "rmat10.bin")
    g.write_bin(path)
    plans = [vsgm.Plan.optimize(q) for q in vsgm.Query.graphlets(4)]
    rp, cl = g.csr()
    want = [oracle.count_plan(rp, cl, p.serialize()) for p in plans]
    hop = max(p.hop for p in plans)
    vsgm.ViewBinManager.write_vbmap(path + f".{hop}.hop.vbmap", 2, np.random.default_rng(1).integers(0, 2, size=g.vertex_count))
    exe = os.path.join(ROOT, "vsgm_amd", "bin", "pm_pipeline")
    for sp in ("0", "1"):
        out = subprocess.run([exe, "-f", path, "-m", "0.1", "-km", "4", "-sp", sp], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr[-2000:]
        got = [int(x.split(":")[1]) for x in out.stdout.splitlines() if x.startswith("total count for query ")]
        assert got == want, (sp, got, want)


def test_shared_time_on_five_vertex_graphlets(vsgm, gpu):
    """Printed: the 21 five-vertex graphlet plans on R-MAT s12, one at a time
    and as a prefix tree (same counts)."""
    g = vsgm.Graph.rmat(12, 16, seed=26)
    plans = [vsgm.Plan.optimize(q) for q in vsgm.Query.graphlets(5)]
    t = time.perf_counter()
    plain, _ = _run(vsgm, g, plans, False)
    t_plain = time.perf_counter() - t
    t = time.perf_counter()
    shared, st = _run(vsgm, g, plans, True)
    t_shared = time.perf_counter() - t
    print(f"[gpu] 21 five-vertex graphlets, R-MAT s12: one at a time {t_plain:.2f}s, prefix tree {t_shared:.2f}s "
          f"({st['saved_runs']} level runs saved)", flush=True)
    assert shared == plain
