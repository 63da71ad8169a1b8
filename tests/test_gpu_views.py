"""GPU: views built on the device from the resident graph are identical to
the host's ViewBin::Materialize views (same CSR, byte for byte) for every
hop and root-degree filter, and enumerate the same counts."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("hop,rd,hubs", [(1, 0, False), (2, 0, False), (2, 3, False), (3, 2, False), (2, 0, True), (3, 2, True)])
def test_device_view_equals_host_view(vsgm, gpu, hop, rd, hubs):
    # hubs: a Chung-Lu graph whose 15 largest lists (up to 41,017 entries)
    # are long enough to be spread over the whole grid when expanded and copied
    g = vsgm.Graph.chung_lu(1 << 16, 1 << 21, 2.1, seed=5) if hubs else vsgm.Graph.rmat(10, 16, seed=7)
    V = g.vertex_count
    k = 4
    assign = np.random.default_rng(hop * 10 + rd).integers(0, k, size=V).astype(np.int32)
    vbm = vsgm.ViewBinManager.from_assignment(g, hop, k, assign)
    holder = vsgm.ViewBinHolder(V, g.entries * 4 + 4)
    p = vsgm.Plan.optimize(vsgm.Query.clique(3))
    ex = vsgm.PipelineExecutor(0, V, [p])
    ex.set_graph(g)
    for b in range(vbm.n_bins):
        vbm.materialize(b, holder, rd)
        rp, cl, src = holder.csr(V)
        i = ex.add_device_view(src, hop, rd)
        drp, dcl = ex.view_csr(i, V)
        assert np.array_equal(drp, rp), (hop, rd, b)
        assert np.array_equal(dcl, cl), (hop, rd, b)


def test_device_views_count_like_host_views(vsgm, oracle, gpu):
    g = vsgm.Graph.rmat(11, 16, seed=8)
    V = g.vertex_count
    plans = [vsgm.Plan.optimize(q) for q in (vsgm.Query.clique(4), vsgm.Query.preset(1), vsgm.Query.preset(4), vsgm.Query.preset(2))]
    hop = max(p.hop for p in plans)
    rd = min(p.root_degree for p in plans)
    rp, cl = g.csr()
    want = [oracle.count_plan(rp, cl, p.serialize()) for p in plans]
    assign = np.random.default_rng(3).integers(0, 6, size=V).astype(np.int32)
    ex = vsgm.PipelineExecutor(0, V, plans)
    ex.set_graph(g)
    for b in range(6):
        ex.add_device_view(np.nonzero(assign == b)[0].astype(np.uint32), hop, rd)
    ex.match_all()
    assert ex.total_counts == want
    # one bin at a time, each replacing the last (the -dv 1 consumer loop)
    ex.reset()
    for b in range(6):
        ex.transfer_device(np.nonzero(assign == b)[0].astype(np.uint32), hop, rd)
        ex.match()
    assert ex.total_counts == want
    # an empty bin builds an empty view and counts nothing
    ex.reset()
    i = ex.add_device_view(np.zeros(0, np.uint32), hop, rd)
    drp, dcl = ex.view_csr(i, V)
    assert drp[-1] == 0 and dcl.size == 0


def test_device_view_build_time(vsgm, gpu):
    """Host materialisation + upload vs device build for 8 bins (printed)."""
    g = vsgm.Graph.rmat(18, 16, seed=1)
    V = g.vertex_count
    assign = np.random.default_rng(0).integers(0, 8, size=V).astype(np.int32)
    vbm = vsgm.ViewBinManager.from_assignment(g, 2, 8, assign)
    p = vsgm.Plan.optimize(vsgm.Query.clique(3))
    holder = vsgm.ViewBinHolder(V, g.entries * 4 + 4)
    ex = vsgm.PipelineExecutor(0, V, [p])
    ex.set_option("hub_bytes", 0)
    t = time.perf_counter()
    for b in range(vbm.n_bins):
        vbm.materialize(b, holder, 2)
        ex.add_view(holder)
    host_s = time.perf_counter() - t
    ex2 = vsgm.PipelineExecutor(0, V, [p])
    ex2.set_option("hub_bytes", 0)
    ex2.set_graph(g)
    srcs = [np.nonzero(assign == b)[0].astype(np.uint32) for b in range(8)]
    t = time.perf_counter()
    for s in srcs:
        ex2.add_device_view(s, 2, 2)
    dev_s = time.perf_counter() - t
    print(f"[gpu] 8 two-hop bins of R-MAT s18: host build + upload {host_s:.3f}s, device build {dev_s:.3f}s", flush=True)
    ex.match_all()
    ex2.match_all()
    assert ex.total_counts == ex2.total_counts


def test_prefetched_views_count_like_built_views(vsgm, oracle, gpu):
    """Each bin's view built in the background (second context and stream)
    while the previous bin matches: per-bin counts equal the synchronous
    builds', the total equals the oracle's, a prefetch of another bin is
    dropped (the transfer builds its own view), and the executor can be
    destroyed with a prefetch in flight."""
    g = vsgm.Graph.rmat(12, 16, seed=11)
    V = g.vertex_count
    plans = [vsgm.Plan.optimize(q) for q in (vsgm.Query.clique(4), vsgm.Query.preset(1), vsgm.Query.preset(2))]
    hop = max(p.hop for p in plans)
    rd = min(p.root_degree for p in plans)
    rp, cl = g.csr()
    want = [oracle.count_plan(rp, cl, p.serialize()) for p in plans]
    assign = np.random.default_rng(5).integers(0, 6, size=V).astype(np.int32)
    bins = [np.nonzero(assign == b)[0].astype(np.uint32) for b in range(6)]
    sync = vsgm.PipelineExecutor(0, V, plans)
    sync.set_graph(g)
    per_bin = []
    for s in bins:
        sync.reset()
        sync.transfer_device(s, hop, rd)
        sync.match()
        per_bin.append(sync.total_counts)
    assert [sum(c[i] for c in per_bin) for i in range(len(plans))] == want
    ex = vsgm.PipelineExecutor(0, V, plans)
    ex.set_graph(g)
    print("[gpu] prefetched device views, 6 bins of R-MAT s12", flush=True)
    ex.transfer_device(bins[0], hop, rd)
    for b in range(6):
        if b + 1 < 6:
            ex.prefetch_device(bins[b + 1], hop, rd)
        ex.reset()
        ex.match()
        assert ex.total_counts == per_bin[b], b
        if b + 1 < 6:
            ex.transfer_device(bins[b + 1], hop, rd)
    # a prefetch of bin 2 followed by a transfer of bin 3: bin 3's own view
    ex.prefetch_device(bins[2], hop, rd)
    ex.transfer_device(bins[3], hop, rd)
    ex.reset()
    ex.match()
    assert ex.total_counts == per_bin[3]
    ex.prefetch_device(bins[4], hop, rd)
    del ex
