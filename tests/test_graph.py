"""CPU: CSR .bin format, generators, .vbmap partitions and view bins.

The .bin layout is the reference's (src/graph/graph_io.h:57-100, 313-339):
four size_t header words, row_ptrs u64[V+1], cols u32[E]. Views follow
src/view/view_bin.h:49-142: full lists for vertices within hop-1 of a
source whose degree reaches root_degree, empty lists elsewhere.
"""
import os
import struct

import numpy as np
import pytest


def test_bin_roundtrip_and_byte_layout(vsgm, tmp_path):
    g = vsgm.Graph.rmat(10, 8, seed=3)
    path = str(tmp_path / "g.bin")
    g.write_bin(path)
    raw = open(path, "rb").read()
    V, E = g.vertex_count, g.entries
    assert struct.unpack("<4Q", raw[:32]) == (4, 8, V, E)
    assert len(raw) == 32 + 8 * (V + 1) + 4 * E
    rp, cl = g.csr()
    assert np.array_equal(np.frombuffer(raw[32:32 + 8 * (V + 1)], np.uint64), rp)
    assert np.array_equal(np.frombuffer(raw[32 + 8 * (V + 1):], np.uint32), cl)
    h = vsgm.Graph.read_bin(path)
    r2, c2 = h.csr()
    assert np.array_equal(r2, rp) and np.array_equal(c2, cl)


def test_read_rejects_truncated_file(vsgm, tmp_path):
    g = vsgm.Graph.rmat(6, 4, seed=1)
    path = str(tmp_path / "t.bin")
    g.write_bin(path)
    data = open(path, "rb").read()
    open(path, "wb").write(data[:-4])
    with pytest.raises(vsgm.VsgmError):
        vsgm.Graph.read_bin(path)


def test_rmat_is_clean_deterministic_and_thread_independent(vsgm):
    a = vsgm.Graph.rmat(12, 16, seed=9, threads=1)
    b = vsgm.Graph.rmat(12, 16, seed=9, threads=8)
    c = vsgm.Graph.rmat(12, 16, seed=10, threads=8)
    assert a.validate() == ""
    ra, ca = a.csr()
    rb, cb = b.csr()
    assert np.array_equal(ra, rb) and np.array_equal(ca, cb)
    assert not np.array_equal(c.csr()[1][:1000], ca[:1000])
    assert a.vertex_count == 4096
    assert 0.6 * 16 * 4096 < a.edge_count <= 16 * 4096  # dedup removes the repeats


def test_rmat_skew_follows_the_quadrant_probabilities(vsgm):
    # with a = .57 the low ids collect most edges: vertex 0 is the largest hub
    g = vsgm.Graph.rmat(14, 16, seed=1)
    rp, _ = g.csr()
    deg = np.diff(rp)
    assert int(np.argmax(deg)) == 0
    assert deg[: 1 << 10].sum() > deg[-(1 << 10):].sum() * 20


def test_chung_lu_degree_power_law(vsgm):
    g = vsgm.Graph.chung_lu(1 << 14, 1 << 18, 2.3, seed=4)
    assert g.validate() == ""
    rp, _ = g.csr()
    deg = np.diff(rp).astype(np.float64)
    assert deg[:16].mean() > 10 * deg[-4096:].mean()


def test_from_edges_cleans_input(vsgm):
    g = vsgm.Graph.from_edges(5, [(0, 1), (1, 0), (2, 2), (3, 4), (0, 1)])
    rp, cl = g.csr()
    assert list(rp) == [0, 1, 2, 2, 3, 4]
    assert list(cl) == [1, 0, 4, 3]


def _ball(rp, cl, sources, radius):
    seen = set(sources)
    frontier = list(sources)
    for _ in range(radius):
        nxt = []
        for v in frontier:
            for w in cl[rp[v]:rp[v + 1]]:
                if w not in seen:
                    seen.add(int(w))
                    nxt.append(int(w))
        frontier = nxt
    return seen


@pytest.mark.parametrize("hop", [1, 2, 3])
def test_view_materialize_matches_ball(vsgm, hop):
    g = vsgm.Graph.rmat(9, 8, seed=5)
    rp, cl = g.csr()
    V = g.vertex_count
    rng = np.random.default_rng(hop)
    assign = rng.integers(0, 3, size=V)
    vbm = vsgm.ViewBinManager.from_assignment(g, hop, 3, assign)
    holder = vsgm.ViewBinHolder(V, g.entries * 4 + 4)
    root_degree = 3
    for b in range(3):
        vbm.materialize(b, holder, root_degree)
        vr, vc, src = holder.csr(V)
        assert sorted(src.tolist()) == sorted(np.nonzero(assign == b)[0].tolist())
        qual = [int(s) for s in src if rp[s + 1] - rp[s] >= root_degree]
        ball = _ball(rp, cl, qual, hop - 1)
        for v in range(V):
            if v in ball:
                assert np.array_equal(vc[vr[v]:vr[v + 1]], cl[rp[v]:rp[v + 1]])
            else:
                assert vr[v + 1] == vr[v]


def test_vbmap_roundtrip_and_split_reorder(vsgm, tmp_path):
    g = vsgm.Graph.rmat(10, 8, seed=2)
    V = g.vertex_count
    assign = np.arange(V) % 4
    path = str(tmp_path / "g.bin.2.hop.vbmap")
    vsgm.ViewBinManager.write_vbmap(path, 4, assign)
    lines = open(path).read().split()
    assert lines[0] == "4" and len(lines) == V + 1
    vbm = vsgm.ViewBinManager.load(g, 2, path)
    assert vbm.n_bins == 4 and vbm.max_sources == V // 4
    vbm.split(3, 1)  # (4/3 + 1) * 3 = 6 bins
    assert vbm.n_bins == 6
    holder = vsgm.ViewBinHolder(V, g.entries * 4 + 4)
    seen = []
    for b in range(vbm.n_bins):
        vbm.materialize(b, holder, 0)
        seen += holder.csr(V)[2].tolist()
    assert sorted(seen) == list(range(V))  # every vertex is the source of exactly one bin


def test_view_counts_sum_to_total_on_cpu(vsgm, oracle):
    # the oracle over each bin's view (sources = bin) adds up to the whole count
    g = vsgm.Graph.rmat(9, 16, seed=6)
    rp, cl = g.csr()
    plan = vsgm.Plan.optimize(vsgm.Query.preset(4))
    total = oracle.count_plan(rp, cl, plan.serialize())
    V = g.vertex_count
    assign = np.random.default_rng(0).integers(0, 4, size=V)
    vbm = vsgm.ViewBinManager.from_assignment(g, plan.hop, 4, assign)
    holder = vsgm.ViewBinHolder(V, g.entries * 4 + 4)
    s = 0
    for b in range(4):
        vbm.materialize(b, holder, plan.root_degree)
        vr, vc, src = holder.csr(V)
        s += oracle.count_plan(vr, vc, plan.serialize(), sources=src)
    assert s == total


def test_relabel_by_degree_is_an_isomorphism(vsgm):
    g = vsgm.Graph.rmat(10, 16, seed=7)
    h, new_id = g.relabel_by_degree(threads=3)
    assert h.validate() == ""
    rp, cl = g.csr()
    hr, hc = h.csr()
    n = g.vertex_count
    assert sorted(new_id.tolist()) == list(range(n))
    deg, hdeg = np.diff(rp), np.diff(hr)
    assert np.array_equal(hdeg[new_id], deg)
    assert np.all(np.diff(hdeg.astype(np.int64)) >= 0)
    # ties keep the id order
    old = np.empty(n, np.int64)
    old[new_id] = np.arange(n)
    same = hdeg[1:] == hdeg[:-1]
    assert np.all(old[1:][same] > old[:-1][same])
    # every list is the image of the old list
    for v in range(n):
        assert np.array_equal(hc[hr[new_id[v]]:hr[new_id[v] + 1]], np.sort(new_id[cl[rp[v]:rp[v + 1]]]))
    # and threads do not change the result
    h2, id2 = g.relabel_by_degree(threads=1)
    assert np.array_equal(id2, new_id) and all(np.array_equal(a, b) for a, b in zip(h2.csr(), (hr, hc)))


def test_relabel_by_degree_keeps_counts_on_cpu(vsgm, oracle):
    # the counts of every preset are a property of the graph, not of its ids
    g = vsgm.Graph.rmat(8, 16, seed=8)
    h, _ = g.relabel_by_degree()
    rp, cl = g.csr()
    hr, hc = h.csr()
    for k in range(17):
        text = vsgm.Plan.optimize(vsgm.Query.preset(k)).serialize()
        assert oracle.count_plan(rp, cl, text) == oracle.count_plan(hr, hc, text), k


def test_relabel_by_degree_shrinks_four_cycle_paths(vsgm, oracle):
    # the 4-cycle plan roots each embedding at its smallest id; with hubs last
    # the partial embeddings of its third level (two neighbours above the
    # root) are far fewer on a skewed graph
    g = vsgm.Graph.rmat(10, 16, seed=9)
    h, _ = g.relabel_by_degree()
    text = vsgm.Plan.optimize(vsgm.Query.preset(1)).serialize()
    a = oracle.count_plan(*g.csr(), text, level_paths=True)
    b = oracle.count_plan(*h.csr(), text, level_paths=True)
    assert a[0] == b[0]
    assert sum(b[1]) * 3 < sum(a[1])
