"""CPU: view-bin packing (vsgm_amd/csrc/host/packing.cc) against the
reference's own view_packing tool — byte-identical .vbmap and .vbstat on
the committed vectors (tests/golden/packing_vectors.json, written by
tests/golden/make_packing_vectors.py from oracle/_ref/view_packing), and
live on further seeded cases where that build is present — plus the
library call, error paths and invariants."""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "vsgm_amd", "bin", "view_packing")
REF_TOOL = os.path.join(ROOT, "oracle", "_ref", "view_packing")
VECTORS = json.load(open(os.path.join(ROOT, "tests", "golden", "packing_vectors.json")))


def setup_case(vsgm, d, scale, ef, gseed, part, first, second):
    g = vsgm.Graph.rmat(scale, ef, seed=gseed)
    gf = os.path.join(d, "g.bin")
    g.write_bin(gf)
    pf = f"{gf}.kmeans.{first}x{second}" if second else f"{gf}.kmeans.{first}"
    np.savetxt(pf, np.asarray(part), fmt="%d")
    return g, gf, pf


def run_tool(tool, gf, pf, first, hop, mem, pool, order):
    return subprocess.run([tool, "-gf", gf, "-pf", pf, "-h", str(hop), "-m", repr(float(mem)), "-t", str(first), "-s", str(pool), "-o",
                           str(order), "-d", "1"], capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("i", range(len(VECTORS)))
def test_golden_vectors(vsgm, tmp_path, i):
    v = VECTORS[i]
    scale, ef, gseed, _pseed, first, second, hop, mem, pool, order = v["case"]
    g, gf, pf = setup_case(vsgm, str(tmp_path), scale, ef, gseed, v["partition"], first, second)
    out = run_tool(TOOL, gf, pf, first, hop, mem, pool, order)
    assert out.returncode == 0, out.stderr
    base = f"{gf}.{hop}.hop"
    got = [int(x) for x in open(base + ".vbmap").read().split()]
    assert got[0] == v["bins"] and got[1:] == v["assignment"]
    assert open(base + ".vbstat").read() == v["vbstat"]
    # the library call gives the same bins
    max_entries = int((mem * 1000 * 1000 * 1000) / 4)
    bins, assign, sources, sizes = vsgm.pack_view_bins(g, hop, v["partition"], first, second, max_entries, pool, bool(order))
    assert bins == v["bins"] and assign.tolist() == v["assignment"]
    assert int(sources.sum()) == g.vertex_count
    assert np.array_equal(np.bincount(assign, minlength=bins), sources)


@pytest.mark.skipif(not os.path.exists(REF_TOOL), reason="oracle/_ref/view_packing is built only where the reference tree exists")
@pytest.mark.parametrize("seed", range(6))
def test_live_against_reference_tool(vsgm, tmp_path, seed):
    rng = np.random.default_rng(100 + seed)
    scale, ef = 11, int(rng.choice([4, 8]))
    first = int(rng.integers(1, 6))
    second = int(rng.choice([0, 0, 2, 3]))
    hop = int(rng.choice([1, 2, 2, 3]))
    pool = int(rng.integers(1, 5))
    order = int(rng.integers(0, 2))
    g = vsgm.Graph.rmat(scale, ef, seed=seed + 7)
    rp, cl = g.csr()
    part = rng.integers(0, first * (second or 1), g.vertex_count)
    deg = np.diff(rp).astype(np.int64)
    ball2 = int((deg + np.add.reduceat(np.append(deg[cl], 0), rp[:-1].astype(np.int64)) * (deg > 0)).max())
    mem = float({1: 4 * 1.5 * deg.max(), 2: 4 * 1.3 * ball2, 3: 4 * 1.5 * len(cl)}[hop] / 1e9)
    outs = []
    for name, tool in (("ref", REF_TOOL), ("me", TOOL)):
        d = tmp_path / name
        d.mkdir()
        _, gf, pf = setup_case(vsgm, str(d), scale, ef, seed + 7, part, first, second)
        r = run_tool(tool, gf, pf, first, hop, mem, pool, order)
        base = f"{gf}.{hop}.hop"
        outs.append((r.returncode == 0, open(base + ".vbmap").read() if r.returncode == 0 else None,
                     open(base + ".vbstat").read() if r.returncode == 0 else None, open(gf + ".degrees.bin", "rb").read()))
    bins = outs[0][1].split()[0] if outs[0][0] else "-"
    print(f"case scale={scale} ef={ef} layers={first}x{second} hop={hop} pool={pool} order={order}: bins={bins}")
    assert outs[0] == outs[1]


def test_errors(vsgm, tmp_path):
    g, gf, pf = setup_case(vsgm, str(tmp_path), 9, 8, 1, np.zeros(512, int), 1, 0)
    # a hub's 2-hop view cannot fit a 100-entry bin
    assert run_tool(TOOL, gf, pf, 1, 2, 4e-7, 1, 0).returncode == 1
    # -t must match the partition's first layer
    assert run_tool(TOOL, gf, pf, 2, 2, 1.0, 1, 0).returncode == 2
    with pytest.raises(vsgm.VsgmError):
        vsgm.pack_view_bins(g, 2, np.full(g.vertex_count, 3), 2)
    with pytest.raises(vsgm.VsgmError):
        vsgm.pack_view_bins(g, 2, np.zeros(g.vertex_count), 1, 0, 100)


def test_one_bin_when_everything_fits(vsgm):
    g = vsgm.Graph.rmat(9, 8, seed=2)
    part = np.random.default_rng(0).integers(0, 3, g.vertex_count)
    bins, assign, sources, sizes = vsgm.pack_view_bins(g, 2, part, 3, 0, 1 << 40)
    # one bin per cluster, merged into one
    assert bins == 1 and not assign.any() and sources.tolist() == [g.vertex_count]


def test_cluster_reorder_and_compaction(vsgm):
    # two disjoint triangles {0,1,2}, {3,4,5} and a 5-vertex path 6..10
    edges = [(0, 1), (1, 2), (0, 2), (3, 4), (4, 5), (3, 5), (6, 7), (7, 8), (8, 9), (9, 10)]
    g = vsgm.Graph.from_edges(11, edges)
    a = [0, 0, 0, 2, 2, 2, 3, 3, 3, 3, 3]  # cluster 1 empty
    # distinct closed-neighbourhood vertices: cluster 0 -> 3, 2 -> 3, 3 -> 5, 1 -> 0
    r = vsgm.reorder_clusters(g, 4, a)
    assert r.tolist() == [1, 1, 1, 2, 2, 2, 0, 0, 0, 0, 0]
    kept, c = vsgm.compact_clusters(4, a)
    assert kept == 3 and c.tolist() == [0, 0, 0, 1, 1, 1, 2, 2, 2, 2, 2]
    with pytest.raises(vsgm.VsgmError):
        vsgm.compact_clusters(2, a)


def test_kmeans_tool_rejects_bad_levels_before_any_work(tmp_path):
    exe = os.path.join(ROOT, "vsgm_amd", "bin", "kmeans")
    for levels in ("2^0", "2^17", "4x", "0x", "17"):
        out = subprocess.run([exe, "-f", str(tmp_path / "missing.bin"), "-k", levels], capture_output=True, text=True, timeout=60)
        assert out.returncode == 2, (levels, out.stderr)


def test_degree_cache_must_match_the_graph(vsgm, tmp_path):
    g, gf, pf = setup_case(vsgm, str(tmp_path), 9, 8, 1, np.zeros(512, int), 1, 0)
    assert run_tool(TOOL, gf, pf, 1, 2, 1.0, 1, 0).returncode == 0
    raw = np.fromfile(gf + ".degrees.bin", dtype=np.uint64)
    rp, _ = g.csr()
    assert raw[0] == g.vertex_count and np.array_equal(raw[1:], np.diff(rp))
    raw[5] += 1
    raw.tofile(gf + ".degrees.bin")
    assert run_tool(TOOL, gf, pf, 1, 2, 1.0, 1, 0).returncode == 1


REF_FEATURES = os.path.join(ROOT, "oracle", "_ref", "features")


@pytest.mark.skipif(not os.path.exists(REF_FEATURES), reason="oracle/_ref/features is built only where the reference tree exists")
@pytest.mark.parametrize("scale,ef,seed,flags,out", [(9, 8, 1, ["-h", "1"], "1hop"), (11, 4, 2, ["-h", "1"], "1hop"),
                                                     (9, 8, 3, ["-h", "2"], "2hop"), (10, 4, 4, ["-h", "2"], "2hop"),
                                                     (8, 8, 5, ["-dn", "3"], "x3"), (10, 4, 6, ["-dn", "2"], "x2")])
def test_features_tool_matches_reference_tool(vsgm, tmp_path, scale, ef, seed, flags, out):
    g = vsgm.Graph.rmat(scale, ef, seed=seed)
    outs = []
    for name, tool in (("ref", REF_FEATURES), ("me", os.path.join(ROOT, "vsgm_amd", "bin", "features"))):
        d = tmp_path / name
        d.mkdir()
        g.write_bin(str(d / "g.bin"))
        r = subprocess.run([tool, "-f", str(d / "g.bin"), "-t", "4"] + flags, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        outs.append((d / f"g.bin.{out}.bin").read_bytes())
    assert outs[0] == outs[1]


def test_features_tool_writes_closed_rows(vsgm, tmp_path):
    g = vsgm.Graph.from_edges(5, [(0, 1), (1, 2), (3, 4)])
    g.write_bin(str(tmp_path / "g.bin"))
    exe = os.path.join(ROOT, "vsgm_amd", "bin", "features")
    assert subprocess.run([exe, "-f", str(tmp_path / "g.bin"), "-h", "1"], capture_output=True, timeout=60).returncode == 0
    rp, cols = vsgm.Graph.read_bin(str(tmp_path / "g.bin.1hop.bin")).csr()
    assert rp.tolist() == [0, 2, 5, 7, 9, 11] and cols.tolist() == [0, 1, 0, 1, 2, 1, 2, 3, 4, 3, 4]
    assert subprocess.run([exe, "-f", str(tmp_path / "g.bin"), "-h", "2"], capture_output=True, timeout=60).returncode == 0
    rp, cols = vsgm.Graph.read_bin(str(tmp_path / "g.bin.2hop.bin")).csr()
    assert rp.tolist() == [0, 3, 6, 9, 11, 13] and cols.tolist() == [0, 1, 2, 0, 1, 2, 0, 1, 2, 3, 4, 3, 4]
    assert subprocess.run([exe, "-f", str(tmp_path / "g.bin"), "-h", "-2"], capture_output=True, timeout=60).returncode == 2


FENNEL = json.load(open(os.path.join(ROOT, "tests", "golden", "fennel_vectors.json")))
REF_FENNEL = os.path.join(ROOT, "oracle", "_ref", "fennel")


@pytest.mark.parametrize("i", range(len(FENNEL)))
def test_fennel_golden_vectors(vsgm, tmp_path, i):
    scale, ef, seed, parts, hop = FENNEL[i]["case"]
    g = vsgm.Graph.rmat(scale, ef, seed=seed)
    gf = str(tmp_path / "g.bin")
    g.write_bin(gf)
    out = subprocess.run([os.path.join(ROOT, "vsgm_amd", "bin", "fennel"), "-f", gf, "-p", str(parts), "-h", str(hop)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert [int(x) for x in open(f"{gf}.fennel.part.{parts}").read().split()] == FENNEL[i]["parts"]
    assert vsgm.fennel(g, parts, hop).tolist() == FENNEL[i]["parts"]


@pytest.mark.skipif(not os.path.exists(REF_FENNEL), reason="oracle/_ref/fennel is built only where the reference tree exists")
@pytest.mark.parametrize("seed", range(4))
def test_fennel_live_against_reference_tool(vsgm, tmp_path, seed):
    rng = np.random.default_rng(300 + seed)
    g = vsgm.Graph.rmat(int(rng.integers(8, 12)), int(rng.choice([4, 8, 16])), seed=seed + 40)
    parts, hop = int(rng.integers(1, 12)), int(rng.integers(1, 3))
    outs = []
    for name, tool in (("ref", REF_FENNEL), ("me", os.path.join(ROOT, "vsgm_amd", "bin", "fennel"))):
        d = tmp_path / name
        d.mkdir()
        g.write_bin(str(d / "g.bin"))
        r = subprocess.run([tool, "-f", str(d / "g.bin"), "-p", str(parts), "-h", str(hop)], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        outs.append((d / f"g.bin.fennel.part.{parts}").read_bytes())
    assert outs[0] == outs[1]


@pytest.mark.parametrize("flags", [["-p", "2"], ["-kc", "5"], ["-km", "4"]])
def test_query_test_tool_prints_the_planner_plans(vsgm, flags):
    out = subprocess.run([os.path.join(ROOT, "vsgm_amd", "bin", "query_test")] + flags, capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    if flags[0] == "-p":
        queries = [vsgm.Query.preset(int(flags[1]))]
    elif flags[0] == "-kc":
        queries = [vsgm.Query.clique(int(flags[1]))]
    else:
        queries = vsgm.Query.graphlets(int(flags[1]))
    plans = [vsgm.Plan.optimize(q) for q in queries]
    for p in plans:
        assert p.serialize() in out.stdout
    if flags[0] == "-km":
        assert f"largest hop: {max(p.hop for p in plans)}" in out.stdout


from hypothesis import given, settings, HealthCheck, strategies as st  # noqa: E402


@pytest.mark.skipif(not (os.path.exists(REF_TOOL) and os.path.exists(REF_FENNEL)),
                    reason="the reference tools are built only where the reference tree exists")
@settings(max_examples=25, deadline=None, suppress_health_check=[HealthCheck.function_scoped_fixture])
@given(edges=st.lists(st.tuples(st.integers(0, 40), st.integers(0, 40)), min_size=1, max_size=150),
       first=st.integers(1, 4), hop=st.integers(1, 3), pool=st.integers(1, 3), order=st.integers(0, 1), slack=st.floats(1.0, 3.0),
       parts=st.integers(1, 6), fhop=st.integers(1, 2), seed=st.integers(0, 1000))
def test_random_graphs_pack_and_partition_like_the_reference_tools(vsgm, tmp_path, edges, first, hop, pool, order, slack, parts, fhop,
                                                                   seed):
    g = vsgm.Graph.from_edges(41, np.array(edges, np.uint32))
    rp, cl = g.csr()
    # a limit every single source fits: its ball is at most the whole graph
    limit = max(int(len(cl) * slack), 1)
    part = np.random.default_rng(seed).integers(0, first, g.vertex_count)
    outs = []
    for name, pack_tool, fennel_tool in (("ref", REF_TOOL, REF_FENNEL),
                                         ("me", TOOL, os.path.join(ROOT, "vsgm_amd", "bin", "fennel"))):
        # a fresh directory per example: each tool keeps a degree cache beside the graph
        d = tempfile.mkdtemp(prefix=name, dir=str(tmp_path))
        gf, pf = os.path.join(d, "g.bin"), os.path.join(d, f"g.bin.kmeans.{first}")
        g.write_bin(gf)
        np.savetxt(pf, part, fmt="%d")
        r = run_tool(pack_tool, gf, pf, first, hop, limit * 4 / 1e9, pool, order)
        f = subprocess.run([fennel_tool, "-f", gf, "-p", str(parts), "-h", str(fhop)], capture_output=True, timeout=60)
        base = f"{gf}.{hop}.hop"
        outs.append((r.returncode, open(base + ".vbmap").read() if r.returncode == 0 else None,
                     open(base + ".vbstat").read() if r.returncode == 0 else None, f.returncode,
                     open(f"{gf}.fennel.part.{parts}").read()))
    assert outs[0] == outs[1]


FEATURES = json.load(open(os.path.join(ROOT, "tests", "golden", "features_vectors.json")))


@pytest.mark.parametrize("i", range(len(FEATURES)))
def test_features_golden_vectors(vsgm, tmp_path, i):
    import hashlib

    scale, ef, seed, flags, out = FEATURES[i]["case"]
    gf = str(tmp_path / "g.bin")
    vsgm.Graph.rmat(scale, ef, seed=seed).write_bin(gf)
    r = subprocess.run([os.path.join(ROOT, "vsgm_amd", "bin", "features"), "-f", gf] + flags, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert hashlib.sha256(open(f"{gf}.{out}.bin", "rb").read()).hexdigest() == FEATURES[i]["sha256"]
