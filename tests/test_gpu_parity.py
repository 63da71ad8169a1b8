"""GPU parity of the whole enumeration path (PipelineExecutor over the
C-ABI) against the oracle: bit-exact counts on seeded graphs for every
preset, clique, graphlet and connected 3..6-vertex shape, closed forms,
batching / view-bin / matching-order invariance, edge cases, and the
pm_pipeline CLI with the reference's file formats."""
import math
import os
import subprocess

import numpy as np
import pytest

from vsgm_amd.shapes import all_shapes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_gpu(vsgm, graph, plans, sources=None, batch_bytes=None, **options):
    ex = vsgm.PipelineExecutor(0, graph.vertex_count, plans)
    if batch_bytes:
        ex.set_option("batch_bytes", batch_bytes)
    for k, v in options.items():
        ex.set_option(k, v)
    ex.transfer_graph(graph, sources)
    ex.match()
    return ex.total_counts


def oracle_counts(oracle, graph, plans, sources=None):
    rp, cl = graph.csr()
    return [oracle.count_plan(rp, cl, p.serialize(), sources=sources) for p in plans]


@pytest.fixture(scope="module")
def rmat9(vsgm, gpu):
    return vsgm.Graph.rmat(9, 16, seed=21)


def test_presets_and_cliques(vsgm, oracle, rmat9):
    qs = [vsgm.Query.preset(k) for k in range(19)] + [vsgm.Query.clique(k) for k in range(3, 7)]
    plans = [vsgm.Plan.optimize(q) for q in qs]
    print("[gpu] presets+cliques on R-MAT s9", flush=True)
    assert run_gpu(vsgm, rmat9, plans) == oracle_counts(oracle, rmat9, plans)


def test_graphlets(vsgm, oracle, rmat9):
    for k in (3, 4, 5):
        plans = [vsgm.Plan.optimize(q) for q in vsgm.Query.graphlets(k)]
        print(f"[gpu] graphlets k={k}", flush=True)
        assert run_gpu(vsgm, rmat9, plans) == oracle_counts(oracle, rmat9, plans)


def test_all_connected_shapes(vsgm, oracle, gpu):
    g = vsgm.Graph.rmat(8, 8, seed=4)
    plans = [vsgm.Plan.optimize(vsgm.Query.from_matrix(a)) for _, _, a in all_shapes()]
    print(f"[gpu] {len(plans)} shapes on R-MAT s8", flush=True)
    assert run_gpu(vsgm, g, plans) == oracle_counts(oracle, g, plans)


def _complete(vsgm, n):
    return vsgm.Graph.from_edges(n, [(i, j) for i in range(n) for j in range(i + 1, n)])


def test_closed_forms(vsgm, oracle, gpu):
    K = 9
    g = _complete(vsgm, K)
    qs = [vsgm.Query.preset(k) for k in range(19)]
    counts = run_gpu(vsgm, g, [vsgm.Plan.optimize(q) for q in qs])
    for k, (q, c) in enumerate(zip(qs, counts)):
        n, _, adj, _ = q.info()
        assert c == math.factorial(K) // math.factorial(K - n) // oracle.automorphisms(adj), f"P{k}"
    # K_{5,6}: no triangles, C(5,2)*C(6,2) 4-cycles
    kb = vsgm.Graph.from_edges(11, [(i, 5 + j) for i in range(5) for j in range(6)])
    tri, sq = run_gpu(vsgm, kb, [vsgm.Plan.optimize(vsgm.Query.clique(3)), vsgm.Plan.optimize(vsgm.Query.preset(1))])
    assert (tri, sq) == (0, 10 * 15)
    # 6x7 grid: (6-1)*(7-1) 4-cycles; Petersen: 0 triangles, 0 4-cycles, 12 5-cycles
    grid = vsgm.Graph.from_edges(42, [(r * 7 + c, r * 7 + c + 1) for r in range(6) for c in range(6)] +
                                 [(r * 7 + c, (r + 1) * 7 + c) for r in range(5) for c in range(7)])
    assert run_gpu(vsgm, grid, [vsgm.Plan.optimize(vsgm.Query.preset(1))]) == [30]
    pet = vsgm.Graph.from_edges(10, [(i, (i + 1) % 5) for i in range(5)] + [(5 + i, 5 + (i + 2) % 5) for i in range(5)] +
                                [(i, i + 5) for i in range(5)])
    plans = [vsgm.Plan.optimize(vsgm.Query.clique(3)), vsgm.Plan.optimize(vsgm.Query.preset(1)),
             vsgm.Plan.optimize(vsgm.Query.preset(16))]
    assert run_gpu(vsgm, pet, plans) == [0, 0, 12]
    # wheel W_12 (hub + 12-cycle): 12 triangles
    wheel = vsgm.Graph.from_edges(13, [(0, i) for i in range(1, 13)] + [(i, i % 12 + 1) for i in range(1, 13)])
    assert run_gpu(vsgm, wheel, [vsgm.Plan.optimize(vsgm.Query.clique(3))]) == [12]


def test_hub_index_invariance(vsgm, oracle, rmat9):
    """Counts are the same with the membership index off, on for the
    largest lists only, and on for every non-empty list."""
    plans = [vsgm.Plan.optimize(q) for q in [vsgm.Query.preset(k) for k in range(19)] + [vsgm.Query.clique(5)]]
    want = oracle_counts(oracle, rmat9, plans)
    for opts in ({"hub_bytes": 0}, {"hub_min_degree": 32}, {"hub_min_degree": 1}, {"hub_min_degree": 1, "hub_bytes": 64 * 40},
                 {"isect_kernel": 0}, {"isect_kernel": 0, "hub_bytes": 0}):
        print(f"[gpu] hub index {opts}", flush=True)
        assert run_gpu(vsgm, rmat9, plans, **opts) == want, opts
    ex = vsgm.PipelineExecutor(0, rmat9.vertex_count, plans[:1])
    ex.set_option("hub_min_degree", 1)
    ex.transfer_graph(rmat9)
    assert ex.hub_count() == int(np.count_nonzero(np.diff(rmat9.csr()[0])))
    ex.set_option("hub_bytes", 64 * 40)  # 40 rows of 512 bits
    ex.transfer_graph(rmat9)
    assert ex.hub_count() == 40


def test_batching_invariance(vsgm, oracle, rmat9):
    plans = [vsgm.Plan.optimize(q) for q in (vsgm.Query.clique(4), vsgm.Query.preset(1), vsgm.Query.preset(4),
                                             vsgm.Query.preset(2), vsgm.Query.preset(13))]
    want = oracle_counts(oracle, rmat9, plans)
    for bb in (1 << 14, 1 << 17):
        print(f"[gpu] batch_bytes={bb}", flush=True)
        assert run_gpu(vsgm, rmat9, plans, batch_bytes=bb) == want


def test_view_bins_sum_to_total(vsgm, oracle, gpu):
    g = vsgm.Graph.rmat(10, 16, seed=8)
    plans = [vsgm.Plan.optimize(q) for q in (vsgm.Query.clique(4), vsgm.Query.preset(1), vsgm.Query.preset(4))]
    hop = max(p.hop for p in plans)
    rd = min(p.root_degree for p in plans)
    want = oracle_counts(oracle, g, plans)
    rng = np.random.default_rng(2)
    k = 5
    assign = rng.integers(0, k, size=g.vertex_count).astype(np.int32)
    vbm = vsgm.ViewBinManager.from_assignment(g, hop, k, assign)
    vbm.reorder()
    ex = vsgm.PipelineExecutor(0, g.vertex_count, plans, max_sources=vbm.max_sources)
    holder = vsgm.ViewBinHolder(g.vertex_count, g.entries * 4 + 4)
    for b in range(vbm.n_bins):
        vbm.materialize(b, holder, rd)
        ex.transfer(holder)
        ex.match()
    assert ex.total_counts == want


def test_matching_orders_agree(vsgm, oracle, rmat9):
    import itertools

    for q in (vsgm.Query.preset(4), vsgm.Query.preset(1), vsgm.Query.preset(11)):
        base = oracle_counts(oracle, rmat9, [vsgm.Plan.optimize(q)])[0]
        n, _, adj, _ = q.info()
        plans = []
        for perm in itertools.permutations(range(n)):
            if all(any(adj[perm[i], perm[j]] for j in range(i)) for i in range(1, n)):
                plans.append(vsgm.Plan.from_order(q, perm))
        plans = plans[:: max(1, len(plans) // 12)]
        assert run_gpu(vsgm, rmat9, plans) == [base] * len(plans)


def test_edge_cases(vsgm, oracle, gpu):
    tri = vsgm.Plan.optimize(vsgm.Query.clique(3))
    sq = vsgm.Plan.optimize(vsgm.Query.preset(1))
    # no edges, isolated vertices only
    empty = vsgm.Graph.from_edges(17, [])
    assert run_gpu(vsgm, empty, [tri, sq]) == [0, 0]
    # a single edge
    one = vsgm.Graph.from_edges(3, [(0, 2)])
    assert run_gpu(vsgm, one, [tri, sq]) == [0, 0]
    # star with a hub of degree 5000 plus a triangle among leaves
    star = vsgm.Graph.from_edges(5001, [(0, i) for i in range(1, 5001)] + [(1, 2)])
    assert run_gpu(vsgm, star, [tri, sq]) == [1, 0]
    # sources restricted to a subset, and an empty subset
    g = vsgm.Graph.rmat(8, 16, seed=1)
    src = list(range(0, g.vertex_count, 3))
    assert run_gpu(vsgm, g, [tri, sq], sources=src) == oracle_counts(oracle, g, [tri, sq], sources=src)
    # a source outside the vertex range is refused on the host, before any upload or launch,
    # on every route a view takes (whole graph, device-built, host holder)
    ex = vsgm.PipelineExecutor(0, g.vertex_count, [tri])
    bad = [0, 5, g.vertex_count]
    with pytest.raises(vsgm.VsgmError, match="is not a vertex"):
        ex.transfer_graph(g, bad)
    ex.set_graph(g)
    with pytest.raises(vsgm.VsgmError, match="is not a vertex"):
        ex.transfer_device(bad, 2, 0)
    with pytest.raises(vsgm.VsgmError, match="is not a vertex"):
        ex.add_device_view(bad, 2, 0)
    ex.transfer_graph(g, src)  # the executor still works afterwards
    ex.match()
    assert ex.total_counts == oracle_counts(oracle, g, [tri], sources=src)


def test_large_hub_graph(vsgm, oracle, gpu):
    # hubs with degrees in the tens of thousands exercise the long-segment paths
    g = vsgm.Graph.rmat(16, 16, seed=9)
    assert g.max_degree > 5000
    plans = [vsgm.Plan.optimize(q) for q in (vsgm.Query.clique(3), vsgm.Query.clique(4), vsgm.Query.preset(2))]
    rp, cl = g.csr()
    got = run_gpu(vsgm, g, plans)
    assert got[0] == oracle.triangles(rp, cl)
    assert got[1] == oracle.four_cliques(rp, cl)
    assert got[2] == oracle.diamonds(rp, cl)


def test_full_size_triangles_and_cliques(vsgm, oracle, gpu):
    """Full-size inputs: R-MAT scale 22 triangles against the textbook counter
    on the host cores, scale 20 diamonds and scale 18 4-cliques likewise,
    plus count invariance under a forced multi-batch run."""
    g = vsgm.Graph.rmat(22, 16, seed=1)
    rp, cl = g.csr(copy=False)
    tri = vsgm.Plan.optimize(vsgm.Query.clique(3))
    print("[gpu] s22 triangles", flush=True)
    assert run_gpu(vsgm, g, [tri]) == [oracle.triangles(rp, cl)]
    del g, rp, cl
    g = vsgm.Graph.rmat(20, 16, seed=1)
    rp, cl = g.csr(copy=False)
    dia = vsgm.Plan.optimize(vsgm.Query.preset(2))
    print("[gpu] s20 diamonds", flush=True)
    got = run_gpu(vsgm, g, [dia])
    assert got == [oracle.diamonds(rp, cl)]
    assert run_gpu(vsgm, g, [dia], batch_bytes=64 << 20) == got
    del g, rp, cl
    g = vsgm.Graph.rmat(18, 16, seed=1)
    rp, cl = g.csr(copy=False)
    print("[gpu] s18 4-cliques", flush=True)
    assert run_gpu(vsgm, g, [vsgm.Plan.optimize(vsgm.Query.clique(4))]) == [oracle.four_cliques(rp, cl)]


def _golden(vsgm, scale):
    """Golden counts of R-MAT scale `scale` (tests/golden/make_rmat_counts.py)
    and the graph they were computed on (its CSR digest checked)."""
    import hashlib
    import json

    rec = json.load(open(os.path.join(ROOT, "tests", "golden", "rmat_counts.json")))[f"rmat{scale}"]
    g = vsgm.Graph.rmat(scale, rec["edge_factor"], *rec["abc"], seed=rec["seed"])
    rp, cl = g.csr(copy=False)
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(rp, dtype=np.uint64).tobytes())
    h.update(np.ascontiguousarray(cl, dtype=np.uint32).tobytes())
    assert h.hexdigest() == rec["csr_sha256"], "the generator no longer makes the graph of the golden counts"
    return g, rec["counts"]


def test_benchmark_graph_counts_against_golden(vsgm, gpu):
    """The benchmark pipeline (k-means with k = 1, 2-hop view with the
    root-degree filter, heuristic plans) on R-MAT scale 22: 4-cliques and
    diamonds at full size, and all three benchmark patterns at scale 20,
    equal the golden counts of the oracle's textbook counters (the scale-22
    4-cycle, ~200 s on the GPU, is checked by every bench run)."""
    from vsgm_amd import partition

    for scale, keys in ((22, ("kc4", "p2")), (20, ("kc4", "p2", "p1"))):
        g, want = _golden(vsgm, scale)
        qs = {"kc4": vsgm.Query.clique(4), "p2": vsgm.Query.preset(2), "p1": vsgm.Query.preset(1)}
        plans = [vsgm.Plan.optimize(qs[k]) for k in keys]
        hop = max(p.hop for p in plans)
        rd = min(p.root_degree for p in plans)
        assign, _, _ = partition.kmeans(g, 1, iters=1, seed=1)
        vbm = vsgm.ViewBinManager.from_assignment(g, hop, 1, assign)
        holder = vsgm.ViewBinHolder(g.vertex_count, max(4, g.entries * 4))
        vbm.materialize(0, holder, rd)
        ex = vsgm.PipelineExecutor(0, g.vertex_count, plans, max_sources=vbm.max_sources)
        ex.transfer(holder)
        print(f"[gpu] R-MAT s{scale} {keys} against the golden counts", flush=True)
        ex.match()
        assert ex.total_counts == [want[k] for k in keys], (scale, ex.total_counts)
        del ex, holder, vbm, g


def test_degree_relabelled_graph_counts(vsgm, oracle, gpu):
    """Relabelling the ids in ascending degree order changes no count: every
    preset and clique on a relabelled R-MAT graph equals the oracle on the
    graph as generated, over the whole graph and over three view bins, and
    the scale-20 benchmark patterns equal the golden counts."""
    g = vsgm.Graph.rmat(10, 16, seed=27)
    h, new_id = g.relabel_by_degree()
    plans = [vsgm.Plan.optimize(vsgm.Query.preset(k)) for k in range(17)] + [vsgm.Plan.optimize(vsgm.Query.clique(k)) for k in (3, 4, 5)]
    want = oracle_counts(oracle, g, plans)
    print("[gpu] degree-relabelled R-MAT s10, presets + cliques", flush=True)
    assert run_gpu(vsgm, h, plans) == want
    assign = np.random.default_rng(3).integers(0, 3, size=h.vertex_count)
    total = np.zeros(len(plans), np.int64)
    for b in range(3):
        total += run_gpu(vsgm, h, plans, sources=np.nonzero(assign == b)[0].astype(np.uint32))
    assert total.tolist() == want
    g20, gold = _golden(vsgm, 20)
    h20, _ = g20.relabel_by_degree()
    keys = ("kc4", "p2", "p1")
    qs = {"kc4": vsgm.Query.clique(4), "p2": vsgm.Query.preset(2), "p1": vsgm.Query.preset(1)}
    print("[gpu] degree-relabelled R-MAT s20 against the golden counts", flush=True)
    assert run_gpu(vsgm, h20, [vsgm.Plan.optimize(qs[k]) for k in keys]) == [gold[k] for k in keys]


def test_pm_pipeline_cli(vsgm, oracle, gpu, tmp_path):
    g = vsgm.Graph.rmat(10, 16, seed=12)
    path = str(tmp_path / "rmat10.bin")
    g.write_bin(path)
    rng = np.random.default_rng(0)
    vsgm.ViewBinManager.write_vbmap(path + ".2.hop.vbmap", 3, rng.integers(0, 3, size=g.vertex_count))
    exe = os.path.join(ROOT, "vsgm_amd", "bin", "pm_pipeline")
    for flags, q in ((["-kc", "4"], vsgm.Query.clique(4)), (["-p", "4"], vsgm.Query.preset(4)),
                     (["-p", "1", "-rl", "1"], vsgm.Query.preset(1))):  # ids relabelled, .vbmap mapped
        out = subprocess.run([exe, "-f", path, "-m", "0.1", "-qs", "2", "-pn", "2", "-cn", "1", "-t", "4"] + flags,
                             capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr[-2000:]
        line = [x for x in out.stdout.splitlines() if x.startswith("total count for query 0:")][0]
        assert int(line.split(":")[1]) == oracle_counts(oracle, g, [vsgm.Plan.optimize(q)])[0]


def test_pm_pipeline_cli_device_views(vsgm, oracle, gpu, tmp_path):
    """-gk partitions with the GPU k-means instead of the .vbmap file, -gp
    packs its clusters into size-bounded bins, -dv 1 builds every view on the
    consumer's GPU. Counts stay those of the graph."""
    g = vsgm.Graph.rmat(10, 16, seed=13)
    path = str(tmp_path / "rmat10.bin")
    g.write_bin(path)
    exe = os.path.join(ROOT, "vsgm_amd", "bin", "pm_pipeline")
    want = oracle_counts(oracle, g, [vsgm.Plan.optimize(vsgm.Query.clique(4))])[0]
    for extra in (["-gk", "3"], ["-gk", "3", "-dv", "1"], ["-gk", "1", "-dv", "1"], ["-gk", "3", "-gp", "2", "-dv", "1", "-m", "0.0001"]):
        print("[gpu] pm_pipeline", extra, flush=True)
        out = subprocess.run([exe, "-f", path, "-m", "0.1", "-qs", "2", "-pn", "2", "-cn", "1", "-t", "4", "-kc", "4"] + extra,
                             capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr[-2000:]
        line = [x for x in out.stdout.splitlines() if x.startswith("total count for query 0:")][0]
        assert int(line.split(":")[1]) == want


def test_offline_tool_chain(vsgm, oracle, gpu, tmp_path):
    """SNAP text -> preprocess_graph -g 0 -> kmeans (GPU) -> view_packing ->
    pm_pipeline, the reference's tool chain, ends in the graph's count."""
    g0 = vsgm.Graph.rmat(10, 16, seed=21)
    rp, cl = g0.csr()
    src = np.repeat(np.arange(g0.vertex_count), np.diff(rp).astype(np.int64))
    keep = src < cl
    text = "# FromNodeId\tToNodeId\n" + "".join(f"{a + 3}\t{b + 3}\n" for a, b in zip(src[keep], cl[keep]))
    path = str(tmp_path / "g.ungraph.txt")
    open(path, "w").write(text)
    bindir = os.path.join(ROOT, "vsgm_amd", "bin")

    def run(tool, *args):
        print("[gpu]", tool, *args, flush=True)
        out = subprocess.run([os.path.join(bindir, tool)] + list(args), capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr[-2000:]
        return out.stdout

    run("preprocess_graph", "-f", path, "-d", "0", "-g", "0")
    g = vsgm.Graph.read_bin(path + ".bin")
    grp, gcl = g.csr()
    deg = np.diff(grp).astype(np.int64)
    ball2 = int((deg + np.add.reduceat(np.append(deg[gcl], 0), grp[:-1].astype(np.int64)) * (deg > 0)).max())
    run("kmeans", "-f", path + ".bin", "-k", "3", "-i", "5", "-rc", "1")
    assert len(open(path + ".bin.kmeans.3").read().split()) == g.vertex_count
    run("view_packing", "-gf", path + ".bin", "-pf", path + ".bin.kmeans.3", "-h", "2", "-m", repr(4 * 1.3 * ball2 / 1e9), "-t", "3",
        "-s", "2")
    out = run("pm_pipeline", "-f", path + ".bin", "-m", "0.1", "-qs", "2", "-pn", "2", "-cn", "1", "-t", "4", "-kc", "4")
    line = [x for x in out.splitlines() if x.startswith("total count for query 0:")][0]
    assert int(line.split(":")[1]) == oracle_counts(oracle, g, [vsgm.Plan.optimize(vsgm.Query.clique(4))])[0]


def test_kmeans_tool_reads_features_graphs_as_the_reference_does(vsgm, gpu, tmp_path):
    """`features -h 1` then `kmeans -f G.bin.1hop.bin` (the reference's
    documented steps) gives the partition `kmeans -f G.bin` gives: both
    cluster the closed rows."""
    g = vsgm.Graph.rmat(10, 8, seed=31)
    path = str(tmp_path / "g.bin")
    g.write_bin(path)
    bindir = os.path.join(ROOT, "vsgm_amd", "bin")
    for args in (["features", "-f", path, "-h", "1"], ["kmeans", "-f", path, "-k", "4", "-i", "4"],
                 ["kmeans", "-f", path + ".1hop.bin", "-k", "4", "-i", "4"]):
        print("[gpu]", *args, flush=True)
        out = subprocess.run([os.path.join(bindir, args[0])] + args[1:], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr[-2000:]
    plain = np.loadtxt(path + ".kmeans.4", dtype=np.int64)
    closed = np.loadtxt(path + ".1hop.bin.kmeans.4", dtype=np.int64)
    # the two runs see the same rows (opening closed rows is exact, checked on
    # the CPU) and a symmetric graph's k-means is deterministic
    assert len(plain) == g.vertex_count and np.array_equal(plain, closed)


def test_pm_pipeline_checkpoint_resume(vsgm, oracle, gpu, tmp_path):
    """-ck FILE: each finished bin appends its counts under a header keyed
    to the run (plans, graph, hop, root-degree filter); a rerun skips the
    bins listed (if their sources still match) and totals stay exact; a run
    with another query, or a file without the header, is refused."""
    g = vsgm.Graph.rmat(10, 16, seed=14)
    path = str(tmp_path / "g.bin")
    g.write_bin(path)
    vsgm.ViewBinManager.write_vbmap(path + ".2.hop.vbmap", 4, np.random.default_rng(1).integers(0, 4, size=g.vertex_count))
    exe = os.path.join(ROOT, "vsgm_amd", "bin", "pm_pipeline")
    ck = str(tmp_path / "ck.txt")
    want = oracle_counts(oracle, g, [vsgm.Plan.optimize(vsgm.Query.clique(4))])[0]

    def run(*extra, query=("-kc", "4"), ok=True):
        print("[gpu] pm_pipeline -ck", *query, *extra, flush=True)
        out = subprocess.run([exe, "-f", path, "-m", "0.1", "-qs", "2", "-pn", "2", "-cn", "1", "-t", "4", *query, "-ck", ck] + list(extra),
                             capture_output=True, text=True, timeout=120)
        if not ok:
            assert out.returncode != 0
            return None, out.stdout + out.stderr
        assert out.returncode == 0, out.stderr[-2000:]
        line = [x for x in out.stdout.splitlines() if x.startswith("total count for query 0:")][0]
        return int(line.split(":")[1]), out.stdout

    total, _ = run()
    lines = open(ck).read().splitlines()
    assert total == want and len(lines) == 5 and lines[0].startswith("vsgm-checkpoint ")
    total, out = run()
    assert total == want and "4 of 4 bins already counted" in out
    # the same file under another query (5-cliques) is refused, untouched
    _, msg = run(query=("-kc", "5"), ok=False)
    assert "written for another run" in msg and open(ck).read().splitlines() == lines
    lines = open(ck).read().splitlines()
    bad = lines[3].split()
    bad[1] = str(int(bad[1]) ^ 1)  # a bin whose sources no longer match is counted again
    open(ck, "w").write("\n".join(lines[:3] + [" ".join(bad)]) + "\n")
    total, out = run("-dv", "1")
    assert total == want and "2 of 4 bins already counted" in out
    assert len(open(ck).read().splitlines()) == 6
    # a file without the header line (records of unknown origin) is refused
    open(ck, "w").write("\n".join(lines[1:]) + "\n")
    _, msg = run(ok=False)
    assert "no vsgm-checkpoint header" in msg
