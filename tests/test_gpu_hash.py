"""GPU parity of the map kernels that keep a run's stable adjacency list in
an LDS hash table (isect_kernel 3, and the flat map of isect_kernel 4)
against the oracle and set algebra:
every preset, clique and connected 3..6-vertex shape; tickets of 1..64
chunks and every pivot weight; batching; lists longer than the table (the
kernel falls back to the membership index / binary search for those); the
membership index off; and the per-path count operator."""
import ctypes as C

import numpy as np
import pytest

from vsgm_amd.shapes import all_shapes

pytestmark = pytest.mark.gpu

HASH_KERNELS = pytest.mark.parametrize("kernel", [3, 4])


def run_gpu(vsgm, graph, plans, batch_bytes=None, **options):
    ex = vsgm.PipelineExecutor(0, graph.vertex_count, plans)
    if batch_bytes:
        ex.set_option("batch_bytes", batch_bytes)
    for k, v in options.items():
        ex.set_option(k, v)
    ex.transfer_graph(graph)
    ex.match()
    return ex.total_counts


def oracle_counts(oracle, graph, plans):
    rp, cl = graph.csr()
    return [oracle.count_plan(rp, cl, p.serialize()) for p in plans]


@pytest.fixture(scope="module")
def rmat9(vsgm, gpu):
    return vsgm.Graph.rmat(9, 16, seed=21)


@HASH_KERNELS
def test_presets_cliques_every_ticket_and_weight(vsgm, oracle, rmat9, kernel):
    qs = [vsgm.Query.preset(k) for k in range(17)] + [vsgm.Query.clique(k) for k in range(3, 7)]
    plans = [vsgm.Plan.optimize(q) for q in qs]
    want = oracle_counts(oracle, rmat9, plans)
    for tc, cw in ((64, 2), (1, 1), (16, 1024), (3, 4)):
        print(f"[gpu] hash kernel {kernel} on R-MAT s9, ticket_chunks={tc} cold_weight={cw}", flush=True)
        assert run_gpu(vsgm, rmat9, plans, isect_kernel=kernel, ticket_chunks=tc, cold_weight=cw) == want, (tc, cw)


@HASH_KERNELS
def test_six_and_seven_vertex_presets(vsgm, oracle, gpu, kernel):
    g = vsgm.Graph.rmat(8, 8, seed=4)
    plans = [vsgm.Plan.optimize(q) for q in (vsgm.Query.preset(17), vsgm.Query.preset(18), vsgm.Query.clique(7))]
    print(f"[gpu] hash kernel {kernel}: presets 17, 18 and the 7-clique on R-MAT s8", flush=True)
    assert run_gpu(vsgm, g, plans, isect_kernel=kernel) == oracle_counts(oracle, g, plans)


@HASH_KERNELS
def test_all_connected_shapes(vsgm, oracle, gpu, kernel):
    g = vsgm.Graph.rmat(8, 8, seed=4)
    plans = [vsgm.Plan.optimize(vsgm.Query.from_matrix(a)) for _, _, a in all_shapes()]
    print(f"[gpu] hash kernel {kernel}: {len(plans)} shapes on R-MAT s8", flush=True)
    assert run_gpu(vsgm, g, plans, isect_kernel=kernel) == oracle_counts(oracle, g, plans)


@HASH_KERNELS
def test_batching_and_index_off(vsgm, oracle, rmat9, kernel):
    plans = [vsgm.Plan.optimize(q) for q in (vsgm.Query.clique(4), vsgm.Query.preset(1), vsgm.Query.preset(4),
                                             vsgm.Query.preset(2), vsgm.Query.preset(13))]
    want = oracle_counts(oracle, rmat9, plans)
    for opts in ({"batch_bytes": 1 << 15}, {"batch_bytes": 1 << 18}, {"hub_bytes": 0}, {"hub_bytes": 0, "batch_bytes": 1 << 16}):
        print(f"[gpu] hash kernel {kernel} {opts}", flush=True)
        assert run_gpu(vsgm, rmat9, plans, isect_kernel=kernel, **opts) == want, opts


@HASH_KERNELS
def test_lists_longer_than_the_table(vsgm, oracle, gpu, kernel):
    """Two vertices adjacent to 15,000 others each (longer than the
    14,336-entry table; 5,000 common neighbours) plus a sparse random graph:
    4-cycles, 4-cliques and diamonds whose stable lists do and do not fit,
    against the textbook counters."""
    rng = np.random.default_rng(5)
    n = 26_000
    edges = [(0, j) for j in range(2, 15_002)] + [(1, j) for j in range(10_000, 25_000)]
    edges += [tuple(e) for e in rng.integers(2, n, size=(60_000, 2)) if e[0] != e[1]]
    g = vsgm.Graph.from_edges(n, edges)
    assert g.max_degree == 15_000
    rp, cl = g.csr()
    plans = [vsgm.Plan.optimize(q) for q in (vsgm.Query.preset(1), vsgm.Query.clique(4), vsgm.Query.preset(2))]
    want = [oracle.four_cycles(rp, cl), oracle.four_cliques(rp, cl), oracle.diamonds(rp, cl)]
    print(f"[gpu] hash kernel {kernel}, lists above the table size: want {want}", flush=True)
    assert run_gpu(vsgm, g, plans, isect_kernel=kernel) == want
    assert run_gpu(vsgm, g, plans, isect_kernel=kernel, hub_bytes=0) == want


@HASH_KERNELS
def test_vertex_ids_above_24_bits(vsgm, oracle, gpu, kernel):
    """Vertex ids that agree in their low 24 bits (the flat map's first
    table hash sees only those): 6,144 core vertices k * 2^24 + j (k < 3,
    j < 2048) in a graph of 33,556,480 vertices, random edges among them and
    one vertex adjacent to every other core vertex; 4-cliques, diamonds and
    4-cycles against the textbook counters (4-cycles: the per-lane binary
    search kernel and the oracle's plan restatement)."""
    rng = np.random.default_rng(9)
    core = np.array([k * (1 << 24) + j for k in range(3) for j in range(2048)], np.int64)
    n = int(core.max()) + 1
    pairs = rng.integers(0, len(core), size=(60_000, 2))
    edges = [(int(core[a]), int(core[b])) for a, b in pairs.tolist() if a != b]
    edges += [(int(core[0]), int(v)) for v in core[1:]]
    g = vsgm.Graph.from_edges(n, edges)
    rp, cl = g.csr()
    plans = [vsgm.Plan.optimize(q) for q in (vsgm.Query.clique(4), vsgm.Query.preset(2), vsgm.Query.preset(1))]
    want = [oracle.four_cliques(rp, cl), oracle.diamonds(rp, cl), oracle.count_plan(rp, cl, plans[2].serialize(), sources=core.astype(np.uint32))]
    print(f"[gpu] hash kernel {kernel}, ids above 2^24: want {want}", flush=True)
    assert run_gpu(vsgm, g, plans, isect_kernel=kernel) == want
    assert run_gpu(vsgm, g, plans, isect_kernel=kernel, cold_weight=64) == want


@HASH_KERNELS
def test_views_with_self_loops(vsgm, oracle, rmat9, kernel):
    """A view with self-loops (every third vertex of R-MAT s9 adjacent to
    itself): the '!=' conditions against backward vertices are no longer
    implied and must be tested; counts against the oracle's plan
    restatement, and the device self-loop count against numpy."""
    from vsgm_amd.hip import DeviceContext, hiplib

    rp, cl = rmat9.csr()
    n = len(rp) - 1
    rows = [list(cl[rp[v]:rp[v + 1]]) for v in range(n)]
    loops = [v for v in range(0, n, 3) if rows[v]]
    for v in loops:
        rows[v] = sorted(rows[v] + [v])
    rp2 = np.zeros(n + 1, np.uint64)
    rp2[1:] = np.cumsum([len(r) for r in rows])
    cl2 = np.array([x for r in rows for x in r], np.uint32)
    g = vsgm.Graph.from_csr(rp2, cl2)
    ctx = DeviceContext(0)
    d_rp, d_cl = ctx.array(rp2), ctx.array(cl2)
    cnt = C.c_uint64(0)
    ctx.check(hiplib().vsgm_view_self_loops(ctx.h, d_rp.ptr, d_cl.ptr, n, C.byref(cnt)), "self loops")
    assert cnt.value == len(loops)
    ctx.close()
    plans = [vsgm.Plan.optimize(q) for q in (vsgm.Query.preset(1), vsgm.Query.clique(4), vsgm.Query.preset(2), vsgm.Query.preset(16))]
    want = [oracle.count_plan(rp2, cl2, p.serialize()) for p in plans]
    print(f"[gpu] hash kernel {kernel}, view with {len(loops)} self-loops: want {want}", flush=True)
    assert run_gpu(vsgm, g, plans, isect_kernel=kernel) == want
    assert run_gpu(vsgm, g, plans, isect_kernel=kernel, cold_weight=64, ticket_chunks=4) == want


def test_per_path_counts(vsgm, gpu):
    """COMPUTE_PATH_COUNT through the hash kernel: for edge paths (v0, v1)
    with v1 > v0, |N(v0) ∩ N(v1) ∩ (> v1)| per path, equal to set algebra
    and to isect_kernel 0."""
    from vsgm_amd.hip import DeviceContext, LevelArgs, hiplib

    rng = np.random.default_rng(17)
    n = 4000
    e = rng.integers(0, n, size=(60_000, 2))
    e[:6000, 0] = rng.integers(0, 8, size=6000)
    e = e[e[:, 0] != e[:, 1]]
    adj = [set() for _ in range(n)]
    for a, b in e.tolist():
        adj[a].add(b)
        adj[b].add(a)
    rp = np.zeros(n + 1, np.uint64)
    rp[1:] = np.cumsum([len(x) for x in adj])
    cols = np.array([v for x in adj for v in sorted(x)], np.uint32)
    v0 = np.repeat(np.arange(n, dtype=np.uint32), np.diff(rp).astype(np.int64))
    v1 = cols.copy()
    keep = v1 > v0
    v0, v1 = v0[keep], v1[keep]
    P = len(v0)
    want = np.array([len({x for x in adj[a] & adj[b] if x > b}) for a, b in zip(v0.tolist(), v1.tolist())], np.uint64)
    ctx = DeviceContext(0)
    d_rp, d_cols, d_v0, d_v1 = ctx.array(rp), ctx.array(cols), ctx.array(v0), ctx.array(v1)
    a = LevelArgs()
    a.row_ptrs, a.cols, a.vertex_count, a.path_num = d_rp.ptr, d_cols.ptr, n, P
    a.other = -1
    a.inst[0], a.inst[1] = d_v0.ptr, d_v1.ptr
    a.target, a.n_back = 2, 2
    a.back[0], a.back[1] = 0, 1
    a.n_cond = 1
    a.cond[0].op, a.cond[0].operand = 1, 1  # > v1
    for kernel, tc in ((0, 16), (3, 1), (3, 16), (4, 1), (4, 16)):
        ctx.tuning(isect_kernel=kernel, ticket_chunks=tc)
        cnt = ctx.empty(P, np.uint64)
        ctx.check(hiplib().vsgm_fill_u64(ctx.h, cnt.ptr, P, 0), "fill")
        ctx.check(hiplib().vsgm_intersect_count_paths(ctx.h, C.byref(a), cnt.ptr), "count_paths")
        assert np.array_equal(cnt.numpy(), want), kernel
        tot = ctx.array(np.zeros(1, np.uint64))
        ctx.check(hiplib().vsgm_intersect_count_total(ctx.h, C.byref(a), tot.ptr), "count_total")
        assert int(tot.numpy()[0]) == int(want.sum()), kernel
    ctx.close()


def test_tuning_bounds(vsgm, gpu):
    from vsgm_amd.hip import DeviceContext

    ctx = DeviceContext(0)
    assert ctx.tuning() == {"isect_kernel": 4, "max_grid": 8192, "cold_weight": 16, "ticket_chunks": 64}
    for bad in ({"isect_kernel": 5}, {"ticket_chunks": 0}, {"ticket_chunks": 65}, {"cold_weight": 0}):
        with pytest.raises(Exception):
            ctx.tuning(**bad)
    ctx.close()
