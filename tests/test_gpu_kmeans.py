"""GPU: the k-means partitioner's kernels against the CPU checker
(oracle/kmeans_ref.py) — one assignment pass, one centroid update, whole
runs — and the view sizes of k-means bins against random bins.

Float64 sums are accumulated in different orders on the two sides, so
similarities and centroids agree to ~1e-12 relative; an assignment may
differ only where two clusters are within that of each other."""
import numpy as np
import pytest

from oracle import kmeans_ref as K

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(vsgm, gpu):
    from vsgm_amd.partition import DeviceGraph

    g = vsgm.Graph.rmat(8, 8, seed=3)
    rp, cl = g.csr()
    return g, rp, cl, DeviceGraph(g)


def test_assign_pass(small):
    from vsgm_amd.partition import assign_step

    g, rp, cl, dg = small
    X = K.features(rp, cl)
    rng = np.random.default_rng(0)
    C, _ = K.update(X, rng.integers(0, 5, g.vertex_count), 5)
    norm = np.sqrt((C * C).sum(1))
    a, s = assign_step(dg, C.T.copy(), norm)
    S = K.similarities(X, C)
    want_a, want_s = K.assign(X, C)
    assert np.allclose(s, want_s, rtol=1e-12, atol=1e-14)
    # the chosen cluster is a maximum; it is the reference's wherever the top two differ
    assert np.all(S[np.arange(len(a)), a] >= want_s - 1e-12)
    top2 = np.sort(S, axis=1)[:, -2:]
    clear = top2[:, 1] - top2[:, 0] > 1e-9
    assert np.array_equal(a[clear], want_a[clear])


def test_update_pass(small):
    from vsgm_amd.partition import update_step

    g, rp, cl, dg = small
    X = K.features(rp, cl)
    a = np.random.default_rng(1).integers(0, 4, g.vertex_count)
    a[:10] = 3
    cent, norm, sizes = update_step(dg, a.astype(np.int32), 4)
    C, want_sizes = K.update(X, a, 4)
    assert np.array_equal(sizes, want_sizes.astype(np.uint64))
    assert np.allclose(cent, C.T, rtol=1e-12, atol=1e-15)
    assert np.allclose(norm, np.sqrt((C * C).sum(1)), rtol=1e-12)


def test_whole_run_matches_checker(vsgm, small):
    from vsgm_amd.partition import first_vertex, kmeans

    g, rp, cl, _ = small
    for k, seed in ((2, 1), (5, 7), (8, 3)):
        a, sizes, obj = kmeans(g, k, iters=4, seed=seed)
        want_a, want_sizes, want_obj = K.kmeans(rp, cl, k, 4, first=first_vertex(seed, g.vertex_count))
        assert int(sizes.sum()) == g.vertex_count and np.array_equal(np.bincount(a, minlength=k), sizes)
        assert np.allclose(obj, want_obj, rtol=1e-9), (k, obj, want_obj)
        assert np.mean(a == want_a) > 0.99, k


def test_disjoint_cliques(vsgm, gpu):
    from vsgm_amd.partition import kmeans

    m, k = 7, 5
    g = vsgm.Graph.from_edges(m * k, [(c * m + i, c * m + j) for c in range(k) for i in range(m) for j in range(i + 1, m)])
    a, sizes, obj = kmeans(g, k, iters=3, seed=11)
    assert sorted(sizes.tolist()) == [m] * k
    assert all(len(set(a[c * m:(c + 1) * m].tolist())) == 1 for c in range(k))


def test_kmeans_bins_make_smaller_views(vsgm, gpu):
    """The point of the partition: 2-hop views of k-means bins hold fewer
    list entries than views of random bins of the same sizes."""
    from vsgm_amd.partition import kmeans

    g = vsgm.Graph.rmat(14, 16, seed=2)
    k = 8
    a, sizes, obj = kmeans(g, k, iters=10, seed=1)
    assert all(y >= x - 1e-6 * abs(x) for x, y in zip(obj, obj[1:]))
    rand = np.random.default_rng(0).permutation(np.repeat(np.arange(k), sizes.astype(np.int64)))

    def view_entries(assign):
        vbm = vsgm.ViewBinManager.from_assignment(g, 2, k, assign.astype(np.int32))
        h = vsgm.ViewBinHolder(g.vertex_count, g.entries * 4 + 4)
        tot = 0
        for b in range(vbm.n_bins):
            vbm.materialize(b, h, 0)
            tot += h.info()["entries"]
        return tot

    km, rnd = view_entries(a), view_entries(rand)
    print(f"[gpu] R-MAT s14, k=8: view entries k-means {km} vs random {rnd} ({km / rnd:.3f}); objective {obj[0]:.1f} -> {obj[-1]:.1f}",
          flush=True)
    assert km < rnd


def test_reference_self_test_pairs(vsgm, gpu):
    """The reference's own k-means self-test (src/tools/kmeans.cc:697-771) on
    the GPU: each pair (2i, 2i+1) of the 8 directed lists lands in one
    cluster of 2, for every seed."""
    from vsgm_amd.partition import kmeans

    lists = [[2, 3], [2, 3], [4, 5], [4, 5], [6, 7], [6, 7], [0, 1], [0, 1]]
    rp = np.zeros(9, np.uint64)
    rp[1:] = np.cumsum([len(x) for x in lists])
    g = vsgm.Graph.from_csr(rp, np.array([y for x in lists for y in x], np.uint32))
    for seed in range(1, 9):
        a, sizes, _ = kmeans(g, 4, iters=10, seed=seed)
        assert all(a[i] == a[i + 1] for i in range(0, 8, 2)), (seed, a)
        assert sorted(sizes.tolist()) == [2, 2, 2, 2]


def test_member_run_matches_checker(vsgm, small):
    """The member-subset run (each level of multi-level clustering) against
    the checker, and with every vertex a member, equal to the whole run."""
    from vsgm_amd.partition import first_vertex, kmeans, kmeans_members

    g, rp, cl, _ = small
    V = g.vertex_count
    a_all, s_all, o_all = kmeans_members(g, np.arange(V), 4, iters=3, seed=5)
    a, s, o = kmeans(g, 4, iters=3, seed=5)
    # float64 atomics sum in any order: objectives agree to rounding, assignments up to near-ties
    assert np.allclose(o_all, o, rtol=1e-9) and np.mean(a_all == a) >= 0.97
    members = np.nonzero(np.random.default_rng(2).random(V) < 0.4)[0]
    for k, seed in ((2, 1), (5, 9)):
        got_a, got_s, got_o = kmeans_members(g, members, k, iters=4, seed=seed)
        want_a, want_s, want_o = K.kmeans_members(rp, cl, k, 4, first_vertex(seed, len(members)), members)
        assert int(got_s.sum()) == len(members)
        assert np.allclose(got_o, want_o, rtol=1e-9), (k, got_o, want_o)
        assert np.mean(got_a == want_a) >= 0.97, k  # ~100 members: a near-tie or two may go either way
    with pytest.raises(vsgm.VsgmError):
        kmeans_members(g, np.array([0, V]), 1)


def test_multi_level_nests_and_keeps_cliques_whole(vsgm, gpu):
    """"3x2" on six disjoint 5-cliques: level 2 splits level 1's clusters
    (id // 2 is the level-1 cluster), no clique is ever split, and "3"
    equals the single-level run. "2^3" (recursive bisection) on eight
    cliques keeps every clique whole and numbers its clusters 0..n-1."""
    from vsgm_amd.partition import kmeans

    m, n = 5, 6
    g = vsgm.Graph.from_edges(m * n, [(c * m + i, c * m + j) for c in range(n) for i in range(m) for j in range(i + 1, m)])
    one, k1 = vsgm.kmeans_levels(g, "3", iters=3, seed=4)
    assert k1 == 3 and np.array_equal(one, kmeans(g, 3, iters=3, seed=4)[0])
    two, k2 = vsgm.kmeans_levels(g, "3x2", iters=3, seed=4)
    assert k2 == 6 and np.array_equal(two // 2, one)
    assert all(len(set(two[c * m:(c + 1) * m].tolist())) == 1 for c in range(n))
    m, n = 4, 8
    g8 = vsgm.Graph.from_edges(m * n, [(c * m + i, c * m + j) for c in range(n) for i in range(m) for j in range(i + 1, m)])
    b, nb = vsgm.kmeans_levels(g8, "2^3", iters=3, seed=2)
    assert 2 <= nb <= 8 and sorted(set(b.tolist())) == list(range(nb))
    assert all(len(set(b[c * m:(c + 1) * m].tolist())) == 1 for c in range(n))
    with pytest.raises(vsgm.VsgmError):
        vsgm.kmeans_levels(g, "2^0")


def test_symmetric_graphs_partition_the_same_every_run(vsgm, gpu):
    """On a symmetric graph the update gathers in a fixed order and the
    norms fold in wave order: two runs with one seed give the same partition
    bit for bit (whole runs and member runs)."""
    from vsgm_amd.partition import kmeans, kmeans_members

    g = vsgm.Graph.rmat(12, 8, seed=9)
    a1, s1, _ = kmeans(g, 6, iters=5, seed=3)
    a2, s2, _ = kmeans(g, 6, iters=5, seed=3)
    assert np.array_equal(a1, a2) and np.array_equal(s1, s2)
    members = np.arange(0, g.vertex_count, 3)
    m1, _, _ = kmeans_members(g, members, 4, iters=5, seed=8)
    m2, _, _ = kmeans_members(g, members, 4, iters=5, seed=8)
    assert np.array_equal(m1, m2)
