"""CPU: pinning the k-means checker (oracle/kmeans_ref.py) — cosine
similarity against the set formula, centroid sums against their closed
form, exact recovery of disjoint cliques, and a non-decreasing objective."""
import numpy as np

from oracle import kmeans_ref as K


def csr(n, edges):
    adj = [set() for _ in range(n)]
    for a, b in edges:
        if a != b:
            adj[a].add(b)
            adj[b].add(a)
    rp = np.zeros(n + 1, np.uint64)
    rp[1:] = np.cumsum([len(x) for x in adj])
    cols = np.array([y for x in adj for y in sorted(x)], np.uint32)
    return rp, cols, adj


def test_similarity_is_the_set_cosine():
    rng = np.random.default_rng(0)
    n = 25
    rp, cols, adj = csr(n, [tuple(rng.integers(0, n, 2)) for _ in range(60)])
    X = K.features(rp, cols)
    for u in (0, 3, 11):
        s = K.similarities(X, X[u:u + 1])[:, 0]
        Nu = adj[u] | {u}
        for v in range(n):
            Nv = adj[v] | {v}
            assert abs(s[v] - len(Nu & Nv) / np.sqrt(len(Nu) * len(Nv))) < 1e-12


def test_centroid_sums_closed_form():
    rng = np.random.default_rng(1)
    n = 30
    rp, cols, adj = csr(n, [tuple(rng.integers(0, n, 2)) for _ in range(80)])
    X = K.features(rp, cols)
    a = rng.integers(0, 3, n)
    C, sizes = K.update(X, a, 3)
    for k in range(3):
        members = np.nonzero(a == k)[0]
        # each feature sums to sqrt(|N[v]|): the centroid sums to their mean
        want = np.mean([np.sqrt(len(adj[v]) + 1) for v in members])
        assert abs(C[k].sum() - want) < 1e-12 and sizes[k] == len(members)


def test_disjoint_cliques_are_recovered():
    m, k = 6, 4
    edges = [(c * m + i, c * m + j) for c in range(k) for i in range(m) for j in range(i + 1, m)]
    rp, cols, _ = csr(m * k, edges)
    a, sizes, obj = K.kmeans(rp, cols, k, 3, first=7)
    assert sorted(sizes.tolist()) == [m] * k
    for c in range(k):
        assert len(set(a[c * m:(c + 1) * m].tolist())) == 1  # one cluster per clique
    assert abs(obj[-1] - m * k) < 1e-9  # every vertex identical to its centroid


def test_objective_does_not_decrease():
    rng = np.random.default_rng(2)
    n = 80
    rp, cols, _ = csr(n, [tuple(rng.integers(0, n, 2)) for _ in range(300)])
    _, _, obj = K.kmeans(rp, cols, 4, 6, first=5)
    assert all(b >= a - 1e-9 for a, b in zip(obj, obj[1:]))


def test_reference_self_test_pairs():
    """The reference's own k-means self-test (src/tools/kmeans.cc:697-771):
    on the 8 directed lists {2,3},{2,3},{4,5},{4,5},{6,7},{6,7},{0,1},{0,1},
    k = 4 with 10 rounds puts each pair (2i, 2i+1) in one cluster and every
    cluster holds 2 vertices, whatever the first centroid."""
    lists = [[2, 3], [2, 3], [4, 5], [4, 5], [6, 7], [6, 7], [0, 1], [0, 1]]
    rp = np.zeros(9, np.uint64)
    rp[1:] = np.cumsum([len(x) for x in lists])
    cols = np.array([y for x in lists for y in x], np.uint32)
    for first in range(8):
        a, sizes, _ = K.kmeans(rp, cols, 4, 10, first)
        assert all(a[i] == a[i + 1] for i in range(0, 8, 2)), (first, a)
        assert sorted(sizes.tolist()) == [2, 2, 2, 2]


def test_member_run_equals_whole_run_and_ignores_other_rows():
    rng = np.random.default_rng(3)
    n = 40
    rp, cols, adj = csr(n, [tuple(rng.integers(0, n, 2)) for _ in range(120)])
    a, sizes, obj = K.kmeans(rp, cols, 3, 4, 5)
    am, sm, om = K.kmeans_members(rp, cols, 3, 4, 5, np.arange(n))
    assert np.array_equal(a, am) and np.array_equal(sizes, sm) and np.allclose(obj, om)
    # members' features only read their own rows: rewriting the other rows changes nothing
    members = np.arange(0, n, 2)
    a1, s1, o1 = K.kmeans_members(rp, cols, 3, 4, 2, members)
    lists = [list(cols[rp[v]:rp[v + 1]]) for v in range(n)]
    for v in range(1, n, 2):
        lists[v] = sorted(set(rng.integers(0, n, 5).tolist()) - {v})
    rp2 = np.zeros(n + 1, np.uint64)
    rp2[1:] = np.cumsum([len(x) for x in lists])
    cols2 = np.array([y for x in lists for y in x], np.uint32)
    a2, s2, o2 = K.kmeans_members(rp2, cols2, 3, 4, 2, members)
    assert np.array_equal(a1, a2) and np.allclose(o1, o2)
    assert int(s1.sum()) == len(members)


def test_member_run_recovers_disjoint_cliques_among_members():
    """Six disjoint 5-cliques; the members are the vertices of cliques 1, 3
    and 4: k = 3 over them returns exactly those three cliques, whatever the
    first centroid."""
    m, n_cliques = 5, 6
    edges = [(c * m + i, c * m + j) for c in range(n_cliques) for i in range(m) for j in range(i + 1, m)]
    rp, cols, _ = csr(m * n_cliques, edges)
    members = np.array([c * m + i for c in (1, 3, 4) for i in range(m)])
    for first in range(len(members)):
        a, sizes, _ = K.kmeans_members(rp, cols, 3, 3, first, members)
        assert sorted(sizes.tolist()) == [m] * 3
        assert all(len(set(a[j * m:(j + 1) * m].tolist())) == 1 for j in range(3))
