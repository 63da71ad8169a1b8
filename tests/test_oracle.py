"""CPU: pinning the oracle.

Every oracle function is checked against something that does not share its
code: closed forms on complete / bipartite / grid / Petersen / wheel graphs,
brute-force automorphism counts of textbook shapes, the textbook counters
against brute force on random graphs, and the plan-driven restatement of
the level algorithm against brute force for every preset and every
connected 3..5-vertex shape (and a sample of 6-vertex shapes). A mistake
planted into a plan (a dropped symmetry condition) must change the count.
"""
import math

import numpy as np
import pytest

from vsgm_amd.shapes import all_shapes


def csr(n, edges):
    adj = [set() for _ in range(n)]
    for a, b in edges:
        if a != b:
            adj[a].add(b)
            adj[b].add(a)
    rp = np.zeros(n + 1, np.uint64)
    rp[1:] = np.cumsum([len(x) for x in adj])
    cols = np.array([y for x in adj for y in sorted(x)], np.uint32)
    return rp, cols


def complete(n):
    return csr(n, [(i, j) for i in range(n) for j in range(i + 1, n)])


def bipartite(m, n):
    return csr(m + n, [(i, m + j) for i in range(m) for j in range(n)])


def grid(r, c):
    e = [(i * c + j, i * c + j + 1) for i in range(r) for j in range(c - 1)]
    e += [(i * c + j, (i + 1) * c + j) for i in range(r - 1) for j in range(c)]
    return csr(r * c, e)


def petersen():
    return csr(10, [(i, (i + 1) % 5) for i in range(5)] + [(5 + i, 5 + (i + 2) % 5) for i in range(5)] + [(i, i + 5) for i in range(5)])


def wheel(k):
    return csr(k + 1, [(0, i) for i in range(1, k + 1)] + [(i, i % k + 1) for i in range(1, k + 1)])


def random_graph(n, p, seed):
    rng = np.random.default_rng(seed)
    return csr(n, [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < p])


def mat(n, edges):
    a = np.zeros((n, n), np.uint8)
    for i, j in edges:
        a[i, j] = a[j, i] = 1
    return a


TRI = mat(3, [(0, 1), (1, 2), (0, 2)])
SQUARE = mat(4, [(0, 1), (1, 2), (2, 3), (3, 0)])
DIAMOND = mat(4, [(0, 1), (0, 2), (0, 3), (1, 2), (2, 3)])
K4 = mat(4, [(i, j) for i in range(4) for j in range(i + 1, 4)])
HOUSE = mat(5, [(0, 1), (0, 2), (1, 2), (1, 3), (2, 4), (3, 4)])
C5 = mat(5, [(i, (i + 1) % 5) for i in range(5)])


def test_automorphism_counts(oracle):
    assert oracle.automorphisms(TRI) == 6
    assert oracle.automorphisms(SQUARE) == 8
    assert oracle.automorphisms(DIAMOND) == 4
    assert oracle.automorphisms(K4) == 24
    assert oracle.automorphisms(HOUSE) == 2
    assert oracle.automorphisms(C5) == 10
    star = mat(5, [(0, i) for i in range(1, 5)])
    assert oracle.automorphisms(star) == 24
    path = mat(4, [(0, 1), (1, 2), (2, 3)])
    assert oracle.automorphisms(path) == 2


@pytest.mark.parametrize("n", [4, 6, 7])
def test_brute_force_on_complete_graphs(oracle, n):
    rp, cl = complete(n)
    for pat in (TRI, SQUARE, DIAMOND, K4, HOUSE, C5):
        k = pat.shape[0]
        want = math.factorial(n) // math.factorial(n - k) // oracle.automorphisms(pat) if n >= k else 0
        assert oracle.count_brute(rp, cl, pat) == want


def test_brute_force_closed_forms(oracle):
    rp, cl = bipartite(5, 6)
    assert oracle.count_brute(rp, cl, TRI) == 0
    assert oracle.count_brute(rp, cl, SQUARE) == math.comb(5, 2) * math.comb(6, 2)
    rp, cl = grid(5, 7)
    assert oracle.count_brute(rp, cl, SQUARE) == 4 * 6
    rp, cl = petersen()
    assert oracle.count_brute(rp, cl, TRI) == 0
    assert oracle.count_brute(rp, cl, SQUARE) == 0
    assert oracle.count_brute(rp, cl, C5) == 12
    rp, cl = wheel(9)
    assert oracle.count_brute(rp, cl, TRI) == 9


@pytest.mark.parametrize("seed", range(4))
def test_textbook_counters_against_brute_force(oracle, seed):
    rp, cl = random_graph(24, 0.35, seed)
    assert oracle.triangles(rp, cl) == oracle.count_brute(rp, cl, TRI)
    assert oracle.four_cliques(rp, cl) == oracle.count_brute(rp, cl, K4)
    assert oracle.four_cliques_ranked(rp, cl) == oracle.count_brute(rp, cl, K4)
    assert oracle.diamonds(rp, cl) == oracle.count_brute(rp, cl, DIAMOND)
    assert oracle.four_cycles(rp, cl) == oracle.count_brute(rp, cl, SQUARE)


@pytest.mark.parametrize("scale", [8, 11, 13])
def test_ranked_four_cliques_equal_the_id_ordered_counter(vsgm, oracle, scale):
    # skewed R-MAT graphs, where degree order and id order differ most
    rp, cl = vsgm.Graph.rmat(scale, 16, seed=scale).csr()
    assert oracle.four_cliques_ranked(rp, cl) == oracle.four_cliques(rp, cl)


def test_textbook_counters_closed_forms(oracle):
    rp, cl = complete(9)
    assert oracle.triangles(rp, cl) == math.comb(9, 3)
    assert oracle.four_cliques(rp, cl) == math.comb(9, 4)
    assert oracle.four_cliques_ranked(rp, cl) == math.comb(9, 4)
    assert oracle.diamonds(rp, cl) == 6 * math.comb(9, 4)
    assert oracle.four_cycles(rp, cl) == 3 * math.comb(9, 4)


@pytest.mark.parametrize("seed", range(3))
def test_plan_restatement_presets_vs_brute(vsgm, oracle, seed):
    rp, cl = random_graph(18, 0.4, 100 + seed)
    for k in range(19):
        q = vsgm.Query.preset(k)
        n, _, adj, _ = q.info()
        got = oracle.count_plan(rp, cl, vsgm.Plan.optimize(q).serialize())
        assert got == oracle.count_brute(rp, cl, adj), f"P{k}"


def test_plan_restatement_all_shapes_vs_brute(vsgm, oracle):
    rp, cl = random_graph(16, 0.45, 7)
    shapes = all_shapes((3, 4, 5)) + all_shapes((6,))[::7]
    for n, mask, adj in shapes:
        q = vsgm.Query.from_matrix(adj)
        got = oracle.count_plan(rp, cl, vsgm.Plan.optimize(q).serialize())
        assert got == oracle.count_brute(rp, cl, adj), (n, mask)


def test_plan_restatement_with_sources(vsgm, oracle):
    # counts split over any partition of the roots sum to the total
    rp, cl = random_graph(30, 0.3, 3)
    txt = vsgm.Plan.optimize(vsgm.Query.preset(4)).serialize()
    total = oracle.count_plan(rp, cl, txt)
    parts = [list(range(0, 30, 3)), list(range(1, 30, 3)), list(range(2, 30, 3))]
    assert sum(oracle.count_plan(rp, cl, txt, sources=s) for s in parts) == total
    assert oracle.count_plan(rp, cl, txt, sources=[]) == 0


def test_planted_mistake_is_detected(vsgm, oracle):
    # dropping a symmetry condition from the triangle plan counts every
    # triangle more than once; the oracle check must see it
    rp, cl = complete(6)
    txt = vsgm.Plan.optimize(vsgm.Query.clique(3)).serialize()
    good = oracle.count_plan(rp, cl, txt)
    broken = txt.replace("compute_cond:[][LARGER_THAN(0)]", "compute_cond:[][NON_EQUAL(0)]")
    assert broken != txt
    assert good == math.comb(6, 3)
    assert oracle.count_plan(rp, cl, broken) != good


def test_level_paths_are_reported(vsgm, oracle):
    rp, cl = complete(5)
    txt = vsgm.Plan.optimize(vsgm.Query.clique(3)).serialize()
    total, paths = oracle.count_plan(rp, cl, txt, level_paths=True)
    assert total == 10
    assert paths[0] == 5  # one partial embedding per root
    assert paths[3] == 10  # (v0 < v1) edge paths reaching the last level


def test_plan_restatement_equals_golden_counts_at_scale_16(vsgm, oracle):
    """The level-by-level restatement (heuristic plans) against the golden
    textbook counts of the R-MAT scale-16 graph (tests/golden/rmat_counts.json,
    graph identified by its CSR digest), on the graph as generated for the
    triangles and on its degree-order relabelling for the triangle, 4-clique
    and diamond (the 4-cycle's plan takes minutes on the host here; its
    golden counts are checked on the GPU)."""
    import hashlib
    import json
    import os

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rmat_counts.json")
    rec = json.load(open(path))["rmat16"]
    g = vsgm.Graph.rmat(16, rec["edge_factor"], *rec["abc"], seed=rec["seed"])
    rp, cl = g.csr()
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(rp, dtype=np.uint64).tobytes())
    h.update(np.ascontiguousarray(cl, dtype=np.uint32).tobytes())
    assert h.hexdigest() == rec["csr_sha256"]
    qs = {"kc3": vsgm.Query.clique(3), "kc4": vsgm.Query.clique(4), "p2": vsgm.Query.preset(2)}
    text = {k: vsgm.Plan.optimize(q).serialize() for k, q in qs.items()}
    assert oracle.count_plan(rp, cl, text["kc3"]) == rec["counts"]["kc3"]
    hr, hc = g.relabel_by_degree()[0].csr()
    for k in qs:
        assert oracle.count_plan(hr, hc, text[k]) == rec["counts"][k], k
