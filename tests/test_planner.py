"""Planner parity: the build's planner against the REFERENCE planner.

tests/golden/ref_plans.txt holds the reference's own plans (remapped
adjacency, conditions, hop, root degree, exec sequence, per-level vertex
groups, backward connectivity and the compute / materialize / filter
conditions) for the 19 presets, 3..7-cliques, every 3..5-graphlet and every
connected 3..6-vertex shape, dumped by oracle/_ref/ref_plan_dump built from
the reference sources (tests/golden/make_plan_fixtures.py). Symmetry
breaking is further pinned by counting on complete graphs: each pattern
must be found exactly n!/((n-k)! |Aut|) times.
"""
import math
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN
from shapes import all_shapes


def _records():
    recs = []
    cur = None
    with open(os.path.join(GOLDEN, "ref_plans.txt")) as f:
        for line in f:
            if line.startswith("query: "):
                cur = [line[len("query: "):].strip(), []]
                recs.append(cur)
            else:
                cur[1].append(line.rstrip("\n"))
    return [(spec, "\n".join(body) + "\n") for spec, body in recs]


RECORDS = _records()


def _query_for(vsgm, spec, cache={}):
    parts = spec.split()
    if parts[0] == "preset":
        return vsgm.Query.preset(int(parts[1]))
    if parts[0] == "clique":
        return vsgm.Query.clique(int(parts[1]))
    if parts[0] == "graphlets":
        k = int(parts[1])
        if k not in cache:
            cache[k] = vsgm.Query.graphlets(k)
        return cache[k][int(parts[2].lstrip("#"))]
    if parts[0] == "adj":
        n = int(parts[1])
        a = np.array([int(c) for c in parts[2]], np.uint8).reshape(n, n)
        return vsgm.Query.from_matrix(a)
    raise ValueError(spec)


def test_fixture_inventory():
    kinds = [s.split()[0] for s, _ in RECORDS]
    assert kinds.count("preset") == 19
    assert kinds.count("clique") == 5
    assert kinds.count("graphlets") == 2 + 6 + 21
    assert kinds.count("adj") == 141


@pytest.mark.parametrize("spec,expected", RECORDS, ids=[s for s, _ in RECORDS])
def test_plan_matches_reference(vsgm, spec, expected):
    q = _query_for(vsgm, spec)
    p = vsgm.Plan.optimize(q)
    assert p.serialize() == expected


def _count_on_complete(oracle, plan_text, n):
    rp = np.arange(0, n * (n - 1) + 1, n - 1, dtype=np.uint64)
    cols = np.array([j for i in range(n) for j in range(n) if j != i], np.uint32)
    return oracle.count_plan(rp, cols, plan_text)


@pytest.mark.parametrize("n,mask,adj", all_shapes((3, 4, 5)), ids=lambda x: str(x) if isinstance(x, int) else "")
def test_symmetry_breaking_counts_each_instance_once(vsgm, oracle, n, mask, adj):
    q = vsgm.Query.from_matrix(adj)
    p = vsgm.Plan.optimize(q)
    aut = oracle.automorphisms(adj)
    K = 7
    expect = math.factorial(K) // math.factorial(K - n) // aut
    assert _count_on_complete(oracle, p.serialize(), K) == expect


def test_preset_counts_on_complete_graph(vsgm, oracle):
    K = 8
    for k in range(19):
        q = vsgm.Query.preset(k)
        n, _, adj, _ = q.info()
        expect = math.factorial(K) // math.factorial(K - n) // oracle.automorphisms(adj)
        assert _count_on_complete(oracle, vsgm.Plan.optimize(q).serialize(), K) == expect, f"P{k}"


def test_from_order_plans_count_the_same(vsgm, oracle):
    # every connected matching order gives the same count (the invariant the
    # learned planner relies on)
    g = vsgm.Graph.rmat(9, 8, seed=5)
    rp, cl = g.csr()
    for q in (vsgm.Query.preset(4), vsgm.Query.preset(1), vsgm.Query.clique(4), vsgm.Query.preset(2)):
        base = oracle.count_plan(rp, cl, vsgm.Plan.optimize(q).serialize())
        n = q.vertex_count
        adj = q.adjacency
        import itertools

        tried = 0
        for perm in itertools.permutations(range(n)):
            ok = all(any(adj[perm[i], perm[j]] for j in range(i)) for i in range(1, n))
            if not ok:
                continue
            p = vsgm.Plan.from_order(q, perm)
            assert p.order == list(perm)
            assert oracle.count_plan(rp, cl, p.serialize()) == base, perm
            tried += 1
        assert tried > 1


def test_from_order_rejects_disconnected_order(vsgm):
    q = vsgm.Query.preset(1)  # 4-cycle: 0-1, 0-2, 1-3, 2-3 after remap
    adj = q.adjacency
    bad = None
    import itertools

    for perm in itertools.permutations(range(4)):
        if not adj[perm[0], perm[1]]:
            bad = perm
            break
    with pytest.raises(vsgm.VsgmError):
        vsgm.Plan.from_order(q, bad)


def test_hop_and_root_degree_of_baseline_patterns(vsgm):
    want = {"clique 3": (2, 2), "clique 4": (2, 3), "preset 2": (2, 3), "preset 1": (2, 2), "preset 4": (2, 3)}
    for spec, (hop, rd) in want.items():
        p = vsgm.Plan.optimize(_query_for(vsgm, spec))
        assert (p.hop, p.root_degree) == (hop, rd), spec
