"""Planner parity: vsgm_amd's planner against the REFERENCE planner.

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
from vsgm_amd.shapes import all_shapes


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


def test_hop_and_root_degree_of_benchmark_patterns(vsgm):
    want = {"clique 3": (2, 2), "clique 4": (2, 3), "preset 2": (2, 3), "preset 1": (2, 2), "preset 4": (2, 3)}
    for spec, (hop, rd) in want.items():
        p = vsgm.Plan.optimize(_query_for(vsgm, spec))
        assert (p.hop, p.root_degree) == (hop, rd), spec


def test_from_order_hop_covers_every_list_it_reads(vsgm, oracle):
    """Over view bins (each a ball of radius hop-1 around its sources) the
    per-bin counts depend on the root, but their sum must be the pattern's
    count for every matching order: an order that reads the list of a
    deeper vertex needs a larger hop than the breadth-first one."""
    import itertools

    g = vsgm.Graph.rmat(8, 16, seed=4)
    rp, cl = g.csr()
    path4 = np.zeros((4, 4), np.uint8)
    for a, b in ((0, 1), (1, 2), (2, 3)):
        path4[a, b] = path4[b, a] = 1
    for adj in (path4, [s[2] for s in all_shapes((5,))][7]):
        q = vsgm.Query.from_matrix(adj)
        n, _, a, _ = q.info()
        orders = [p for p in itertools.permutations(range(n))
                  if all(any(a[p[i], p[j]] for j in range(i)) for i in range(1, n))][::3]
        plans = [vsgm.Plan.from_order(q, o) for o in orders]
        want = oracle.count_plan(rp, cl, plans[0].serialize())
        hop = max(p.hop for p in plans)
        rd = min(p.root_degree for p in plans)
        assign = np.random.default_rng(1).integers(0, 3, size=g.vertex_count).astype(np.int32)
        vbm = vsgm.ViewBinManager.from_assignment(g, hop, 3, assign)
        h = vsgm.ViewBinHolder(g.vertex_count, g.entries * 4 + 4)
        tot = np.zeros(len(plans), np.int64)
        for b in range(vbm.n_bins):
            vbm.materialize(b, h, rd)
            vr, vc, src = h.csr(g.vertex_count)
            tot += [oracle.count_plan(vr, vc, p.serialize(), sources=src) for p in plans]
        assert tot.tolist() == [want] * len(plans)
    # on the 4-path some connected order reads the list of a vertex two steps
    # from its root (e.g. an end vertex matched before its neighbour), which
    # needs hop 3, one more than any breadth-first order of the path
    q = vsgm.Query.from_matrix(path4)
    n, _, a, _ = q.info()
    assert max(vsgm.Plan.from_order(q, o).hop for o in itertools.permutations(range(4))
               if all(any(a[o[i], o[j]] for j in range(i)) for i in range(1, 4))) == 3


def _plan_fields(text):
    """The serialized plan's fields as lists of bracket groups (strings)."""
    out = {}
    for line in text.splitlines():
        if ":" not in line:
            continue
        key, val = line.split(":", 1)
        if key == "exec":
            out[key] = re.findall(r"\((\w+),(\d+)\)", val)
        elif val.startswith("["):
            out[key] = re.findall(r"\[([^\]]*)\]", val)
    return out


def _signature(f, level):
    """Level signature restated from the plan text: operator, vertex,
    materialised and pending groups, and what the op applies (backward set
    and compute conditions, with the condition against the one pending set
    of a COMPUTE_COUNT; materialise or filter conditions; the condition
    between the two sets of a COUNT)."""
    op, u = f["exec"][level]
    u = int(u)
    cu = [int(x) for x in f["computed_unmaterialized"][level].split(",") if x]

    def cond(a, b):  # first entry of a's order list naming b, else NON_EQUAL
        for c in re.findall(r"(\w+)\((\d+)\)", f["order"][a]):
            if int(c[1]) == b:
                return c[0]
        return "NON_EQUAL"

    sig = [op, u, f["materialized"][level], f["computed_unmaterialized"][level]]
    if op in ("COMPUTE", "COMPUTE_COUNT", "COMPUTE_PATH_COUNT"):
        sig += [f["backward"][u], f["compute_cond"][u]]
        if op == "COMPUTE_COUNT" and len(cu) == 1:
            sig.append(cond(u, cu[0]))
    elif op == "MATERIALIZE":
        sig.append(f["materialize_cond"][u])
    elif op == "FILTER_COMPUTE":
        sig.append(f["filter_cond"][u])
    elif op == "COUNT" and len(cu) == 2:
        sig.append(cond(cu[0], cu[1]))
    return sig


def test_shared_prefix_matches_the_plan_text(vsgm):
    """Prefix-shared matching groups plans by their level signatures: the
    library's shared-prefix length equals the one restated from the plan
    texts for every pair of graphlet (3..5), preset and clique plans, and a
    plan shares all its levels with itself."""
    qs = [q for k in (3, 4, 5) for q in vsgm.Query.graphlets(k)]
    qs += [vsgm.Query.preset(k) for k in range(19)] + [vsgm.Query.clique(k) for k in range(3, 8)]
    plans = [vsgm.Plan.optimize(q) for q in qs]
    fields = [_plan_fields(p.serialize()) for p in plans]
    sigs = [[_signature(f, l) for l in range(len(f["exec"]))] for f in fields]
    shared_somewhere = 0
    for i, a in enumerate(plans):
        assert a.shared_prefix(a) == len(sigs[i])
        for j, b in enumerate(plans):
            n = 0
            while n < min(len(sigs[i]), len(sigs[j])) and sigs[i][n] == sigs[j][n]:
                n += 1
            assert a.shared_prefix(b) == n, (i, j)
            shared_somewhere += i != j and n >= 2
    assert shared_somewhere > 0  # some plans do share more than the root level
