"""ctypes bindings of oracle/liboracle.so (test infrastructure only)."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

MAXV = 8
MAXOPS = 32
_OPS = {"COMPUTE": 0, "COMPUTE_PATH_COUNT": 1, "MATERIALIZE": 2, "FILTER_COMPUTE": 3, "COUNT": 4, "COMPUTE_COUNT": 5}
_CONDS = {"LESS_THAN": 0, "LARGER_THAN": 1, "NON_EQUAL": 2}


class CondList(C.Structure):
    _fields_ = [("n", C.c_int), ("op", C.c_int * MAXV), ("operand", C.c_int * MAXV)]


class OraclePlan(C.Structure):
    _fields_ = [
        ("n_vertices", C.c_int),
        ("n_ops", C.c_int),
        ("op", C.c_int * MAXOPS),
        ("op_vertex", C.c_int * MAXOPS),
        ("back_n", C.c_int * MAXV),
        ("back", (C.c_int * MAXV) * MAXV),
        ("compute_cond", CondList * MAXV),
        ("materialize_cond", CondList * MAXV),
        ("filter_cond", CondList * MAXV),
        ("cu_n", C.c_int * MAXOPS),
        ("cu", (C.c_int * MAXV) * MAXOPS),
        ("order", (C.c_int * MAXV) * MAXV),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        u64 = C.c_uint64
        for name in ("oracle_triangles", "oracle_four_cliques", "oracle_four_cliques_ranked", "oracle_diamonds", "oracle_four_cycles"):
            f = getattr(L, name)
            f.argtypes = [u64, P, P, C.c_int]
            f.restype = u64
        L.oracle_automorphisms.argtypes = [C.c_int, P]
        L.oracle_automorphisms.restype = u64
        L.oracle_count_brute.argtypes = [u64, P, P, C.c_int, P]
        L.oracle_count_brute.restype = u64
        L.oracle_count_plan.argtypes = [u64, P, P, P, u64, C.POINTER(OraclePlan), P, C.c_int]
        L.oracle_count_plan.restype = u64
        _lib = L
    return _lib


def _csr(row_ptrs, cols):
    rp = np.ascontiguousarray(row_ptrs, dtype=np.uint64)
    cl = np.ascontiguousarray(cols, dtype=np.uint32)
    if cl.size == 0:
        cl = np.zeros(1, np.uint32)
    return rp, cl


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def automorphisms(adj) -> int:
    a = np.ascontiguousarray(np.asarray(adj, dtype=np.uint8))
    return int(_load().oracle_automorphisms(a.shape[0], _p(a)))


def count_brute(row_ptrs, cols, adj) -> int:
    rp, cl = _csr(row_ptrs, cols)
    a = np.ascontiguousarray(np.asarray(adj, dtype=np.uint8))
    return int(_load().oracle_count_brute(len(rp) - 1, _p(rp), _p(cl), a.shape[0], _p(a)))


def _textbook(name, row_ptrs, cols, threads=0):
    rp, cl = _csr(row_ptrs, cols)
    return int(getattr(_load(), name)(len(rp) - 1, _p(rp), _p(cl), threads))


def triangles(row_ptrs, cols, threads=0):
    return _textbook("oracle_triangles", row_ptrs, cols, threads)


def four_cliques(row_ptrs, cols, threads=0):
    return _textbook("oracle_four_cliques", row_ptrs, cols, threads)


def four_cliques_ranked(row_ptrs, cols, threads=0):
    return _textbook("oracle_four_cliques_ranked", row_ptrs, cols, threads)


def diamonds(row_ptrs, cols, threads=0):
    return _textbook("oracle_diamonds", row_ptrs, cols, threads)


def four_cycles(row_ptrs, cols, threads=0):
    return _textbook("oracle_four_cycles", row_ptrs, cols, threads)


def _groups(text):
    return [[int(x) for x in g.split(",") if x] for g in re.findall(r"\[([^\]]*)\]", text)]


def _conds(text):
    out = []
    for g in re.findall(r"\[([^\]]*)\]", text):
        out.append([(_CONDS[m.group(1)], int(m.group(2))) for m in re.finditer(r"(\w+)\((\d+)\)", g)])
    return out


def parse_plan(serialized: str) -> OraclePlan:
    """Plan (in the key/value text of Plan::Serialize) -> oracle_plan."""
    kv = {}
    for line in serialized.strip().splitlines():
        k, _, v = line.partition(":")
        kv[k.strip()] = v.strip()
    p = OraclePlan()
    conn = _groups(kv["conn"])
    n = len(conn)
    p.n_vertices = n
    ops = re.findall(r"\((\w+),(\d+)\)", kv["exec"])
    p.n_ops = len(ops)
    for i, (op, v) in enumerate(ops):
        p.op[i] = _OPS[op]
        p.op_vertex[i] = int(v)
    for u, g in enumerate(_groups(kv["backward"])):
        p.back_n[u] = len(g)
        for j, x in enumerate(g):
            p.back[u][j] = x
    for key, field in (("compute_cond", "compute_cond"), ("materialize_cond", "materialize_cond"), ("filter_cond", "filter_cond")):
        for u, cl in enumerate(_conds(kv[key])):
            lst = getattr(p, field)[u]
            lst.n = len(cl)
            for j, (op, x) in enumerate(cl):
                lst.op[j] = op
                lst.operand[j] = x
    for lvl, g in enumerate(_groups(kv["computed_unmaterialized"])):
        p.cu_n[lvl] = len(g)
        for j, x in enumerate(g):
            p.cu[lvl][j] = x
    # operator between a and b: the first entry of a's list naming b, else NON_EQUAL
    order = _conds(kv["order"])
    for a in range(n):
        for b in range(n):
            ops_ab = [op for op, x in (order[a] if a < len(order) else []) if x == b]
            p.order[a][b] = ops_ab[0] if ops_ab else _CONDS["NON_EQUAL"]
    return p


def count_plan(row_ptrs, cols, serialized_plan: str, sources=None, threads=0, level_paths=False):
    rp, cl = _csr(row_ptrs, cols)
    plan = parse_plan(serialized_plan)
    lp = np.zeros(MAXOPS, np.uint64)
    if sources is None:
        src_p, n_src = None, 0
    else:
        s = np.ascontiguousarray(np.asarray(sources, dtype=np.uint32))
        if s.size == 0:
            return (0, [0] * plan.n_ops) if level_paths else 0
        src_p, n_src = _p(s), len(s)
    r = int(_load().oracle_count_plan(len(rp) - 1, _p(rp), _p(cl), src_p, n_src, C.byref(plan), _p(lp), threads))
    if r == 2**64 - 1:
        raise ValueError("malformed plan")
    if level_paths:
        return r, [int(x) for x in lp[: plan.n_ops]]
    return r
