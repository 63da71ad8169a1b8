"""CPU restatement of graph ingestion (include/vsgm_hip.h vsgm_csr_from_pairs,
include/vsgm.h vsgm_graph_read_snap).

TEST INFRASTRUCTURE ONLY: the product runs vsgm_amd/csrc/hip/ingest.hip and
vsgm_amd/csrc/host/ingest.cc; this numpy version is the checker. It follows
GraphIO::ReadSNAPFile (src/graph/graph_io.h:193-274: lines not starting with
a digit skipped, ids shifted by the smallest id, both directions of each
undirected edge; ReadSNAPLargeVIDFile, :106-191, for large ids),
SortCSRArray, and PreprocessGraph::Preprocess
(src/preprocess/preprocess_graph.h:12-92: self loops and parallel edges
dropped, then vertices with empty lists removed and the rest renumbered in
order). Pinned by tests/test_oracle_ingest.py against a set-of-edges brute
force and closed forms (complete graphs, stars, id gaps).
"""
import numpy as np


def parse_snap(text: str):
    """(pairs u32[n][2] shifted by the smallest id, vertex_count)."""
    rows = []
    for line in text.split("\n"):
        if not line or not line[0].isdigit():
            continue
        a, b = line.split()[:2]
        rows.append((int(a), int(b)))
    if not rows:
        raise ValueError("no edge lines")
    arr = np.array(rows, dtype=np.int64)
    lo, hi = int(arr.min()), int(arr.max())
    return (arr - lo).astype(np.uint32), hi - lo + 1


def parse_snap_lvid(text: str):
    """(pairs u32[n][2], vertex_count) of a large-id SNAP text
    (graph_io.h:106-191): every digit run of a line that starts with a
    digit is an id, numbered by first appearance; consecutive ids pair."""
    import re

    ids, ends = {}, []
    for line in text.split("\n"):
        if not line or not line[0].isdigit():
            continue
        for tok in re.findall(r"[0-9]+", line):
            ends.append(ids.setdefault(tok, len(ids)))
    if not ends or len(ends) % 2:
        raise ValueError("no edges or an odd number of ids")
    return np.array(ends, np.uint32).reshape(-1, 2), len(ids)


def csr_from_pairs(pairs, vertex_count: int, symmetric: bool = True, drop_isolated: bool = False):
    """(row_ptrs u64[V'+1], cols u32[M]) of the cleaned, sorted CSR."""
    p = np.asarray(pairs, dtype=np.uint64).reshape(-1, 2)
    u, v = p[:, 0], p[:, 1]
    if p.size and (u.max() >= vertex_count or v.max() >= vertex_count):
        raise ValueError("vertex id out of range")
    if symmetric:
        u, v = np.concatenate([u, v]), np.concatenate([v, u])
    keep = u != v
    keys = np.unique((u[keep] << np.uint64(32)) | v[keep])
    rows = (keys >> np.uint64(32)).astype(np.int64)
    cols = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    deg = np.bincount(rows, minlength=vertex_count)[:vertex_count]
    if drop_isolated:
        if not symmetric:
            raise ValueError("isolated-vertex removal needs a symmetric graph")
        alive = deg > 0
        new_id = np.cumsum(alive) - 1
        cols = new_id[cols].astype(np.uint32)
        deg = deg[alive]
    rp = np.zeros(len(deg) + 1, np.uint64)
    rp[1:] = np.cumsum(deg)
    return rp, cols
