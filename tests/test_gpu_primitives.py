"""GPU: device primitives and level operators through the C-ABI of
include/vsgm_hip.h, against numpy / brute force on the same inputs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(gpu):
    from vsgm_amd.hip import DeviceContext

    c = DeviceContext(0)
    yield c
    c.close()


@pytest.mark.parametrize("n", [0, 1, 2, 255, 2047, 2048, 2049, 100003, 5_000_011])
def test_scan_exclusive(ctx, n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 1 << 20, size=n, dtype=np.uint64)
    out = ctx.scan(ctx.array(a)).numpy()
    want = np.concatenate([[0], np.cumsum(a, dtype=np.uint64)]).astype(np.uint64)
    assert np.array_equal(out, want)


def test_scan_in_place_and_large_values(ctx):
    a = np.full(70_000, (1 << 40) + 7, np.uint64)
    d = ctx.empty(a.size + 1, np.uint64)
    from vsgm_amd.hip import hiplib

    ctx.check(hiplib().vsgm_h2d(ctx.h, d.ptr, a.ctypes.data, a.nbytes), "h2d")
    ctx.check(hiplib().vsgm_scan_excl_u64(ctx.h, d.ptr, d.ptr, a.size), "scan")
    out = d.numpy()
    assert out[-1] == a.sum(dtype=np.uint64)  # exceeds 2^32
    assert np.array_equal(out[:-1], np.concatenate([[0], np.cumsum(a, dtype=np.uint64)[:-1]]))


@pytest.mark.parametrize("n", [0, 1, 777, 1 << 20])
def test_reduce_gather_iota_fill(ctx, n):
    rng = np.random.default_rng(7 + n)
    a = rng.integers(0, 1 << 30, size=n, dtype=np.uint64)
    assert ctx.reduce(ctx.array(a)) == int(a.sum(dtype=np.uint64))
    if n:
        idx = rng.integers(0, n, size=n // 2 + 1, dtype=np.uint64)
        src32 = rng.integers(0, 1 << 31, size=n, dtype=np.uint32)
        assert np.array_equal(ctx.gather(ctx.array(idx), ctx.array(src32)).numpy(), src32[idx])
        assert np.array_equal(ctx.gather(ctx.array(idx), ctx.array(a)).numpy(), a[idx])
    assert np.array_equal(ctx.iota(n, 5).numpy(), np.arange(5, 5 + n, dtype=np.uint64))
    assert np.array_equal(ctx.fill(n, 9).numpy(), np.full(n, 9, np.uint64))


def test_split_batch_partitions_parents(ctx):
    rng = np.random.default_rng(3)
    children = rng.integers(0, 1000, size=50_000, dtype=np.uint64)
    children[123] = 10**7  # one parent larger than the budget on its own
    pf, cf, budget = 16, 5, 1 << 20
    ends = ctx.split_batch(ctx.array(children), pf, cf, budget)
    assert ends[-1] == children.size and np.all(np.diff(ends.astype(np.int64)) > 0) and ends[0] > 0
    cost = children * cf + pf
    starts = np.concatenate([[0], ends[:-1]]).astype(np.int64)
    for s, e in zip(starts, ends.astype(np.int64)):
        c = int(cost[s:e].sum())
        # a batch stays within the budget plus one parent
        assert c <= budget + int(cost[s:e].max())
    assert len(ends) >= int(cost.sum()) // budget


def _csr_from_edges(n, edges):
    adj = [[] for _ in range(n)]
    for a, b in edges:
        if a != b:
            adj[a].append(b)
            adj[b].append(a)
    adj = [sorted(set(x)) for x in adj]
    rp = np.zeros(n + 1, np.uint64)
    rp[1:] = np.cumsum([len(x) for x in adj])
    cols = np.array([y for x in adj for y in x], np.uint32)
    return rp, cols, adj


def _args(ctx, d_rp, d_cols, n, paths):
    from vsgm_amd.hip import LevelArgs

    a = LevelArgs()
    a.row_ptrs, a.cols, a.vertex_count, a.path_num = d_rp.ptr, d_cols.ptr, n, paths
    a.other = -1
    return a


def test_intersect_write_expand_filter_count(ctx):
    """Triangle-style levels on a random graph, checked against set algebra:
    COMPUTE u2 = N(v0) ∩ N(v1) ∩ (> v1) over edge paths, then FILTER,
    MATERIALIZE and the count ops on the result."""
    import ctypes as C

    from vsgm_amd.hip import CsrOut, ExpandOut, hiplib

    rng = np.random.default_rng(11)
    n = 3000
    edges = rng.integers(0, n, size=(40_000, 2))
    edges[:2000, 0] = rng.integers(0, 20, size=2000)  # a few hubs
    rp, cols, adj = _csr_from_edges(n, edges)
    sets = [set(x) for x in adj]
    v0 = np.repeat(np.arange(n, dtype=np.uint32), np.diff(rp).astype(np.int64))
    v1 = cols.copy()
    keep = v1 > v0
    v0, v1 = v0[keep], v1[keep]
    P = len(v0)
    d_rp, d_cols, d_v0, d_v1 = ctx.array(rp), ctx.array(cols), ctx.array(v0), ctx.array(v1)

    a = _args(ctx, d_rp, d_cols, n, P)
    a.inst[0], a.inst[1] = d_v0.ptr, d_v1.ptr
    a.target, a.n_back = 2, 2
    a.back[0], a.back[1] = 0, 1
    a.n_cond = 2
    a.cond[0].op, a.cond[0].operand = 2, 0  # != v0
    a.cond[1].op, a.cond[1].operand = 1, 1  # > v1
    out = CsrOut()
    ctx.check(hiplib().vsgm_intersect_write(ctx.h, C.byref(a), C.byref(out)), "intersect_write")
    offs = ctx.adopt(out.offsets, P + 1, np.uint64).numpy()
    vals = ctx.adopt(out.values, out.count, np.uint32).numpy()
    want_rows = [sorted(x for x in sets[a0] & sets[a1] if x > a1) for a0, a1 in zip(v0.tolist(), v1.tolist())]
    want_off = np.concatenate([[0], np.cumsum([len(r) for r in want_rows])]).astype(np.uint64)
    assert np.array_equal(offs, want_off)
    assert np.array_equal(vals, np.array([x for r in want_rows for x in r], np.uint32))

    # per-path counts and the reduced total
    cnt = ctx.empty(P, np.uint64)
    ctx.check(hiplib().vsgm_intersect_count_paths(ctx.h, C.byref(a), cnt.ptr), "count_paths")
    assert np.array_equal(cnt.numpy(), np.diff(want_off))
    tot = ctx.array(np.zeros(1, np.uint64))
    ctx.check(hiplib().vsgm_intersect_count_total(ctx.h, C.byref(a), tot.ptr), "count_total")
    assert int(tot.numpy()[0]) == int(want_off[-1])

    # FILTER the candidate rows against a random bound column: x < r3, x != r4
    d_off, d_val = ctx.array(offs), ctx.array(vals)
    r3 = rng.integers(0, n, size=P, dtype=np.uint32)
    r4 = np.array([r[len(r) // 2] if r else 0 for r in want_rows], np.uint32)
    f = _args(ctx, d_rp, d_cols, n, P)
    d_r3, d_r4 = ctx.array(r3), ctx.array(r4)
    f.inst[3], f.inst[4] = d_r3.ptr, d_r4.ptr
    f.cand[2], f.cand_off[2] = d_val.ptr, d_off.ptr
    f.target = 2
    f.n_cond = 2
    f.cond[0].op, f.cond[0].operand = 0, 3  # candidate < r3
    f.cond[1].op, f.cond[1].operand = 2, 4  # candidate != r4
    fo = CsrOut()
    ctx.check(hiplib().vsgm_filter(ctx.h, C.byref(f), C.byref(fo)), "filter")
    f_rows = [[x for x in r if x < b3 and x != b4] for r, b3, b4 in zip(want_rows, r3.tolist(), r4.tolist())]
    assert 0 < sum(map(len, f_rows)) < int(want_off[-1])
    f_off = ctx.adopt(fo.offsets, P + 1, np.uint64).numpy()
    assert np.array_equal(f_off, np.concatenate([[0], np.cumsum([len(r) for r in f_rows])]).astype(np.uint64))
    assert np.array_equal(ctx.adopt(fo.values, fo.count, np.uint32).numpy(), np.array([x for r in f_rows for x in r], np.uint32))

    # MATERIALIZE target 2 with condition "!= v0", carrying inst[0], inst[1]
    e = _args(ctx, d_rp, d_cols, n, P)
    e.inst[0], e.inst[1] = d_v0.ptr, d_v1.ptr
    e.cand[2], e.cand_off[2] = d_val.ptr, d_off.ptr
    e.target = 2
    e.n_cond = 1
    e.cond[0].op, e.cond[0].operand = 2, 0
    mat = np.array([0, 1], np.uint32)
    eo = ExpandOut()
    ctx.check(hiplib().vsgm_expand(ctx.h, C.byref(e), mat.ctypes.data, 2, None, 0, C.byref(eo)), "expand")
    k = eo.path_num
    got = np.stack([ctx.adopt(eo.inst[u], k, np.uint32).numpy() for u in (0, 1, 2)], axis=1)
    want = np.array([(a0, a1, x) for a0, a1, r in zip(v0.tolist(), v1.tolist(), want_rows) for x in r], np.uint32).reshape(-1, 3)
    assert np.array_equal(got, want)

    # COUNT of one set, and of two sets with "x < y"
    c1 = _args(ctx, d_rp, d_cols, n, P)
    c1.cand[2], c1.cand_off[2], c1.target = d_val.ptr, d_off.ptr, 2
    tot = ctx.array(np.zeros(1, np.uint64))
    ctx.check(hiplib().vsgm_count_sets(ctx.h, C.byref(c1), tot.ptr), "count1")
    assert int(tot.numpy()[0]) == int(want_off[-1])
    c2 = _args(ctx, d_rp, d_cols, n, P)
    c2.cand[2], c2.cand_off[2] = d_val.ptr, d_off.ptr
    c2.cand[3], c2.cand_off[3] = d_val.ptr, d_off.ptr
    c2.target, c2.other, c2.other_op = 2, 3, 0
    tot = ctx.array(np.zeros(1, np.uint64))
    ctx.check(hiplib().vsgm_count_sets(ctx.h, C.byref(c2), tot.ptr), "count2")
    assert int(tot.numpy()[0]) == sum(len(r) * (len(r) - 1) // 2 for r in want_rows)


def test_hub_index_bits_and_indexed_intersections(ctx):
    """The dense membership index holds exactly each indexed list, and
    intersections that probe it give the same rows as binary search."""
    import ctypes as C

    from vsgm_amd.hip import CsrOut, DeviceArray, hiplib

    rng = np.random.default_rng(21)
    n = 5000
    edges = rng.integers(0, n, size=(60_000, 2))
    edges[:6000, 0] = rng.integers(0, 30, size=6000)
    rp, cols, adj = _csr_from_edges(n, edges)
    sets = [set(x) for x in adj]
    deg = np.diff(rp).astype(np.int64)
    hubs = np.nonzero(deg >= 40)[0].astype(np.uint32)
    rng.shuffle(hubs)  # slot order is the caller's order
    assert 20 < hubs.size < n // 4
    d_rp, d_cols = ctx.array(rp), ctx.array(cols)
    idx = ctx.hub_index(d_rp, d_cols, n, hubs)
    try:
        assert idx.n_hubs == hubs.size and idx.words % 2 == 0 and idx.words * 32 >= n
        slot = DeviceArray(ctx, idx.slot, n, np.uint32, owned=False).numpy()
        want_slot = np.full(n, 0xFFFFFFFF, np.uint32)
        want_slot[hubs] = np.arange(hubs.size, dtype=np.uint32)
        assert np.array_equal(slot, want_slot)
        bits = DeviceArray(ctx, idx.bits, hubs.size * idx.words, np.uint32, owned=False).numpy().reshape(hubs.size, idx.words)
        for s, v in enumerate(hubs.tolist()):
            members = np.nonzero(np.unpackbits(bits[s].view(np.uint8), bitorder="little"))[0]
            assert members.tolist() == adj[v], v

        # wedge paths (v0, v1, v2) with v1 < v2 in N(v0); u3 in N(v0) ∩ N(v1) ∩ N(v2), u3 > v0
        v0s, v1s, v2s = [], [], []
        for v in range(0, n, 7):
            nb = adj[v]
            for i in range(min(len(nb), 12)):
                for j in range(i + 1, min(len(nb), 12)):
                    v0s.append(v), v1s.append(nb[i]), v2s.append(nb[j])
        P = len(v0s)
        cols_d = [ctx.array(np.array(x, np.uint32)) for x in (v0s, v1s, v2s)]
        rows = [sorted(x for x in sets[a] & sets[b] & sets[c] if x > a) for a, b, c in zip(v0s, v1s, v2s)]
        want_off = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.uint64)
        assert want_off[-1] > 0
        for use_index in (False, True):
            a = _args(ctx, d_rp, d_cols, n, P)
            for u in range(3):
                a.inst[u] = cols_d[u].ptr
            a.target, a.n_back = 3, 3
            a.back[0], a.back[1], a.back[2] = 0, 1, 2
            a.n_cond = 1
            a.cond[0].op, a.cond[0].operand = 1, 0  # > v0
            if use_index:
                a.hub_slot, a.hub_bits, a.hub_words = idx.slot, idx.bits, idx.words
            out = CsrOut()
            ctx.check(hiplib().vsgm_intersect_write(ctx.h, C.byref(a), C.byref(out)), "intersect_write")
            assert np.array_equal(ctx.adopt(out.offsets, P + 1, np.uint64).numpy(), want_off)
            assert np.array_equal(ctx.adopt(out.values, out.count, np.uint32).numpy(), np.array([x for r in rows for x in r], np.uint32))
            tot = ctx.array(np.zeros(1, np.uint64))
            ctx.check(hiplib().vsgm_intersect_count_total(ctx.h, C.byref(a), tot.ptr), "count_total")
            assert int(tot.numpy()[0]) == int(want_off[-1])
    finally:
        ctx.hub_index_free(idx)


def test_hub_index_rejects_bad_lists(ctx):
    from vsgm_amd.hip import VsgmError

    rp, cols, _ = _csr_from_edges(4, [(0, 1), (1, 2)])
    d_rp, d_cols = ctx.array(rp), ctx.array(cols)
    for bad in ([1, 1], [4]):
        with pytest.raises(VsgmError):
            ctx.hub_index(d_rp, d_cols, 4, np.array(bad, np.uint32))
    empty = ctx.hub_index(d_rp, d_cols, 4, np.zeros(0, np.uint32))
    assert empty.n_hubs == 0
    ctx.hub_index_free(empty)


def test_level_args_are_validated(ctx):
    import ctypes as C

    from vsgm_amd.hip import CsrOut, hiplib

    a = _args(ctx, ctx.array(np.zeros(2, np.uint64)), ctx.array(np.zeros(1, np.uint32)), 1, 1)
    a.target, a.n_back = 1, 1
    a.back[0] = 0  # inst[0] is NULL: must be rejected, not dereferenced
    out = CsrOut()
    assert hiplib().vsgm_intersect_write(ctx.h, C.byref(a), C.byref(out)) != 0
    assert b"not materialized" in hiplib().vsgm_last_error(ctx.h)
