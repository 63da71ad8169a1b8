"""The allocation-free COMPUTE_COUNT entry point
(vsgm_intersect_count_total_noalloc) through the C-ABI: equal to
vsgm_intersect_count_total (itself equal to set algebra here) on two calls
in a row, the second with a smaller input, for both map kernels; it leaves
the context's pool untouched, and sizes beyond its limit are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _level(ctx, n, m, seed):
    """Edge paths (v0, v1), v1 > v0, of a random graph with a few hubs; the
    level counts |N(v0) ∩ N(v1) ∩ (> v1)| (triangles) plus the device arrays
    that must stay alive."""
    from vsgm_amd.hip import LevelArgs

    rng = np.random.default_rng(seed)
    e = rng.integers(0, n, size=(m, 2))
    e[: m // 10, 0] = rng.integers(0, 8, size=m // 10)
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
    want = sum(len({x for x in adj[a] & adj[b] if x > b}) for a, b in zip(v0.tolist(), v1.tolist()))
    keepalive = [ctx.array(rp), ctx.array(cols), ctx.array(v0), ctx.array(v1)]
    a = LevelArgs()
    a.row_ptrs, a.cols, a.vertex_count, a.path_num = keepalive[0].ptr, keepalive[1].ptr, n, len(v0)
    a.other = -1
    a.inst[0], a.inst[1] = keepalive[2].ptr, keepalive[3].ptr
    a.target, a.n_back = 2, 2
    a.back[0], a.back[1] = 0, 1
    a.n_cond = 1
    a.cond[0].op, a.cond[0].operand = 1, 1  # > v1
    return a, want, keepalive


def _count(ctx, fn, a, tot):
    from vsgm_amd.hip import hiplib

    ctx.check(hiplib().vsgm_fill_u64(ctx.h, tot.ptr, 1, 0), "fill")
    ctx.check(getattr(hiplib(), fn)(ctx.h, C.byref(a), tot.ptr), fn)
    return int(tot.numpy()[0])


def test_noalloc_matches_the_allocating_count(gpu):
    from vsgm_amd.hip import DeviceContext

    ctx = DeviceContext(0)
    big, want_big, keep_big = _level(ctx, 6000, 90_000, 3)
    small, want_small, keep_small = _level(ctx, 900, 8_000, 4)
    assert small.path_num < big.path_num
    tot = ctx.empty(1, np.uint64)
    for kernel in (0, 3, 4):
        ctx.tuning(isect_kernel=kernel)
        print(f"[gpu] noalloc isect_kernel={kernel}: paths {big.path_num} then {small.path_num}", flush=True)
        assert _count(ctx, "vsgm_intersect_count_total", big, tot) == want_big
        assert _count(ctx, "vsgm_intersect_count_total", small, tot) == want_small
        live0 = ctx.mem_info()
        assert _count(ctx, "vsgm_intersect_count_total_noalloc", big, tot) == want_big
        assert _count(ctx, "vsgm_intersect_count_total_noalloc", small, tot) == want_small
        assert _count(ctx, "vsgm_intersect_count_total_noalloc", big, tot) == want_big
        # the pool saw no allocation: available bytes and the high-water mark are unchanged
        assert ctx.mem_info() == live0
    ctx.close()


def test_noalloc_refuses_sizes_beyond_its_limit(gpu):
    from vsgm_amd.hip import DeviceContext, hiplib

    ctx = DeviceContext(0)
    a, _, keep = _level(ctx, 300, 2_000, 5)
    a.path_num = (1 << 21) + 1  # checked before anything reads the columns
    tot = ctx.array(np.zeros(1, np.uint64))
    assert hiplib().vsgm_intersect_count_total_noalloc(ctx.h, C.byref(a), tot.ptr) == 4  # VSGM_ERR_ARG
    assert b"allocation-free limit" in hiplib().vsgm_last_error(ctx.h)
    assert int(tot.numpy()[0]) == 0
    ctx.close()
