"""GPU: the device CSR build (vsgm_csr_from_pairs) against the ingestion
checker (oracle/ingest_ref.py) and the host build — random edge lists with
loops, duplicates and id gaps, edge cases, an R-MAT-sized skewed list — and
the SNAP reader and preprocess_graph tool with -g 0 against the host path.
All integer work: results must be identical."""
import os
import subprocess
import time

import numpy as np
import pytest

from oracle import ingest_ref as I

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def ctx(vsgm, gpu):
    from vsgm_amd.hip import DeviceContext

    c = DeviceContext(0)
    yield c
    c.close()


def build(ctx, pairs, V, symmetric=True, drop=False):
    pr = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint32).reshape(-1))
    d = ctx.array(pr) if pr.size else ctx.empty(0, np.uint32)
    rp, cols = ctx.csr_from_pairs(d, V, symmetric, drop)
    out = rp.numpy(), cols.numpy()
    rp.free(), cols.free(), d.free()
    return out


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("mode", [(True, False), (True, True), (False, False)])
def test_random_lists_match_oracle(ctx, seed, mode):
    rng = np.random.default_rng(seed)
    V = int(rng.integers(1, 3000))
    m = int(rng.integers(0, 20000))
    pairs = rng.integers(0, V, size=(m, 2))
    pairs[: m // 10, 1] = pairs[: m // 10, 0]  # self loops
    pairs = np.concatenate([pairs, pairs[: m // 5]])  # duplicates
    print(f"[gpu] csr_from_pairs V={V} pairs={len(pairs)} mode={mode}", flush=True)
    got = build(ctx, pairs, V, *mode)
    want = I.csr_from_pairs(pairs, V, *mode)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def test_edge_cases(ctx, vsgm):
    print("[gpu] csr_from_pairs edge cases", flush=True)
    rp, cols = build(ctx, np.zeros((0, 2)), 5)
    assert rp.tolist() == [0] * 6 and cols.size == 0
    rp, cols = build(ctx, [(2, 2), (0, 0)], 3)
    assert rp.tolist() == [0] * 4 and cols.size == 0
    rp, cols = build(ctx, [(2, 2), (0, 0)], 3, True, True)
    assert rp.tolist() == [0] and cols.size == 0
    rp, cols = build(ctx, [(0, 0)], 1)
    assert rp.tolist() == [0, 0]
    rp, cols = build(ctx, [(0, 1)] * 5, 2, True, True)
    assert rp.tolist() == [0, 1, 2] and cols.tolist() == [1, 0]
    with pytest.raises(vsgm.VsgmError):
        build(ctx, [(0, 5)], 5)
    with pytest.raises(vsgm.VsgmError):
        build(ctx, [(0, 1)], 5, False, True)


def test_skewed_list_matches_host_build(ctx, vsgm):
    rng = np.random.default_rng(22)
    V = 1 << 22
    m = 8 << 22
    # one heavy-tailed endpoint (hub lists of 10^5..10^6 entries), one uniform
    pairs = np.empty((m, 2), np.uint32)
    pairs[:, 0] = np.minimum((rng.pareto(0.8, size=m) * 8).astype(np.int64), V - 1)
    pairs[:, 1] = rng.integers(0, V, size=m)
    print(f"[gpu] skewed list V={V} pairs={m}", flush=True)
    host = vsgm.Graph.from_edges(V, pairs)
    hrp, hcols = host.csr()
    d = ctx.array(pairs.reshape(-1))
    ctx.sync()
    t = time.perf_counter()
    rp, cols = ctx.csr_from_pairs(d, V, True, False)
    ms = (time.perf_counter() - t) * 1e3
    print(f"[gpu] csr_from_pairs {m} pairs -> {cols.n} entries in {ms:.1f} ms", flush=True)
    assert np.array_equal(rp.numpy(), hrp) and np.array_equal(cols.numpy(), hcols)
    rp.free(), cols.free()
    rp, cols = ctx.csr_from_pairs(d, V, True, True)
    drp, dcols = host.drop_isolated().csr()
    assert np.array_equal(rp.numpy(), drp) and np.array_equal(cols.numpy(), dcols)
    rp.free(), cols.free(), d.free()


def test_read_snap_on_device(vsgm, gpu, tmp_path):
    rng = np.random.default_rng(5)
    lines = ["# FromNodeId\tToNodeId"] + [f"{a + 100}\t{b + 100}" for a, b in rng.integers(0, 5000, size=(40000, 2))]
    path = tmp_path / "g.ungraph.txt"
    path.write_text("\n".join(lines) + "\n")
    print("[gpu] read_snap device 0", flush=True)
    dev = vsgm.Graph.read_snap(str(path), device=0).csr()
    host = vsgm.Graph.read_snap(str(path)).csr()
    assert np.array_equal(dev[0], host[0]) and np.array_equal(dev[1], host[1])
    exe = os.path.join(ROOT, "vsgm_amd", "bin", "preprocess_graph")
    out = subprocess.run([exe, "-f", str(path), "-d", "0", "-g", "0"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    got = vsgm.Graph.read_bin(str(path) + ".bin").csr()
    assert np.array_equal(got[0], host[0]) and np.array_equal(got[1], host[1])
