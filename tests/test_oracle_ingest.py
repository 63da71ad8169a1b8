"""CPU: pinning the ingestion checker (oracle/ingest_ref.py) against a
set-of-edges brute force and closed forms, then the host SNAP reader,
isolated-vertex removal and the preprocess_graph tool against it."""
import os
import subprocess

import numpy as np
import pytest

from oracle import ingest_ref as I

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def brute(pairs, n, drop_isolated):
    adj = {u: set() for u in range(n)}
    for a, b in pairs:
        if a != b:
            adj[int(a)].add(int(b))
            adj[int(b)].add(int(a))
    keep = [u for u in range(n) if adj[u]] if drop_isolated else list(range(n))
    new = {u: i for i, u in enumerate(keep)}
    rp, cols = [0], []
    for u in keep:
        cols += sorted(new[w] for w in adj[u])
        rp.append(len(cols))
    return np.array(rp, np.uint64), np.array(cols, np.uint32)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("drop", [False, True])
def test_oracle_matches_brute_force(seed, drop):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 40))
    m = int(rng.integers(0, 120))
    pairs = rng.integers(0, n, size=(m, 2))
    rp, cols = I.csr_from_pairs(pairs, n, True, drop)
    brp, bcols = brute(pairs, n, drop)
    assert np.array_equal(rp, brp) and np.array_equal(cols, bcols)


def test_oracle_closed_forms():
    n = 7
    # every ordered pair twice plus loops: the complete graph K_7
    pairs = [(a, b) for a in range(n) for b in range(n)] * 2
    rp, cols = I.csr_from_pairs(pairs, n)
    assert rp[-1] == n * (n - 1) and np.all(np.diff(rp) == n - 1)
    # a star on ids 10, 20, ..., 50 with centre 30 inside 60 ids: 5 vertices remain
    star = [(30, x) for x in (10, 20, 40, 50)] + [(30, 30)]
    rp, cols = I.csr_from_pairs(star, 60, True, True)
    assert len(rp) == 6 and list(np.diff(rp)) == [1, 1, 4, 1, 1]
    assert list(cols[rp[2]:rp[3]]) == [0, 1, 3, 4]
    # only loops: no entries
    rp, cols = I.csr_from_pairs([(3, 3), (1, 1)], 5)
    assert rp[-1] == 0 and len(rp) == 6


def test_oracle_parse_snap():
    text = "# comment\n% other\n\n5\t7\n7 9 1.5\n12\t5\n"
    pairs, V = I.parse_snap(text)
    assert V == 8 and pairs.tolist() == [[0, 2], [2, 4], [7, 0]]


def snap_text(rng, n, m, offset):
    lines = ["# Directed graph (each unordered pair of nodes is saved once)", "# FromNodeId\tToNodeId"]
    for a, b in rng.integers(0, n, size=(m, 2)):
        lines.append(f"{a + offset}\t{b + offset}")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_host_read_snap_matches_oracle(vsgm, tmp_path, threads):
    rng = np.random.default_rng(threads)
    text = snap_text(rng, 300, 900, 17) + "5\t5\n"  # a loop below the smallest id shifts every id
    path = tmp_path / "g.txt"
    path.write_text(text)
    g = vsgm.Graph.read_snap(str(path), threads=threads)
    pairs, V = I.parse_snap(text)
    rp, cols = I.csr_from_pairs(pairs, V, True, True)
    grp, gcols = g.csr()
    assert np.array_equal(grp, rp) and np.array_equal(gcols, cols)
    assert g.validate() == ""


def test_host_read_snap_errors(vsgm, tmp_path):
    bad = tmp_path / "bad.txt"
    bad.write_text("1 2\n3\n")
    with pytest.raises(vsgm.VsgmError):
        vsgm.Graph.read_snap(str(bad))
    big = tmp_path / "big.txt"
    big.write_text("1 4294967296\n")
    with pytest.raises(vsgm.VsgmError):
        vsgm.Graph.read_snap(str(big))
    empty = tmp_path / "empty.txt"
    empty.write_text("# nothing\n")
    with pytest.raises(vsgm.VsgmError):
        vsgm.Graph.read_snap(str(empty))
    ok = tmp_path / "ok.txt"
    ok.write_text("1 2\n")
    with pytest.raises(vsgm.VsgmError):
        vsgm.Graph.read_snap(str(ok), directed=True)


def test_drop_isolated_matches_oracle(vsgm):
    rng = np.random.default_rng(4)
    pairs = rng.integers(0, 500, size=(300, 2)).astype(np.uint32)
    g = vsgm.Graph.from_edges(500, pairs).drop_isolated()
    rp, cols = I.csr_from_pairs(pairs, 500, True, True)
    grp, gcols = g.csr()
    assert np.array_equal(grp, rp) and np.array_equal(gcols, cols)


def test_preprocess_graph_cli(vsgm, tmp_path):
    rng = np.random.default_rng(9)
    text = snap_text(rng, 200, 700, 3)
    path = tmp_path / "g.ungraph.txt"
    path.write_text(text)
    exe = os.path.join(ROOT, "vsgm_amd", "bin", "preprocess_graph")
    out = subprocess.run([exe, "-f", str(path), "-d", "0", "-o", "1"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    g = vsgm.Graph.read_bin(str(path) + ".bin")
    pairs, V = I.parse_snap(text)
    rp, cols = I.csr_from_pairs(pairs, V, True, True)
    grp, gcols = g.csr()
    assert np.array_equal(grp, rp) and np.array_equal(gcols, cols)
    assert subprocess.run([exe, "-f", str(path), "-o", "7"], capture_output=True, timeout=60).returncode != 0
    # -rl 1: the same graph numbered in ascending (degree, id) order
    out = subprocess.run([exe, "-f", str(path), "-d", "0", "-o", "1", "-rl", "1"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    h = vsgm.Graph.read_bin(str(path) + ".bin")
    want, _ = vsgm.Graph.from_csr(rp, cols).relabel_by_degree()
    assert all(np.array_equal(x, y) for x, y in zip(h.csr(), want.csr()))


def test_host_read_snap_tiny_file_many_threads(vsgm, tmp_path):
    # fewer bytes than threads: every chunk cut but the last is empty
    path = tmp_path / "tiny.txt"
    path.write_text("1 2\n")
    for threads in (1, 8, 64):
        rp, cols = vsgm.Graph.read_snap(str(path), threads=threads).csr()
        assert rp.tolist() == [0, 1, 2] and cols.tolist() == [1, 0]
    crlf = tmp_path / "crlf.txt"
    crlf.write_bytes(b"# c\r\n3\t4\r\n4 5 extra\r\n")
    rp, cols = vsgm.Graph.read_snap(str(crlf), threads=4).csr()
    assert rp.tolist() == [0, 1, 3, 4] and cols.tolist() == [1, 0, 2, 1]


def test_drop_isolated_rejects_asymmetric_csr(vsgm):
    g = vsgm.Graph.from_csr(np.array([0, 1, 1, 1], np.uint64), np.array([2], np.uint32))
    with pytest.raises(vsgm.VsgmError):
        g.drop_isolated()


def test_oracle_parse_snap_lvid():
    text = "# web graph\n100 2000000000000000000000\n2000000000000000000000\t7\n7,100\n007 7\n"
    pairs, V = I.parse_snap_lvid(text)
    # ids by first appearance: 100 -> 0, 2e21 -> 1, 7 -> 2, "007" -> 3 (a different string)
    assert V == 4 and pairs.tolist() == [[0, 1], [1, 2], [2, 0], [3, 2]]
    with pytest.raises(ValueError):
        I.parse_snap_lvid("1 2 3\n")


def test_load_by_suffix(vsgm, tmp_path):
    rng = np.random.default_rng(11)
    big = [f"{10**19 + int(a) * 7919}\t{10**19 + int(b) * 7919}" for a, b in rng.integers(0, 400, size=(1500, 2))]
    lvid = tmp_path / "g.lvid"
    lvid.write_text("# ids above 2^64\n" + "\n".join(big) + "\n")
    pairs, V = I.parse_snap_lvid(lvid.read_text())
    rp, cols = vsgm.Graph.load(str(lvid)).csr()
    want = I.csr_from_pairs(pairs, V, True, False)
    assert np.array_equal(rp, want[0]) and np.array_equal(cols, want[1])
    # .txt keeps isolated vertices unless asked (ids must match a .vbmap made for the file)
    txt = tmp_path / "g.txt"
    txt.write_text("5 9\n9 12\n12 12\n")
    rp, cols = vsgm.Graph.load(str(txt)).csr()
    assert len(rp) == 9 and rp[-1] == 4 and cols.tolist() == [4, 0, 7, 4]
    rp, cols = vsgm.Graph.load(str(txt), drop_isolated=True).csr()
    assert rp.tolist() == [0, 1, 3, 4] and cols.tolist() == [1, 0, 2, 1]
    g = vsgm.Graph.from_edges(50, rng.integers(0, 50, size=(20, 2)))
    g.write_bin(str(tmp_path / "g.bin"))
    a = vsgm.Graph.load(str(tmp_path / "g.bin"), drop_isolated=True).csr()
    b = g.drop_isolated().csr()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    with pytest.raises(vsgm.VsgmError):
        vsgm.Graph.load(str(tmp_path / "g.unknown"))


def test_preprocess_graph_text_outputs(vsgm, tmp_path):
    path = tmp_path / "g.txt"
    path.write_text("# c\n4 6\n6 8\n8 4\n8 8\n4 6\n10 12\n")
    exe = os.path.join(ROOT, "vsgm_amd", "bin", "preprocess_graph")
    for fmt in ("0", "3"):
        out = subprocess.run([exe, "-f", str(path), "-o", fmt], capture_output=True, text=True, timeout=60)
        assert out.returncode == 0, out.stderr
    # ids 4,6,8,10,12 -> shifted 0..8 -> isolated removed: 0,2,4,6,8 -> 0..4
    assert (tmp_path / "g.txt.snap.txt").read_text() == "0 1\n0 2\n1 2\n3 4\n"
    assert (tmp_path / "g.txt.adj.txt").read_text() == "0\t1 2 \n1\t0 2 \n2\t0 1 \n3\t4 \n4\t3 \n"
    assert (tmp_path / "g.txt.adj.txt.info").read_text() == "5 8"
    assert subprocess.run([exe, "-f", str(path), "-o", "2"], capture_output=True, timeout=60).returncode == 2


from hypothesis import given, settings, HealthCheck, strategies as st  # noqa: E402

_line = st.one_of(
    st.tuples(st.integers(0, 60), st.integers(0, 60), st.sampled_from(["\t", " ", "  ", " \t"]), st.sampled_from(["", " 1.5", "\t7"]),
              st.sampled_from(["\n", "\r\n"])).map(lambda t: f"{t[0] + 1000}{t[2]}{t[1] + 1000}{t[3]}{t[4]}"),
    st.sampled_from(["# comment\n", "% header\n", "\n", "c not an edge\n"]),
)


@settings(max_examples=60, deadline=None, suppress_health_check=[HealthCheck.function_scoped_fixture])
@given(lines=st.lists(_line, min_size=1, max_size=80), threads=st.integers(1, 9))
def test_snap_reader_matches_oracle_on_random_texts(vsgm, tmp_path, lines, threads):
    text = "".join(lines)
    path = tmp_path / "g.txt"
    path.write_text(text)
    try:
        pairs, V = I.parse_snap(text)
    except ValueError:
        with pytest.raises(vsgm.VsgmError):
            vsgm.Graph.read_snap(str(path), threads=threads)
        return
    rp, cols = vsgm.Graph.load(str(path), threads=threads).csr()
    want = I.csr_from_pairs(pairs, V, True, False)
    assert np.array_equal(rp, want[0]) and np.array_equal(cols, want[1])
    rp, cols = vsgm.Graph.read_snap(str(path), threads=threads).csr()
    want = I.csr_from_pairs(pairs, V, True, True)
    assert np.array_equal(rp, want[0]) and np.array_equal(cols, want[1])


REF_PREPROCESS = os.path.join(ROOT, "oracle", "_ref", "preprocess_graph")


@pytest.mark.skipif(not os.path.exists(REF_PREPROCESS), reason="oracle/_ref/preprocess_graph is built only where the reference tree exists")
@pytest.mark.parametrize("seed,kind", [(0, "txt"), (1, "txt"), (2, "txt"), (3, "lvid"), (4, "lvid")])
def test_preprocess_graph_matches_reference_tool(tmp_path, seed, kind):
    """The reference's own preprocess_graph (built from its sources) and ours
    write the same .bin, .snap.txt and .adj.txt(.info) bytes from a SNAP text
    (or a large-id .lvid text) with comments, self loops, duplicate edges and
    missing ids."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(50, 900))
    ids = [str(10 ** 20 + 7919 * i) if kind == "lvid" else str(i + 7) for i in range(n)]
    lines = ["# FromNodeId\tToNodeId"] + [f"{ids[a]}\t{ids[b]}" for a, b in rng.integers(0, n, size=(int(rng.integers(10, 4000)), 2))]
    text = "\n".join(lines) + "\n"
    outs = []
    for name, tool in (("ref", REF_PREPROCESS), ("me", os.path.join(ROOT, "vsgm_amd", "bin", "preprocess_graph"))):
        d = tmp_path / name
        d.mkdir()
        src = d / f"g.{kind}"
        src.write_text(text)
        files = []
        for fmt, suffixes in (("1", [".bin"]), ("0", [".snap.txt"]), ("3", [".adj.txt", ".adj.txt.info"])):
            r = subprocess.run([tool, "-f", str(src), "-d", "0", "-o", fmt], capture_output=True, timeout=120)
            assert r.returncode == 0, (name, fmt)
            files += [(d / (src.name + suf)).read_bytes() for suf in suffixes]
        outs.append(files)
    assert outs[0] == outs[1]


def test_preprocess_graph_golden_vectors(tmp_path):
    """Committed digests of the reference tool's outputs (written by
    tests/golden/make_preprocess_vectors.py) against ours."""
    import hashlib
    import json
    import sys

    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_preprocess_vectors import OUTPUTS, snap_text

    exe = os.path.join(ROOT, "vsgm_amd", "bin", "preprocess_graph")
    for v in json.load(open(os.path.join(ROOT, "tests", "golden", "preprocess_vectors.json"))):
        path = str(tmp_path / f"g{v['case'][0]}.txt")
        open(path, "w").write(snap_text(*v["case"]))
        for fmt, suffix in OUTPUTS:
            if not os.path.exists(path + suffix):
                assert subprocess.run([exe, "-f", path, "-d", "0", "-o", fmt], capture_output=True, timeout=120).returncode == 0
            assert hashlib.sha256(open(path + suffix, "rb").read()).hexdigest() == v["sha256"][suffix], (v["case"], suffix)
