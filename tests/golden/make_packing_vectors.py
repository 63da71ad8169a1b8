"""Writes tests/golden/packing_vectors.json from the REFERENCE view-packing tool.

Runs oracle/_ref/view_packing (built by oracle/Makefile from the reference's
own src/tools/view_packing.cc) on seeded R-MAT graphs with seeded one- and
two-layer partitions, and stores each case's inputs (graph parameters,
partition, flags) with the reference's .vbmap and .vbstat output. Only that
data is committed. Run in the container that has the reference tree:
    make -C oracle && python tests/golden/make_packing_vectors.py
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import vsgm_amd as vsgm  # noqa: E402

# (scale, edge_factor, graph seed, partition seed, first layer, second layer, hop, mem GB, pool, degree order)
CASES = [
    (10, 8, 1, 0, 4, 0, 2, 0.00006, 1, 0),
    (10, 8, 1, 1, 4, 0, 2, 0.00005, 3, 1),
    (10, 8, 2, 2, 2, 3, 2, 0.00005, 2, 0),
    (10, 8, 2, 3, 3, 2, 3, 0.00006, 1, 1),
    (10, 8, 3, 4, 4, 0, 1, 0.000003, 2, 0),
    (10, 16, 4, 5, 8, 0, 2, 0.00008, 4, 1),
]


def run_case(tool, case, d):
    scale, ef, gseed, pseed, first, second, hop, mem, pool, order = case
    g = vsgm.Graph.rmat(scale, ef, seed=gseed)
    gf = os.path.join(d, "g.bin")
    g.write_bin(gf)
    parts = first * (second or 1)
    part = np.random.default_rng(pseed).integers(0, parts, g.vertex_count)
    suffix = f"{first}x{second}" if second else f"{first}"
    pf = f"{gf}.kmeans.{suffix}"
    np.savetxt(pf, part, fmt="%d")
    out = subprocess.run([tool, "-gf", gf, "-pf", pf, "-h", str(hop), "-m", repr(mem), "-t", str(first), "-s", str(pool),
                          "-o", str(order), "-d", "1"], capture_output=True, text=True, timeout=120)
    base = f"{gf}.{hop}.hop"
    res = {"returncode": out.returncode}
    if out.returncode == 0:
        res["vbmap"] = open(base + ".vbmap").read()
        res["vbstat"] = open(base + ".vbstat").read()
    return part, res


def main():
    tool = os.path.join(ROOT, "oracle", "_ref", "view_packing")
    vectors = []
    for case in CASES:
        with tempfile.TemporaryDirectory() as d:
            part, res = run_case(tool, case, d)
        assert res["returncode"] == 0, case
        vbmap = [int(x) for x in res["vbmap"].split()]
        vectors.append({"case": list(case), "partition": part.tolist(), "bins": vbmap[0], "assignment": vbmap[1:],
                        "vbstat": res["vbstat"]})
    with open(os.path.join(HERE, "packing_vectors.json"), "w") as f:
        json.dump(vectors, f, separators=(",", ":"))
    print(f"wrote {len(vectors)} cases")


if __name__ == "__main__":
    main()
