"""Writes tests/golden/fennel_vectors.json from the REFERENCE Fennel tool.

Runs oracle/_ref/fennel (built by oracle/Makefile from the reference's own
src/tools/fennel.cc) on seeded R-MAT graphs and stores each case's inputs
(graph parameters, parts, hop) with the reference's partition. Only that
data is committed. Run in the container that has the reference tree:
    make -C oracle && python tests/golden/make_fennel_vectors.py
"""
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import vsgm_amd as vsgm  # noqa: E402

# (scale, edge_factor, graph seed, parts, hop)
CASES = [(9, 8, 1, 4, 1), (10, 4, 2, 8, 1), (9, 16, 3, 3, 2), (10, 8, 5, 5, 2), (8, 8, 6, 1, 1)]


def main():
    tool = os.path.join(ROOT, "oracle", "_ref", "fennel")
    vectors = []
    for scale, ef, seed, parts, hop in CASES:
        with tempfile.TemporaryDirectory() as d:
            gf = os.path.join(d, "g.bin")
            vsgm.Graph.rmat(scale, ef, seed=seed).write_bin(gf)
            subprocess.run([tool, "-f", gf, "-p", str(parts), "-h", str(hop)], check=True, capture_output=True, timeout=300)
            part = [int(x) for x in open(f"{gf}.fennel.part.{parts}").read().split()]
        vectors.append({"case": [scale, ef, seed, parts, hop], "parts": part})
    with open(os.path.join(HERE, "fennel_vectors.json"), "w") as f:
        json.dump(vectors, f, separators=(",", ":"))
    print(f"wrote {len(vectors)} cases")


if __name__ == "__main__":
    main()
