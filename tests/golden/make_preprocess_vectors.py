"""Writes tests/golden/preprocess_vectors.json from the REFERENCE preprocess_graph.

Runs oracle/_ref/preprocess_graph (built by oracle/Makefile from the
reference's own src/preprocess/preprocess_graph.cc) on seeded SNAP texts and
stores each case's generator parameters with the SHA-256 of the .bin,
.snap.txt, .adj.txt and .adj.txt.info files it wrote. Only that data is
committed. Run in the container that has the reference tree:
    make -C oracle && python tests/golden/make_preprocess_vectors.py
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

# (seed, id range, edge lines, id offset)
CASES = [(1, 300, 1200, 5), (2, 2000, 6000, 0), (3, 40, 500, 100), (4, 1000, 800, 17)]
OUTPUTS = [("1", ".bin"), ("0", ".snap.txt"), ("3", ".adj.txt"), ("3", ".adj.txt.info")]


def snap_text(seed, n, m, offset):
    rng = np.random.default_rng(seed)
    lines = ["# FromNodeId\tToNodeId"] + [f"{a + offset}\t{b + offset}" for a, b in rng.integers(0, n, size=(m, 2))]
    return "\n".join(lines) + "\n"


def main():
    tool = os.path.join(ROOT, "oracle", "_ref", "preprocess_graph")
    vectors = []
    for case in CASES:
        digests = {}
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "g.txt")
            open(path, "w").write(snap_text(*case))
            for fmt, suffix in OUTPUTS:
                if not os.path.exists(path + suffix):
                    subprocess.run([tool, "-f", path, "-d", "0", "-o", fmt], check=True, capture_output=True, timeout=300)
                digests[suffix] = hashlib.sha256(open(path + suffix, "rb").read()).hexdigest()
        vectors.append({"case": list(case), "sha256": digests})
    with open(os.path.join(HERE, "preprocess_vectors.json"), "w") as f:
        json.dump(vectors, f, indent=0)
    print(f"wrote {len(vectors)} cases")


if __name__ == "__main__":
    main()
