"""Writes tests/golden/features_vectors.json from the REFERENCE features tool.

Runs oracle/_ref/features (built by oracle/Makefile from the reference's own
src/tools/features.cc) on seeded R-MAT graphs with -h 1, -h 2 and -dn n, and
stores each case's inputs with the SHA-256 of the file the reference wrote.
Only that data is committed. Run in the container that has the reference tree:
    make -C oracle && python tests/golden/make_features_vectors.py
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import vsgm_amd as vsgm  # noqa: E402

# (scale, edge_factor, graph seed, flags, output suffix)
CASES = [(8, 8, 1, ["-h", "1"], "1hop"), (9, 4, 2, ["-h", "1"], "1hop"), (8, 8, 3, ["-h", "2"], "2hop"),
         (9, 8, 4, ["-h", "2"], "2hop"), (8, 4, 5, ["-dn", "2"], "x2"), (7, 16, 6, ["-dn", "3"], "x3")]


def main():
    tool = os.path.join(ROOT, "oracle", "_ref", "features")
    vectors = []
    for scale, ef, seed, flags, out in CASES:
        with tempfile.TemporaryDirectory() as d:
            gf = os.path.join(d, "g.bin")
            vsgm.Graph.rmat(scale, ef, seed=seed).write_bin(gf)
            subprocess.run([tool, "-f", gf, "-t", "4"] + flags, check=True, capture_output=True, timeout=300)
            digest = hashlib.sha256(open(f"{gf}.{out}.bin", "rb").read()).hexdigest()
        vectors.append({"case": [scale, ef, seed, flags, out], "sha256": digest})
    with open(os.path.join(HERE, "features_vectors.json"), "w") as f:
        json.dump(vectors, f, indent=0)
    print(f"wrote {len(vectors)} cases")


if __name__ == "__main__":
    main()
