"""Writes tests/golden/rmat_counts.json: exact pattern counts of the seeded
R-MAT graphs the benchmark and the full-size GPU tests use, from the
oracle's textbook counters (forward triangles, 4-cliques by nested
intersections over the degree orientation, diamonds as sum over edges of
C(t_e, 2), 4-cycles as half the sum over vertex pairs of C(common, 2)),
which the oracle tests pin against closed forms, brute force and each
other. The graph is identified by its
size and a SHA-256 of its CSR arrays, so a fixture can only be compared
with the same graph.

    python tests/golden/make_rmat_counts.py [scale ...]   (default: 22)

Runs on the host cores (about 15 minutes for scale 22 on 8 cores).
"""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import oracle  # noqa: E402
import vsgm_amd as V  # noqa: E402

OUT = os.path.join(HERE, "rmat_counts.json")


def csr_digest(rp, cl):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(rp, dtype=np.uint64).tobytes())
    h.update(np.ascontiguousarray(cl, dtype=np.uint32).tobytes())
    return h.hexdigest()


def main(scales):
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for s in scales:
        t0 = time.time()
        g = V.Graph.rmat(s, 16, 0.57, 0.19, 0.19, seed=1)
        rp, cl = g.csr(copy=False)
        rec = {"scale": s, "edge_factor": 16, "abc": [0.57, 0.19, 0.19], "seed": 1, "vertices": int(g.vertex_count),
               "edges": int(g.edge_count), "csr_sha256": csr_digest(rp, cl), "counts": {}, "seconds": {}}
        for key, fn in (("kc3", oracle.triangles), ("kc4", oracle.four_cliques_ranked), ("p2", oracle.diamonds), ("p1", oracle.four_cycles)):
            t = time.time()
            rec["counts"][key] = int(fn(rp, cl))
            rec["seconds"][key] = round(time.time() - t, 1)
            print(f"scale {s} {key}: {rec['counts'][key]} in {rec['seconds'][key]} s", flush=True)
        data[f"rmat{s}"] = rec
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        print(f"scale {s} done in {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main([int(x) for x in sys.argv[1:]] or [22])
