"""Writes tests/golden/ref_plans.txt from the REFERENCE planner.

Runs oracle/_ref/ref_plan_dump (built by oracle/Makefile from the
reference's own src/query headers and vendored bliss sources) on:
  preset 0..18, clique 3..7, graphlets 3..5,
  and every connected shape on 3..6 vertices (vsgm_amd/shapes.py).
Only the resulting text (inputs and the reference's plans) is committed.
Run in the container that has the reference tree:
    make -C oracle && python tests/golden/make_plan_fixtures.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from vsgm_amd.shapes import all_shapes  # noqa: E402


def specs():
    out = [f"preset {k}" for k in range(19)]
    out += [f"clique {k}" for k in range(3, 8)]
    out += [f"graphlets {k}" for k in range(3, 6)]
    for n, _mask, a in all_shapes():
        out.append(f"adj {n} " + "".join(str(int(x)) for x in a.reshape(-1)))
    return out


def main():
    tool = os.path.join(ROOT, "oracle", "_ref", "ref_plan_dump")
    if not os.path.exists(tool):
        sys.exit("oracle/_ref/ref_plan_dump is missing: run `make -C oracle` where the reference tree exists")
    dest = os.path.join(HERE, "ref_plans.txt")
    subprocess.run([tool, dest], input="\n".join(specs()) + "\n", text=True, check=True, stdout=subprocess.DEVNULL)
    with open(dest) as f:
        n = sum(1 for line in f if line.startswith("query:"))
    print(f"wrote {n} reference plans to {dest}")


if __name__ == "__main__":
    main()
