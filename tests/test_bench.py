"""CPU: bench.py's host-side pieces — the --gpus guard and the
intersection rates it reports beside the metric."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 2 and "one GPU" in out.stdout


def test_intersection_rates_count_probes_per_backward_list(vsgm):
    p = vsgm.Plan.optimize(vsgm.Query.clique(4))
    # MATERIALIZE 0, COMPUTE 1 (|B|=1), MATERIALIZE 1, COMPUTE 2 (|B|=2), MATERIALIZE 2, COMPUTE_COUNT 3 (|B|=3)
    assert [op for op, _ in p.ops] == ["MATERIALIZE", "COMPUTE", "MATERIALIZE", "COMPUTE", "MATERIALIZE", "COMPUTE_COUNT"]
    levels = [{"work": 10 ** i, "paths": 7 * i} for i in range(6)]
    r = bench.intersection_rates(p.serialize(), p.ops, levels, 0.5)
    probes = 10 ** 1 * 1 + 10 ** 3 * 1 + 10 ** 5 * 2
    sets = 7 * 1 * 0 + 7 * 3 * 1 + 7 * 5 * 2
    assert r == {"probes_per_s": probes / 0.5, "set_intersections_per_s": sets / 0.5}
    assert bench.intersection_rates(p.serialize(), p.ops, levels, 0.0) is None
