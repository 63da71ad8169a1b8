"""CPU: bench.py's host-side pieces — the --gpus guard and the
intersection rates it reports beside the metric."""
import json
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


class _Clock:
    def __init__(self, step_s, start=0.0):
        self.t, self.step_s, self.runs, self.restarts = start, step_s, [], 0

    def now(self):
        return self.t

    def step(self, label):
        self.t += self.step_s
        self.runs.append(label)
        return self.step_s

    def restart(self):
        self.restarts += 1


def test_budget_with_a_step_longer_than_half_the_budget():
    # one 300 s step after 40 s of set-up: no room for a warm-up, one timed step
    c = _Clock(300.0, start=40.0)
    warm, timed, t0, t1 = bench.schedule_steps(c.step, c.step, c.restart, 5, 20, 480.0, c.now)
    assert (warm, timed) == (0, 1) and t1 - t0 == 300.0 and c.runs == ["step 0"] and c.t <= 480.0


def test_budget_warm_up_then_timed_steps():
    # 100 s steps after 40 s: step 0 becomes the warm-up, then timed steps while one more fits
    c = _Clock(100.0, start=40.0)
    warm, timed, t0, t1 = bench.schedule_steps(c.step, c.step, c.restart, 5, 20, 480.0, c.now)
    assert warm == 1 and timed == 3 and c.t <= 480.0 and t1 - t0 == 300.0 and c.restarts == 2
    # short steps: the requested counts are upper bounds that are reached
    c = _Clock(1.0)
    warm, timed, _, _ = bench.schedule_steps(c.step, c.step, c.restart, 5, 20, 480.0, c.now)
    assert (warm, timed) == (5, 20) and len(c.runs) == 25


def test_budget_without_warm_up():
    c = _Clock(50.0)
    warm, timed, t0, t1 = bench.schedule_steps(c.step, c.step, c.restart, 0, 3, 480.0, c.now)
    assert (warm, timed) == (0, 3) and t1 - t0 == 150.0


def test_roofline_uses_traffic_only_for_its_own_configuration(tmp_path, monkeypatch):
    # a traffic profile of map_icount on rmat22 / heuristic / kernel 4, measured on this kernel source
    (tmp_path / "profiles").mkdir()
    (tmp_path / "profiles" / "traffic_map_icount_rmat22_heuristic_k4.json").write_text(
        json.dumps({"kernel_source_sha": bench.kernel_source_sha(), "bytes_per_launch": 2.0e11}))
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    kstats = [{"name": "map_icount", "launches": 10, "total_ms": 100.0, "algo_bytes": 10 * 4_000_000_000}]
    # 4e9 algorithmic bytes per 10 ms launch = 400 GB/s
    r = bench.make_roofline(kstats, "rmat22", "heuristic", 4)
    assert r["achieved"] == 400.0 and r["frac"] == 0.05 and r["bound"] == "hbm"
    assert r["traffic"] == 2.0e11
    assert r["traffic_gbs"] == round(r["traffic"] / 0.010 / 1e9, 1)
    # another workload, plan or map kernel: no traffic
    for args in (("rmat20", "heuristic", 4), ("rmat22", "learned", 4), ("rmat22", "heuristic", 3)):
        assert bench.make_roofline(kstats, *args)["traffic"] is None
    # a profile measured on another kernel source is not used
    monkeypatch.setattr(bench, "kernel_source_sha", lambda: "0000000000000000")
    assert bench.make_roofline(kstats, "rmat22", "heuristic", 4)["traffic"] is None


def test_roofline_frac_never_exceeds_one():
    kstats = [{"name": "map_x", "launches": 1, "total_ms": 1.0, "algo_bytes": 20_000_000_000}]  # 20 TB/s "algorithmic"
    r = bench.make_roofline(kstats, "nowhere", "heuristic", 3)
    assert r["frac"] == 1.0 and "capped" in r["bound_note"]
