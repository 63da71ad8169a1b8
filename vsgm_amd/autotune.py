"""Autotuner for the level-op launch shape and batch budget of each plan.

Knobs (per plan, applied by the executor before the plan's levels run):
  isect_kernel  map kernel variant of the intersection ops: 0 per-lane
                verification with wave-uniform runs and the membership index,
                1 wave-cursor LDS tiles (256-entry tile width per wavefront),
                2 per-lane search with one LDS-staged list per workgroup
                (4096-entry stage), 3 one workgroup per CU with the stable
                list of a run of paths in an LDS hash table, 4 as 3 with the
                flat map (1024-element blocks across paths, software-
                pipelined) for the COMPUTE_COUNT levels it covers;
  max_grid      workgroups per map launch (grid-stride beyond);
  batch_bytes   per-level batch budget (0: budget / remaining levels);
  ticket_chunks work-chunk size of isect_kernel 3 / 4: consecutive 8192-element
                chunks a workgroup takes per ticket (its LDS table is rebuilt
                at most once per run change within a ticket);
  cold_weight   isect_kernel 3 / 4's pivot cost of a test against a list that
                changes per path (streaming a pivot element costs 1).
The element chunk of kernels 0-2 is fixed at 2048 (it sizes the LDS copy of
the chunk's segment offsets).

Search: coordinate descent from the current setting. Each candidate runs the
plan over the executor's resident views (`match_all([plan])`) and is timed
with the executor's HIP events; a candidate replaces the incumbent only when
it is at least `margin` faster. Every candidate's count must equal the
incumbent's (they are the same enumeration); a mismatch raises.
"""
import json
import time
from typing import Dict, List, Optional, Sequence

CANDIDATES = {
    "isect_kernel": [0, 1, 2, 3, 4],
    "max_grid": [2048, 8192, 32768],
    "batch_bytes": [0, 1 << 28, 1 << 31],
    "ticket_chunks": [4, 16, 64],
    "cold_weight": [1, 4, 16, 64],
}


def _run(ex, plan: int, cfg: dict, repeats: int):
    ex.set_plan_tuning(plan, **cfg)
    best, count = None, None
    for _ in range(repeats):
        ex.reset()
        ex.match_all([plan])
        ms = ex.match_ms[plan]
        c = ex.total_counts[plan]
        if count is not None and c != count:
            raise RuntimeError(f"plan {plan}: count changed between repeats ({count} vs {c})")
        count = c
        best = ms if best is None else min(best, ms)
    return best, count


def tune(ex, plans: Optional[Sequence[int]] = None, repeats: int = 2, margin: float = 0.03,
         candidates: Dict[str, List[int]] = CANDIDATES, log=None) -> Dict[int, dict]:
    """Tune every listed plan (all when None) of executor `ex` over its
    resident views; leaves the chosen setting installed and the executor's
    counts reset. Returns {plan: {"config", "ms", "default_ms", "trials"}}."""
    n = len(ex.plans)
    plans = list(range(n)) if plans is None else list(plans)
    out = {}
    for p in plans:
        cfg = dict(ex.plan_tuning(p))
        t0 = time.perf_counter()
        default_ms, count = _run(ex, p, cfg, repeats)
        best_ms = default_ms
        trials = 1
        for key, values in candidates.items():
            start = cfg[key]  # already measured as part of the incumbent
            for v in values:
                if v == start:
                    continue
                trial = dict(cfg, **{key: v})
                ms, c = _run(ex, p, trial, repeats)
                trials += 1
                if c != count:
                    raise RuntimeError(f"plan {p}: {trial} counted {c}, expected {count}")
                if ms < best_ms * (1.0 - margin):
                    cfg, best_ms = trial, ms
        ex.set_plan_tuning(p, **cfg)
        out[p] = {"config": cfg, "ms": best_ms, "default_ms": default_ms, "trials": trials,
                  "tune_s": time.perf_counter() - t0}
        if log:
            log(f"tune plan {p}: {cfg} {best_ms:.3f} ms (default {default_ms:.3f} ms, {trials} trials)")
    ex.reset()
    return out


def save(results: Dict[int, dict], path: str):
    with open(path, "w") as f:
        json.dump({str(k): v for k, v in results.items()}, f, indent=1)


def load(ex, path: str):
    """Install a saved tuning (plan indices must refer to the same plan list)."""
    with open(path) as f:
        data = json.load(f)
    for k, v in data.items():
        ex.set_plan_tuning(int(k), **v["config"])
