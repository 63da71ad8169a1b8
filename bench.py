#!/usr/bin/env python3
"""Benchmark of the VSGM pm_pipeline enumeration path on one MI355X.

Workload (default): seeded R-MAT scale 22, edge factor 16,
a/b/c/d = .57/.19/.19/.05, made undirected, deduplicated, self loops
removed, as a u32 CSR; one k-means partition with k = 1 (every vertex in one
view bin) materialised as a 2-hop view with the root-degree filter; patterns
4-clique (-kc 4), diamond (-p 2) and 4-cycle (-p 1), heuristic plans, one
pattern per launch sequence. A step is one Match of all three patterns over
the view resident in HBM; `value` is embeddings enumerated per second over
the timed steps.

Vertex ids. By default the generated graph is renumbered in ascending
(degree, id) order before the view is built (--relabel degree): the same
graph with other ids, so every count is the same (checked against the
golden counts of the graph as generated), but the plans' symmetry
conditions, which root each embedding at its smallest id, then start from
low-degree vertices and enumerate far fewer partial embeddings. One step
over the ids as generated runs first and is reported as
config.ids_as_generated (--generated-ids-step 0 skips it; --relabel none
times the ids as generated).

Wall-clock budget. The whole run (graph generation, view, CPU baseline,
warm-up, timed steps) is kept within --budget-s seconds counted from the
start of the process: --warmup and --steps are upper bounds, every pattern
gets at least one full timed step, and the line reports the steps actually
warmed up and timed. Progress goes to stderr as each stage ends.

The JSON line also carries
  roofline      for the kernel with the most device time in the timed
                region: algorithmic bytes per launch (every candidate
                element streamed once from its pivot list, plus each path's
                bound vertices, row_ptrs and segment words) / average launch
                time (HIP events on the executor's stream), against 8 TB/s;
                `traffic`: rocprof-measured fabric bytes per launch from
                profiles/, only when that file was measured on this
                workload, plan, map kernel and kernel source;
  cpu_baseline  the oracle's C restatement of the level algorithm on the
                host cores, over a bounded sample of root vertices; the GPU
                counts from the same roots are checked equal to it.
"""
import argparse
import hashlib
import json
import os
import sys
import time

T_START = time.perf_counter()
ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PATTERNS = {"kc3": ("clique", 3), "kc4": ("clique", 4), "kc5": ("clique", 5)}
PATTERNS.update({f"p{k}": ("preset", k) for k in range(19)})
NAMES = {"kc3": "triangle", "kc4": "4-clique", "kc5": "5-clique", "p1": "4-cycle", "p2": "diamond", "p4": "house"}
KERNEL_SOURCE = os.path.join(ROOT, "vsgm_amd", "csrc", "hip", "vsgm_hip.hip")


def elapsed():
    return time.perf_counter() - T_START


def log(msg):
    print(f"[bench {elapsed():7.1f}s] {msg}", file=sys.stderr, flush=True)


def make_query(V, key):
    kind, k = PATTERNS[key]
    return V.Query.clique(k) if kind == "clique" else V.Query.preset(k)


def kernel_source_sha():
    with open(KERNEL_SOURCE, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def intersection_rates(plan_text, ops, levels, seconds):
    """Per second over a pattern's timed match: candidate-vs-list probes
    (every candidate streamed from a pivot list, times max(1, |B| - 1) lists
    it is tested against) and set intersections (paths times |B| - 1), over
    the intersecting levels (COMPUTE, COMPUTE_COUNT, COMPUTE_PATH_COUNT).
    `ops` are the plan's (op, vertex) levels, `levels` their stats."""
    back = {}
    for line in plan_text.splitlines():
        if line.startswith("backward:"):
            rows = line[len("backward:"):].strip()[1:-1].split("][") if line.strip() != "backward:" else []
            back = {u: [x for x in r.split(",") if x] for u, r in enumerate(rows)}
    probes = sets = 0
    for (op, u), st in zip(ops, levels):
        if op in ("COMPUTE", "COMPUTE_COUNT", "COMPUTE_PATH_COUNT"):
            b = len(back.get(u, []))
            probes += st["work"] * max(1, b - 1)
            sets += st["paths"] * max(0, b - 1)
    if seconds <= 0:
        return None
    return {"probes_per_s": probes / seconds, "set_intersections_per_s": sets / seconds}


def cpu_baseline(O, rp, cl, plan_texts, budget_s, threads, seed=12345):
    """Oracle (C restatement of the level algorithm, OpenMP over roots) on a
    seeded uniform sample of root vertices, all patterns. The sample grows
    4x until a pass takes at least budget_s / 4, every root is in, or the
    next pass would not fit in what is left of budget_s; the rate of the
    last pass is reported (embeddings found from the sampled roots per
    second), together with its roots and per-pattern counts."""
    import numpy as np

    n = len(rp) - 1
    rng = np.random.default_rng(seed)
    order = rng.permutation(n).astype(np.uint32)
    m = 64
    spent = 0.0
    while True:
        roots = np.sort(order[: min(m, n)])
        t = time.perf_counter()
        per = [O.count_plan(rp, cl, txt, sources=roots, threads=threads) for txt in plan_texts]
        dt = time.perf_counter() - t
        spent += dt
        total = sum(per)
        log(f"cpu baseline sample {len(roots)} roots: {total} embeddings in {dt:.2f}s")
        if dt * 4 > budget_s - spent or dt >= budget_s / 4 or m >= n:
            return {
                "value": total / dt if dt > 0 else 0.0,
                "unit": "embeddings/s",
                "cores": threads,
                "kind": "port",
                "sample": f"{len(roots)} uniformly drawn roots (seed {seed}) of {n} vertices, all patterns; "
                f"{total} embeddings in {dt:.2f} s",
            }, roots, per
        m *= 4


def schedule_steps(run_first, run_step, restart, warmup, steps, budget_s, now, reserve=20.0):
    """Warm-up and timed steps within budget_s seconds of the clock `now`.
    The first full step is timed like every other; it counts as a warm-up
    step only when one more full step still fits after it, else it is the
    one timed step (so a timed step always exists). Timed steps come first
    in the budget: further warm-up steps (at most `warmup` in all) run only
    in the time left once all `steps` timed steps are provided for; timed
    steps then run while one more still fits (at most `steps`).
    `restart` resets the counts and the kernel records before the timed
    steps. Returns (warm-up steps, timed steps, t0, t1), t0 / t1 bracketing
    the timed steps on the same clock."""
    restart()
    t0 = now()
    est = run_first("step 0")
    timed, warm = 1, 0
    if warmup > 0 and now() + est + reserve <= budget_s:
        warm = 1
        spare = int((budget_s - reserve - now()) // est) - max(steps, 1) if est > 0 else warmup
        while warm < warmup and spare > 0:
            est = run_step(f"warm-up step {warm}")
            warm += 1
            spare -= 1
        restart()
        t0 = now()
        est = run_step("timed step 0")
    while timed < max(steps, 1) and now() + est + reserve <= budget_s:
        est = run_step(f"timed step {timed}")
        timed += 1
    return warm, timed, t0, now()


def golden_check(args, g, keys, counts):
    """Counts of the whole run against the committed golden counts of the
    same graph (tests/golden/rmat_counts.json, written by the oracle's
    textbook counters; the graph is identified by its CSR digest). Returns
    the comparison, or None when no golden record covers this graph; a
    mismatch stops the run."""
    path = os.path.join(ROOT, "tests", "golden", "rmat_counts.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f).get(f"rmat{args.scale}")
    if not rec or rec["edge_factor"] != args.edge_factor or rec["seed"] != args.seed:
        return None
    import numpy as np

    rp, cl = g.csr(copy=False)
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(rp, dtype=np.uint64).tobytes())
    h.update(np.ascontiguousarray(cl, dtype=np.uint32).tobytes())
    if h.hexdigest() != rec["csr_sha256"]:
        return None
    have = {k: c for k, c in zip(keys, counts) if k in rec["counts"]}
    out = {"checked": sorted(have), "equal": all(rec["counts"][k] == c for k, c in have.items())}
    log(f"golden counts: {out}")
    assert out["equal"], f"counts {have} differ from the golden counts {rec['counts']}"
    return out


def load_traffic(kernel, workload, plan, isect_kernel):
    """Measured traffic of `kernel` on this exact configuration, or None: the
    profile file must name the same workload, plan, map kernel variant and
    kernel source (sha256 prefix of vsgm_hip.hip)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}_{plan}_k{isect_kernel}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if d.get("kernel_source_sha") != kernel_source_sha():
        log(f"traffic profile {os.path.basename(path)} was measured on another kernel source; not used")
        return None
    return d


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="upper bound of timed steps (the budget may allow fewer)")
    ap.add_argument("--warmup", type=int, default=1, help="upper bound of untimed warm-up steps")
    ap.add_argument("--budget-s", type=float, default=480.0, help="wall-clock budget of the whole run from process start")
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--patterns", default="kc4,p2,p1")
    ap.add_argument("--cpu-seconds", type=float, default=30.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--detail-out", default="")
    ap.add_argument("--hub-bytes", type=float, default=-1, help="membership index bytes (-1 auto, 0 off)")
    ap.add_argument("--hub-min-degree", type=int, default=64)
    ap.add_argument("--isect-kernel", type=int, default=4)
    ap.add_argument("--cold-weight", type=int, default=None,
                    help="pivot weight of a list that changes from path to path (default: 1 with degree-order ids, "
                         "16 with ids as generated, the fastest of 1/4/16/64 on each)")
    ap.add_argument("--ticket-chunks", type=int, default=64)
    ap.add_argument("--progress", type=float, default=30.0, help="seconds between progress lines of long map launches")
    ap.add_argument("--cl-vertices", type=int, default=65_600_000)
    ap.add_argument("--cl-edges", type=float, default=1.81e9, help="target undirected edges of the Chung-Lu graph")
    ap.add_argument("--cl-sample-s", type=float, default=120.0, help="Chung-Lu: seconds a sampled 4-clique / house pass may grow to")
    ap.add_argument("--cl-parity-roots", type=int, default=8, help="Chung-Lu: roots whose counts are checked against the oracle")
    ap.add_argument("--workload", choices=["single", "batched", "copy-calibration", "chunglu", "triangles-cpu"], default="single",
                    help="single: the 4-clique/diamond/4-cycle line; batched: many random patterns over k view bins, "
                         "one-at-a-time vs batched+tuned vs learned+batched; copy-calibration: kernels of known HBM "
                         "traffic (a 2 GiB fill and a 2 GiB reduction) that scale the rocprof byte counters")
    ap.add_argument("--plan", choices=["heuristic", "learned"], default="heuristic",
                    help="matching orders of the single workload: the heuristic planner (the metric's configuration) "
                         "or the learned planner")
    ap.add_argument("--relabel", choices=["none", "degree"], default="degree",
                    help="vertex ids of the single workload: in ascending degree order (the same graph renumbered, "
                         "same counts; the plans' smallest-id roots then start from low-degree vertices) or as "
                         "generated")
    ap.add_argument("--generated-ids-step", type=int, default=1,
                    help="with --relabel degree: first run one step over the graph with its ids as generated and "
                         "report it beside the timed steps")
    ap.add_argument("--n-patterns", type=int, default=64)
    ap.add_argument("--share-prefixes", type=int, default=0, help="batched: match the plans as a prefix tree")
    ap.add_argument("--bins", type=int, default=8)
    args = ap.parse_args()
    os.environ["VSGM_PROGRESS"] = str(args.progress)

    cold_given = args.cold_weight is not None
    if not cold_given:
        args.cold_weight = 1 if args.relabel == "degree" else 16
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the path shards by view bins "
              f"across GPUs with no data exchange; only the single-GPU share is built and measured here)", flush=True)
        sys.exit(2)

    if args.workload == "batched":
        return batched_main(args)
    if args.workload == "copy-calibration":
        return calibration_main()
    if args.workload == "chunglu":
        return chunglu_main(args)
    if args.workload == "triangles-cpu":
        return triangles_cpu_main(args)

    import numpy as np

    import vsgm_amd as V
    from vsgm_amd import partition

    assert V.device_count() > 0, "no GPU visible"
    keys = [k.strip() for k in args.patterns.split(",") if k.strip()]
    workload = f"rmat{args.scale}" + ("" if args.relabel == "none" else f"_{args.relabel}_ids")

    t = time.perf_counter()
    g = V.Graph.rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, seed=args.seed)
    gen_s = time.perf_counter() - t
    log(f"R-MAT s{args.scale} ef{args.edge_factor}: V={g.vertex_count} E={g.edge_count} maxdeg={g.max_degree} "
        f"generated in {gen_s:.1f}s")
    g_generated, relabel_s = g, None
    if args.relabel == "degree":
        t = time.perf_counter()
        g, _ = g.relabel_by_degree()
        relabel_s = time.perf_counter() - t
        log(f"ids relabelled in ascending degree order in {relabel_s:.1f}s")

    queries = [make_query(V, k) for k in keys]
    plans = [V.Plan.optimize(q) for q in queries]
    learned_info = None
    if args.plan == "learned":
        plans, learned_info = learned_plans(V, args, keys, queries, g)
    hop = max(p.hop for p in plans)
    root_degree = min(p.root_degree for p in plans)

    def single_view(graph, cold_weight):
        """One k-means partition with k = 1 on the GPU (every vertex in
        cluster 0), packed as one view bin (the hop-hop view with the
        root-degree filter) and transferred to a new executor."""
        t = time.perf_counter()
        assign, km_sizes, _ = partition.kmeans(graph, 1, iters=1, seed=args.seed)
        vbm = V.ViewBinManager.from_assignment(graph, hop, 1, assign)
        holder = V.ViewBinHolder(graph.vertex_count, max(4, graph.entries * 4))
        vbm.materialize(0, holder, root_degree)
        view_s = time.perf_counter() - t
        info = holder.info()
        log(f"k-means k=1 ({km_sizes.tolist()}) + view: hop={hop} root_degree={root_degree} vertices={info['vertices']} "
            f"entries={info['entries']} in {view_s:.1f}s")
        ex = V.PipelineExecutor(0, graph.vertex_count, plans, max_sources=vbm.max_sources)
        ex.set_option("hub_bytes", args.hub_bytes)
        ex.set_option("hub_min_degree", args.hub_min_degree)
        ex.set_option("isect_kernel", args.isect_kernel)
        ex.set_option("cold_weight", cold_weight)
        ex.set_option("ticket_chunks", args.ticket_chunks)
        t = time.perf_counter()
        ex.transfer(holder)
        xfer_s = time.perf_counter() - t
        log(f"transfer {info['entries'] * 4 / 1e9 + (graph.vertex_count + 1) * 8 / 1e9:.2f} GB + membership index of "
            f"{ex.hub_count()} lists in {xfer_s:.2f}s")
        return ex, view_s, xfer_s

    # With relabelled ids, one step over the graph with its ids as generated
    # (same plans, same view construction) is run first and reported beside
    # the timed steps; the timed steps get the budget left after it.
    generated_ids = None
    if args.relabel != "none" and args.generated_ids_step:
        gen_cold = args.cold_weight if cold_given else 16
        gx, _, _ = single_view(g_generated, gen_cold)
        t = time.perf_counter()
        gx.match()
        dt = time.perf_counter() - t
        generated_ids = {"ms_per_step": round(dt * 1000.0, 1), "counts": gx.total_counts,
                         "per_pattern_ms": {k: round(m, 1) for k, m in zip(keys, gx.match_ms)},
                         "cold_weight": gen_cold,
                         "note": "one step over the graph with its ids as generated, same plans and view, before the "
                                 "timed steps"}
        log(f"ids as generated: one step {dt:.2f}s counts={gx.total_counts}")
        del gx

    ex, view_s, xfer_s = single_view(g, args.cold_weight)

    # CPU baseline (and the parity sample) before the GPU steps, so the steps
    # get whatever budget is left
    cpu, roots, cpu_counts, sample_check = None, None, None, None
    rp, cl = g.csr(copy=False)
    if not args.no_cpu_baseline:
        import oracle as O

        threads = min(os.cpu_count() or 1, 16)
        cpu, roots, cpu_counts = cpu_baseline(O, rp, cl, [p.serialize() for p in plans], args.cpu_seconds, threads)
        log(f"cpu baseline: {cpu}")
        sx = V.PipelineExecutor(0, g.vertex_count, plans)
        sx.set_option("isect_kernel", args.isect_kernel)
        sx.transfer_graph(g, roots)
        sx.match()
        sample_check = {"roots": int(len(roots)), "gpu": sx.total_counts, "oracle": cpu_counts,
                        "equal": sx.total_counts == cpu_counts}
        del sx
        log(f"sampled-root parity: gpu {sample_check['gpu']} oracle {cpu_counts}")
        assert sample_check["equal"], f"GPU counts {sample_check['gpu']} differ from the oracle {cpu_counts} on the sampled roots"

    def run_step(label):
        ts = time.perf_counter()
        ex.match()
        dt = time.perf_counter() - ts
        log(f"{label}: {dt:.2f}s counts={ex.total_counts}")
        return dt

    def restart():
        ex.reset()
        ex.profile(True)  # drops the records of earlier steps

    first_counts = []

    def run_first(label):
        dt = run_step(label)
        first_counts.extend(ex.total_counts)
        return dt

    warm, steps, t0, t1 = schedule_steps(run_first, run_step, restart, args.warmup, args.steps, args.budget_s, elapsed)
    kstats = ex.kernel_stats()
    ex.profile(False)
    log(f"{warm} warm-up and {steps} timed steps (requested {args.warmup} and {args.steps}) within the {args.budget_s:.0f}s budget")
    ref_counts = first_counts if warm else None

    counts = [c // steps for c in ex.total_counts]
    assert all(c * steps == tc for c, tc in zip(counts, ex.total_counts)), "counts differ between steps"
    if ref_counts is not None:
        assert counts == ref_counts, f"timed counts {counts} differ from warmup counts {ref_counts}"
    golden = golden_check(args, g_generated, keys, counts)  # counts do not depend on the ids
    if generated_ids is not None:
        generated_ids["equal_counts"] = generated_ids["counts"] == counts
        generated_ids["speedup"] = round(generated_ids["ms_per_step"] / ((t1 - t0) * 1000.0 / steps), 2)
        assert generated_ids["equal_counts"], f"counts with ids as generated {generated_ids['counts']} differ from {counts}"
    ms_per_step = (t1 - t0) * 1000.0 / steps
    match_ms = [m / steps for m in ex.match_ms]
    value = sum(counts) / (ms_per_step / 1000.0)
    per_pattern = {f"{NAMES.get(k, k)}({k})": {"count": c, "ms": round(m, 3), "embeddings_per_s": c / (m / 1000.0) if m else 0.0}
                   for k, c, m in zip(keys, counts, match_ms)}

    roofline = make_roofline(kstats, workload, args.plan, args.isect_kernel)
    for s_ in sorted(kstats, key=lambda s: -s["total_ms"]):
        log(f"kernel {s_['name']:16s} launches={s_['launches']:6d} total={s_['total_ms']:10.2f}ms "
            f"algoGB/s={s_['algo_bytes'] / max(s_['total_ms'], 1e-9) / 1e6:8.1f} work={s_['work']:.3e} paths={s_['paths']:.3e}")
    level_stats = {k: ex.level_stats(i) for i, k in enumerate(keys)}
    for k in keys:
        for lvl, st in enumerate(level_stats[k]):
            log(f"{k} L{lvl} batches={st['batches']} paths={st['paths']} work={st['work']} outputs={st['outputs']}")
    for k, p, m in zip(keys, plans, match_ms):
        try:  # reported beside the metric; never allowed to stop the line
            rates = intersection_rates(p.serialize(), p.ops, [{**st, "work": st["work"] // steps,
                                                               "paths": st["paths"] // steps} for st in level_stats[k]],
                                       m / 1000.0)
        except Exception as ex_:  # noqa: BLE001
            log(f"intersection rates of {k} unavailable: {ex_}")
            rates = None
        per_pattern[f"{NAMES.get(k, k)}({k})"]["intersections"] = rates

    line = {
        "metric": "embeddings/sec per pattern, 1 MI355X; learned vs heuristic plan, exact counts",
        "value": value,
        "unit": "embeddings/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": f"synthetic: seeded R-MAT (seed {args.seed}) generated on the host",
        "config": {
            "workload": f"R-MAT scale {args.scale} ef {args.edge_factor} (.57/.19/.19/.05), k-means k=1 single-bin "
                        f"{hop}-hop view, patterns {','.join(keys)}, {args.plan} plans"
                        + ("" if args.relabel == "none" else ", ids relabelled in ascending degree order"),
            "scale": args.scale,
            "vertices": g.vertex_count,
            "edges": g.edge_count,
            "patterns": keys,
            "counts": counts,
            "per_pattern": per_pattern,
            "plan": args.plan,
            "learned_planner": learned_info,
            "steps_requested": args.steps,
            "warmup_requested": args.warmup,
            "budget_s": args.budget_s,
            "wall_s_at_line": round(elapsed(), 1),
            "graph_gen_s": round(gen_s, 2),
            "relabel": args.relabel,
            "relabel_s": None if relabel_s is None else round(relabel_s, 2),
            "ids_as_generated": generated_ids,
            "view_build_s": round(view_s, 2),
            "transfer_s": round(xfer_s, 3),
            "index_lists": ex.hub_count(),
            "isect_kernel": args.isect_kernel,
            "cold_weight": args.cold_weight,
            "ticket_chunks": args.ticket_chunks,
            "sampled_root_parity": sample_check,
            "golden_counts": golden,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if args.detail_out:
        with open(args.detail_out, "w") as f:
            json.dump({"line": line, "kernels": kstats, "levels": level_stats}, f, indent=1)
    print(json.dumps(line), flush=True)


def make_roofline(kstats, workload, plan, isect_kernel):
    """Roofline of the kernel with the most device time: algorithmic bytes
    per launch over the average launch time against the HBM peak. When the
    algorithmic rate exceeds the peak (candidate lists re-read from the
    caches), `frac` is the measured fabric traffic rate over the peak
    instead, and `bound_note` says so; with no matching traffic profile it
    is capped at 1 and labelled as cache-served."""
    if not kstats:
        return None
    dom = max(kstats, key=lambda s: s["total_ms"])
    per_launch_bytes = dom["algo_bytes"] / dom["launches"]
    per_launch_s = dom["total_ms"] / dom["launches"] / 1000.0
    achieved = per_launch_bytes / per_launch_s / 1e9
    tr = load_traffic(dom["name"], workload, plan, isect_kernel)
    traffic = tr["bytes_per_launch"] if tr else None
    frac = achieved / PEAK_HBM_GBS
    note = "algorithmic bytes (each pivot candidate streamed once + per-path words) / launch time"
    if frac > 1.0:
        if traffic:
            frac = traffic / per_launch_s / 1e9 / PEAK_HBM_GBS
            note = "algorithmic rate exceeds HBM peak (lists re-read from cache): frac = measured fabric bytes / launch time / peak"
        else:
            frac = 1.0
            note = "algorithmic rate exceeds HBM peak (lists re-read from cache); no matching traffic profile: frac capped at 1"
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
           "frac": round(min(frac, 1.0), 4), "traffic": int(traffic) if traffic else None,
           "kernel": dom["name"], "launches": dom["launches"], "avg_launch_ms": round(dom["total_ms"] / dom["launches"], 3),
           "algo_bytes_per_launch": int(per_launch_bytes), "bound_note": note}
    if traffic:
        # bytes the launches really moved across the fabric (rocprof FETCH_SIZE + WRITE_SIZE,
        # calibrated), over the same launch time: how busy the memory system was
        out["traffic_gbs"] = round(traffic / per_launch_s / 1e9, 1)
        out["traffic_frac"] = round(traffic / per_launch_s / 1e9 / PEAK_HBM_GBS, 4)
        out["traffic_profile"] = f"profiles/traffic_{dom['name']}_{workload}_{plan}_k{isect_kernel}.json"
    return out


def learned_plans(V, args, keys, queries, g):
    """The learned planner on this workload: records from small seeded R-MAT
    graphs, a HIP-trained cost model, one order per pattern (the heuristic
    plan kept where no order is predicted to gain)."""
    import numpy as np

    from vsgm_amd import learned
    from vsgm_amd.shapes import all_shapes

    t = time.perf_counter()

    def rmat(s):  # record graphs carry the ids the workload uses
        r = V.Graph.rmat(s, 16, seed=100 + s)
        return r.relabel_by_degree()[0] if args.relabel == "degree" else r

    tq = [V.Query.from_matrix(s[2]) for s in all_shapes((3, 4, 5))]
    data = learned.collect([rmat(s) for s in (10, 11)], tq, bins=2, max_orders=24, repeats=1,
                           seed=args.seed)
    # 4-vertex shapes also on larger graphs, where order costs separate
    # from launch overheads (scale 11 times are ~1 ms)
    q4 = [V.Query.from_matrix(s[2]) for s in all_shapes((4,))]
    big = learned.collect([rmat(s) for s in (13, 14)], q4, bins=2, max_orders=24, repeats=1,
                          seed=args.seed + 1)
    big["group"] = big["group"] + int(data["group"].max()) + 1
    data = {k: np.concatenate([data[k], big[k]]) for k in data}
    # a fifth of the record groups is held out to set the margin
    train_set, held = learned.split_groups(data, holdout=0.2, seed=args.seed)
    model, losses = learned.train(train_set, epochs=60, batch=128, lr=2e-3, seed=args.seed)
    margin, table = learned.calibrate_margin(model, held)
    rp_all, _ = g.csr(copy=False)
    vstats = learned.view_stats(rp_all, g.vertex_count)
    planner = learned.LearnedPlanner(model, margin=margin)
    chosen = [planner.plan(q, vstats) for q in queries]
    # a learned order is kept only if it also beats the heuristic order on a
    # seeded sample of this graph's own roots (the records come from far
    # smaller graphs, whose costs can rank orders differently)
    roots = learned.sample_roots(g.vertex_count, seed=args.seed + 7)
    checks = {}
    for i, (k, q, (p, c)) in enumerate(zip(keys, queries, chosen)):
        if not c["learned"]:
            continue
        h = V.Plan.optimize(q)
        keep, ms_l, ms_h = learned.check_on_roots(g, p, h, roots, {"isect_kernel": args.isect_kernel,
                                                                  "cold_weight": args.cold_weight})
        checks[k] = {"roots": int(len(roots)), "learned_ms": round(ms_l, 3), "heuristic_ms": round(ms_h, 3), "kept": keep}
        if not keep:
            chosen[i] = (h, {**c, "learned": False, "rejected_on_sample": True})
    info = {"records": int(len(data["y"])), "train_records": int(len(train_set["y"])), "margin": margin,
            "sample_check": checks,
            "margin_table": table, "loss": [float(losses[0]), float(losses[-1])],
            "orders": {k: (c.get("order") if c["learned"] else "heuristic") for k, (_, c) in zip(keys, chosen)},
            "predicted_gain": {k: (round(c["pred_gain"], 4), c.get("best_order", c.get("order")))
                               for k, (_, c) in zip(keys, chosen)},
            "train_s": round(time.perf_counter() - t, 1)}
    log(f"learned planner: {info}")
    return [p for p, _ in chosen], info


def chunglu_main(args):
    """Friendster-shaped power-law graph, partitioned and packed as the
    reference's offline tools do, then matched bin by bin:
      graph      Chung-Lu with degree exponent 2.3 (weights (i+1)^(-1/1.3)),
                 --cl-vertices vertices, about --cl-edges undirected edges
                 after deduplication, seeded;
      partition  GPU k-means, k = --bins (4), kmeans++ seeding, 10 rounds,
                 over closed 1-hop neighbourhoods (kmeans.cc);
      packing    the reference's view packing with -m 10 (bins of at most
                 2.5e9 view entries) for 2-hop and 3-hop views
                 (view_packing.cc);
      views      every bin's view built on the device from the resident
                 graph, the next bin's build overlapped with the current
                 bin's match (a second context and stream);
      patterns   triangle over every source of every 2-hop bin (and over the
                 3-hop bins when the packing succeeds; where the reference's
                 packer refuses 3-hop bins -m 10, one 2-hop bin's sources
                 are also matched over their 3-hop device view); 4-clique
                 and house over a seeded sample of each 2-hop bin's sources,
                 starting at 1/4096 of them (4-clique) or 4 per bin (house)
                 and growing 4x while a pass takes under an eighth of
                 --cl-sample-s (labelled with the sample).
    Counts: the triangle totals equal the same plan over one whole-graph
    view, and all three patterns equal the oracle's on a seeded sample of
    root vertices."""
    import numpy as np

    import oracle as O
    import vsgm_amd as V
    from vsgm_amd import partition

    assert V.device_count() > 0, "no GPU visible"
    t = time.perf_counter()
    # duplicate draws between hubs and self loops are removed: about 3-5 % of
    # the draws at this exponent, so 1.05 draws per target edge
    g = V.Graph.chung_lu(args.cl_vertices, int(args.cl_edges * 1.05), 2.3, seed=args.seed)
    gen_s = time.perf_counter() - t
    log(f"Chung-Lu: V={g.vertex_count} E={g.edge_count} maxdeg={g.max_degree} generated in {gen_s:.1f}s")
    keys = ["kc3", "kc4", "p4"]
    plans = {k: V.Plan.optimize(make_query(V, k)) for k in keys}
    rd = min(p.root_degree for p in plans.values())
    hop = max(p.hop for p in plans.values())
    info = {"generate_s": round(gen_s, 1)}

    t = time.perf_counter()
    assign, km_sizes, obj = partition.kmeans(g, args.bins, iters=10, seed=args.seed)
    info["kmeans"] = {"k": args.bins, "iters": 10, "sizes": [int(x) for x in km_sizes], "seconds": round(time.perf_counter() - t, 1)}
    log(f"k-means k={args.bins}: sizes {km_sizes.tolist()} in {info['kmeans']['seconds']}s")
    packs = {}
    for h in (2, 3):
        t = time.perf_counter()
        try:
            nb, bin_of, src_n, sizes = V.pack_view_bins(g, h, assign, args.bins, max_bin_entries=int(10e9 / 4))
            packs[h] = (nb, bin_of)
            info[f"packing_{h}hop"] = {"bins": nb, "sources": [int(x) for x in src_n], "entries": [int(x) for x in sizes],
                                       "seconds": round(time.perf_counter() - t, 1)}
        except V.VsgmError as e:
            info[f"packing_{h}hop"] = {"error": str(e), "seconds": round(time.perf_counter() - t, 1)}
        log(f"packing -m 10, {h}-hop: {info[f'packing_{h}hop']}")

    def run_bins(plan_list, bins_sources, view_hop, label):
        """Every bin's view on the device (the next one prefetched while this one matches)."""
        ex = V.PipelineExecutor(0, g.vertex_count, plan_list)
        ex.set_graph(g)
        ex.profile(True)
        t0 = time.perf_counter()
        build = match = 0.0
        for i, srcs in enumerate(bins_sources):
            t = time.perf_counter()
            ex.transfer_device(srcs, view_hop, rd)
            build += time.perf_counter() - t
            if i + 1 < len(bins_sources):
                ex.prefetch_device(bins_sources[i + 1], view_hop, rd)
            t = time.perf_counter()
            ex.match()
            match += time.perf_counter() - t
        wall = time.perf_counter() - t0
        kst = ex.kernel_stats()
        counts = ex.total_counts
        log(f"{label}: {len(bins_sources)} bins in {wall:.2f}s (waiting for views {build:.2f}s, matching {match:.2f}s) counts {counts}")
        del ex
        return counts, wall, match, build, kst

    out = {}
    whole = V.PipelineExecutor(0, g.vertex_count, [plans["kc3"]])
    whole.transfer_graph(g)
    whole.match()
    tri_whole = whole.total_counts[0]
    del whole
    log(f"triangles over one whole-graph view: {tri_whole}")
    kstats_main = None
    for h in (2, 3):
        if h not in packs:
            continue
        nb, bin_of = packs[h]
        bins_sources = [np.nonzero(bin_of == b)[0].astype(np.uint32) for b in range(nb)]
        counts, wall, match, build, kst = run_bins([plans["kc3"]], bins_sources, h, f"triangle, {h}-hop bins")
        assert counts[0] == tri_whole, f"{h}-hop bins count {counts[0]} triangles, the whole graph {tri_whole}"
        out[f"triangle_{h}hop"] = {"count": counts[0], "seconds": round(wall, 3), "match_s": round(match, 3),
                                   "view_wait_s": round(build, 3), "embeddings_per_s": counts[0] / match}
        if h == 2:
            kstats_main = kst
    nb, bin_of = packs[2]
    if 3 not in packs:
        # the smallest 2-hop bin's sources over their 3-hop device view: same triangles as over the 2-hop one
        small = int(np.argmin(np.bincount(bin_of, minlength=nb)))
        srcs = np.nonzero(bin_of == small)[0].astype(np.uint32)
        c2, _, m2, _, _ = run_bins([plans["kc3"]], [srcs], 2, f"triangle, bin {small} ({len(srcs)} sources), 2-hop view")
        c3, _, m3, b3, _ = run_bins([plans["kc3"]], [srcs], 3, f"triangle, bin {small} ({len(srcs)} sources), 3-hop view")
        assert c2 == c3, f"3-hop view counts {c3}, 2-hop view {c2}"
        out["triangle_3hop_view_of_one_bin"] = {"bin": small, "sources": int(len(srcs)), "count": c3[0],
                                                "match_s": round(m3, 3), "view_s": round(b3, 3)}
    # 4-clique and house over growing samples of every 2-hop bin's sources.
    # The house sample is a fraction of each bin rounded down, so bins of a
    # few hubs (a hub root alone enumerates more houses than a pass can
    # count in minutes) join only at large fractions.
    rng = np.random.default_rng(args.seed)
    order = [rng.permutation(np.nonzero(bin_of == b)[0].astype(np.uint32)) for b in range(nb)]
    for k, first in (("kc4", lambda o: max(1, len(o) // 4096)), ("p4", lambda o: len(o) >> 20)):
        take = [first(o) for o in order]
        if sum(take) == 0:
            take[int(np.argmax([len(o) for o in order]))] = 1
        best, prev = None, None
        while True:
            sample = [np.sort(o[:n]) for o, n in zip(order, take)]
            n_src = int(sum(len(x) for x in sample))
            counts, wall, match, build, _ = run_bins([plans[k]], sample, plans[k].hop, f"{NAMES[k]}, {n_src} sampled sources")
            best = {"count": counts[0], "sample_sources": n_src, "sample_fraction": n_src / g.vertex_count,
                    "seconds": round(wall, 3), "match_s": round(match, 3), "embeddings_per_s": counts[0] / match if match else 0.0}
            # the next pass (4x the sources) is predicted to grow like the
            # last one did, at least 4x: counts grow faster than linearly
            # once the sample reaches denser parts of the graph
            grow = max(4.0, wall / prev) if prev else 4.0
            prev = max(wall, 1e-3)
            if n_src >= g.vertex_count or wall * grow > args.cl_sample_s / 2 or elapsed() + wall * grow > args.budget_s:
                break
            grown = [min(len(o), 4 * n) for o, n in zip(order, take)]
            if grown == take:  # every bin the sample draws from is whole
                break
            take = grown
        out[NAMES[k]] = best
    # parity: all three plans against the oracle on seeded roots
    roots = np.sort(rng.permutation(g.vertex_count)[: args.cl_parity_roots]).astype(np.uint32)
    rp, cl = g.csr(copy=False)
    pe = V.PipelineExecutor(0, g.vertex_count, [plans[k] for k in keys])
    pe.transfer_graph(g, roots)
    pe.match()
    want = [O.count_plan(rp, cl, plans[k].serialize(), sources=roots, threads=min(os.cpu_count() or 1, 16)) for k in keys]
    parity = {"roots": int(len(roots)), "gpu": pe.total_counts, "oracle": want, "equal": pe.total_counts == want}
    log(f"parity on {len(roots)} roots: {parity}")
    assert parity["equal"], f"GPU {pe.total_counts} vs oracle {want} on sampled roots"
    dom = max(kstats_main, key=lambda s: s["total_ms"])
    achieved = dom["algo_bytes"] / (dom["total_ms"] / 1000.0) / 1e9
    tri = out["triangle_2hop"]
    line = {
        "metric": "embeddings/sec per pattern, 1 MI355X; learned vs heuristic plan, exact counts",
        "value": tri["embeddings_per_s"], "unit": "embeddings/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": tri["match_s"] * 1000.0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": f"synthetic: seeded Chung-Lu (seed {args.seed}) generated on the host",
        "config": {"workload": f"Chung-Lu gamma 2.3, V={g.vertex_count}, E={g.edge_count}, GPU k-means k={args.bins}, "
                               f"view packing -m 10 (2-hop and 3-hop), device-built views; value: triangles over the 2-hop bins",
                   "per_pattern": out, "pipeline": info, "whole_graph_triangles": tri_whole, "sampled_root_parity": parity},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(min(achieved / PEAK_HBM_GBS, 1.0), 4), "traffic": None, "kernel": dom["name"],
                     "launches": dom["launches"], "avg_launch_ms": round(dom["total_ms"] / dom["launches"], 3)},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def triangles_cpu_main(args):
    """Triangle counting (-kc 3, heuristic plan) on R-MAT scale 16 (or
    --scale), one view bin: the GPU path's count and rate beside a full run
    of the oracle's C restatement of the level algorithm on min(32, cores)
    host threads, both checked against the textbook triangle counter."""
    import numpy as np

    import oracle as O
    import vsgm_amd as V
    from vsgm_amd import partition

    scale = args.scale if args.scale != 22 else 16
    g = V.Graph.rmat(scale, args.edge_factor, 0.57, 0.19, 0.19, seed=args.seed)
    plan = V.Plan.optimize(V.Query.clique(3))
    assign, _, _ = partition.kmeans(g, 1, iters=1, seed=args.seed)
    vbm = V.ViewBinManager.from_assignment(g, plan.hop, 1, assign)
    holder = V.ViewBinHolder(g.vertex_count, max(4, g.entries * 4))
    vbm.materialize(0, holder, plan.root_degree)
    ex = V.PipelineExecutor(0, g.vertex_count, [plan], max_sources=vbm.max_sources)
    ex.transfer(holder)
    ex.match()  # warm-up
    ex.reset()
    steps = max(1, min(args.steps, 20))
    t = time.perf_counter()
    for _ in range(steps):
        ex.match()
    gpu_s = (time.perf_counter() - t) / steps
    gpu_count = ex.total_counts[0] // steps
    rp, cl = g.csr(copy=False)
    # min(32, cores), within the host threads this process is given
    threads = min(32, os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS") or 32))
    t = time.perf_counter()
    cpu_count = O.count_plan(rp, cl, plan.serialize(), threads=threads)
    cpu_s = time.perf_counter() - t
    textbook = int(O.triangles(rp, cl))
    log(f"triangles R-MAT s{scale}: gpu {gpu_count} in {gpu_s * 1000:.2f} ms/step, cpu {cpu_count} in {cpu_s:.2f}s "
        f"({threads} threads), textbook {textbook}")
    assert gpu_count == cpu_count == textbook, (gpu_count, cpu_count, textbook)
    line = {
        "metric": "embeddings/sec per pattern, 1 MI355X; learned vs heuristic plan, exact counts",
        "value": gpu_count / gpu_s, "unit": "embeddings/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": gpu_s * 1000.0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": f"synthetic: seeded R-MAT (seed {args.seed}) generated on the host",
        "config": {"workload": f"triangles (-kc 3), heuristic plan, R-MAT scale {scale} ef {args.edge_factor} "
                               f"(.57/.19/.19/.05), one view bin; the CPU path timed beside it",
                   "count": gpu_count, "textbook_count": textbook, "vertices": g.vertex_count, "edges": g.edge_count},
        "roofline": None,
        "cpu_baseline": {"value": cpu_count / cpu_s, "unit": "embeddings/s", "cores": threads, "kind": "port",
                         "sample": f"full run: every root, {cpu_count} triangles in {cpu_s:.2f} s"},
    }
    print(json.dumps(line), flush=True)


def calibration_main():
    """k_fill writes and k_reduce reads exactly 2 GiB (2^28 u64) per
    dispatch, three dispatches each: what rocprof's WRITE_SIZE / FETCH_SIZE
    report for them, divided by 2 GiB, is the factor that turns those
    counters into bytes for the other kernels."""
    import numpy as np

    from vsgm_amd.hip import DeviceContext, hiplib

    n = 1 << 28
    ctx = DeviceContext(0)
    buf = ctx.empty(n, np.uint64)
    tot = ctx.array(np.zeros(1, np.uint64))
    for _ in range(3):
        ctx.check(hiplib().vsgm_fill_u64(ctx.h, buf.ptr, n, 3), "fill")
        ctx.sync()
    for _ in range(3):
        ctx.check(hiplib().vsgm_reduce_u64(ctx.h, buf.ptr, n, tot.ptr), "reduce")
        ctx.sync()
    print(json.dumps({"calibration": "k_fill writes / k_reduce reads", "bytes_per_dispatch": n * 8,
                      "sum_check": int(tot.numpy()[0]) == 3 * 3 * n}), flush=True)


def batched_main(args):
    """Many patterns over k view bins from a GPU k-means partition (the
    batched configuration):
      plain    one pattern at a time, one bin at a time, heuristic plans
               (host-built view of each bin transferred, then Match one plan
               after another);
      batched  all bins' views built on the device and resident, every plan
               autotuned, the plans run as a prefix tree (levels several
               patterns reach identically run once, --share-prefixes), the
               whole enumeration captured once into one device graph and
               replayed;
      learned  as batched, with matching orders from the learned planner
               (trained here on records from smaller seeded R-MAT graphs).
    Every count is checked equal across the three. Times are the match phase
    (host wall clock around Match calls / the graph launch)."""
    import numpy as np

    import vsgm_amd as V
    from vsgm_amd import autotune, learned
    from vsgm_amd.shapes import all_shapes

    assert V.device_count() > 0, "no GPU visible"
    rng = np.random.default_rng(args.seed)
    shapes = all_shapes()
    pick = rng.integers(0, len(shapes), size=args.n_patterns)  # uniform over the non-isomorphic shapes
    queries = [V.Query.from_matrix(shapes[i][2]) for i in pick]
    heur = [V.Plan.optimize(q) for q in queries]

    t = time.perf_counter()
    g = V.Graph.rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, seed=args.seed)
    log(f"R-MAT s{args.scale}: V={g.vertex_count} E={g.edge_count} maxdeg={g.max_degree} in {time.perf_counter() - t:.1f}s; "
        f"{len(queries)} patterns of sizes {np.bincount([shapes[i][0] for i in pick]).tolist()}")

    # learned planner: records on smaller graphs, 3..5-vertex shapes
    t = time.perf_counter()
    six = all_shapes((6,))
    tq = [V.Query.from_matrix(s[2]) for s in all_shapes((3, 4, 5))]
    tq += [V.Query.from_matrix(six[i][2]) for i in np.random.default_rng(args.seed + 1).choice(len(six), 24, replace=False)]
    data = learned.collect([V.Graph.rmat(s, 16, seed=100 + s) for s in (8, 9)], tq, bins=2, max_orders=12, repeats=1,
                           seed=args.seed)
    train_set, held = learned.split_groups(data, holdout=0.2, seed=args.seed)
    model, losses = learned.train(train_set, epochs=60, batch=128, lr=2e-3, seed=args.seed)
    margin, _ = learned.calibrate_margin(model, held)
    log(f"learned planner: {len(data['y'])} records, loss {losses[0]:.3f} -> {losses[-1]:.3f}, held-out margin {margin} "
        f"in {time.perf_counter() - t:.1f}s")

    # the bins: k-means on the GPU over closed 1-hop neighbourhoods
    from vsgm_amd import partition

    t = time.perf_counter()
    assign, km_sizes, km_obj = partition.kmeans(g, args.bins, iters=10, seed=args.seed)
    log(f"k-means k={args.bins}: sizes {km_sizes.tolist()}, objective {km_obj[0]:.1f} -> {km_obj[-1]:.1f} "
        f"in {time.perf_counter() - t:.2f}s")

    def make_bins(plans):
        hop = max(p.hop for p in plans)
        rd = min(p.root_degree for p in plans)
        vbm = V.ViewBinManager.from_assignment(g, hop, args.bins, assign)
        holders = []
        for b in range(vbm.n_bins):
            h = V.ViewBinHolder(g.vertex_count, g.entries * 4 + 4)
            vbm.materialize(b, h, rd)
            holders.append(h)
        return vbm, holders

    planner = learned.LearnedPlanner(model, margin=margin)
    vbm0, holders0 = make_bins(heur)
    rp0, _, src0 = holders0[0].csr(g.vertex_count)
    vstats = learned.view_stats(rp0, len(src0))
    lplans, n_learned = [], 0
    for q in queries:
        p, info = planner.plan(q, vstats)
        lplans.append(p)
        n_learned += int(info["learned"])
    log(f"learned planner picked a non-heuristic order for {n_learned} of {len(queries)} patterns")
    vbm, holders = make_bins(heur + lplans)  # one view family serves every plan

    # plain: one bin at a time, one plan per Match
    ex = V.PipelineExecutor(0, g.vertex_count, heur, max_sources=vbm.max_sources)
    plain_s = 0.0
    for b, h in enumerate(holders):
        ex.transfer(h)
        for i in range(len(heur)):
            t = time.perf_counter()
            ex.match_all([i])
            plain_s += time.perf_counter() - t
        log(f"plain: bin {b} done, {plain_s:.2f}s so far")
    plain_counts = ex.total_counts
    del ex

    # batched modes build the same bins' views on the device from the resident graph
    all_plans = heur + lplans
    v_hop = max(p.hop for p in all_plans)
    v_rd = min(p.root_degree for p in all_plans)
    bin_sources = [h.csr(g.vertex_count)[2] for h in holders]

    def batched(plans, label):
        # tune on the first bin alone (a sample of the roots), then keep every bin resident
        # with every bin resident and one graph for all plans, a fixed 2 GiB
        # level budget (0 in the tuning below) keeps the pool from fragmenting
        tx = V.PipelineExecutor(0, g.vertex_count, plans, max_sources=vbm.max_sources)
        tx.set_option("batch_bytes", 2 << 30)
        tx.set_graph(g)
        tx.add_device_view(bin_sources[0], v_hop, v_rd)
        t = time.perf_counter()
        res = autotune.tune(tx, repeats=1, candidates={"isect_kernel": [4, 3, 0], "max_grid": [8192], "batch_bytes": [0, 1 << 30],
                                                       "ticket_chunks": [16, 64]})
        tune_s = time.perf_counter() - t
        del tx
        bx = V.PipelineExecutor(0, g.vertex_count, plans, max_sources=vbm.max_sources)
        bx.set_option("batch_bytes", 2 << 30)
        bx.set_option("share_prefixes", int(args.share_prefixes))
        bx.set_graph(g)
        t = time.perf_counter()
        for s in bin_sources:
            bx.add_device_view(s, v_hop, v_rd)
        log(f"{label}: {len(bin_sources)} device views in {time.perf_counter() - t:.3f}s")
        for i, r in res.items():
            bx.set_plan_tuning(i, **r["config"])
        t = time.perf_counter()
        nodes = bx.capture()
        cap_s = time.perf_counter() - t
        bx.reset()
        t = time.perf_counter()
        bx.replay()
        run_s = time.perf_counter() - t
        counts = bx.total_counts
        log(f"{label}: tuned in {tune_s:.1f}s ({sum(r['config'] != {'isect_kernel': 4, 'max_grid': 8192, 'batch_bytes': 0, 'cold_weight': 16, 'ticket_chunks': 64} for r in res.values())} "
            f"plans off the default), captured {nodes} nodes in {cap_s:.1f}s, replay {run_s:.3f}s")
        return counts, run_s, nodes

    b_counts, batched_s, b_nodes = batched(heur, "batched")
    l_counts, learned_s, l_nodes = batched(lplans, "learned+batched")
    assert b_counts == plain_counts, "batched counts differ from the one-at-a-time counts"
    assert l_counts == plain_counts, "learned-plan counts differ from the heuristic counts"
    total = sum(plain_counts)
    line = {
        "metric": "embeddings/sec per pattern, 1 MI355X; learned vs heuristic plan, exact counts",
        "value": total / learned_s,
        "unit": "embeddings/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": learned_s * 1000.0,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": f"synthetic: seeded R-MAT (seed {args.seed}) generated on the host",
        "config": {
            "workload": f"batched: {len(queries)} random connected 3-6-vertex patterns (seed {args.seed}) over "
                        f"{vbm.n_bins} k-means view bins ({v_hop}-hop) of R-MAT scale {args.scale} ef {args.edge_factor}",
            "scale": args.scale, "bins": vbm.n_bins, "patterns": len(queries),
            "total_embeddings": total,
            "plain_s": plain_s, "batched_s": batched_s, "learned_batched_s": learned_s,
            "speedup_batched": plain_s / batched_s, "speedup_learned_batched": plain_s / learned_s,
            "graph_nodes": [b_nodes, l_nodes], "learned_orders": n_learned,
            "counts_identical": True,
        },
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
