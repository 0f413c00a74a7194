"""Benchmark: retro-edges per second of the plain k-cop solve on one MI355X.

Workload (default): planar200_k4 -- the Delaunay triangulation of 200 seeded
uniform points in the unit square, thinned to 400 edges while connected
(tests/golden/matrices/planar200_e400_s1.txt), k = 4 cops, full information
(p = 1): C(203, 4) = 68,685,050 configurations, 2.75e10 states (27.5 GB
table resident in HBM). Other workloads: er200_k3 (ER G(200, 0.03), first
connected seed >= 1, k = 3, 5.41e8 states), the limited-visibility
grid8x8_k2_p3, and batch4096. A step is one full solve with the tables
resident in HBM: initializeCaptures + mainLoop (every depth level until the
fixed point). The device tables are allocated once, before the timed region.

Unit of work: a retro-edge, one predecessor relaxation of the push
formulation, E_retro = E_cop + E_rob (won robber-turn states x |pred(C')| +
won cop-turn states x in-degree(v), depth <= 125). value = E_retro * steps /
elapsed, in GTEPS. The relaxations are counted from the solved table; how a
level performs them (push, factorised pull, or the closed form of depth 1)
does not change the count. `performed_gteps` leaves out the relaxations of
depth-0 states, which the closed-form initialisation replaces.

roofline: the dominant kernel's algorithmic bytes over its own HIP-event time
(events on the solver's stream around each of its launches in the last timed
step; a single-kernel launch such as the push records them itself through
hipExtLaunchKernelGGL), against 8 TB/s. The dominant kernel and the per-kernel
breakdown (kernels_ms_per_step) come from one untimed solve with every launch
timed: a timed dispatch costs ~7 us more, so only the last timed step carries
events, and only around the dominant kernel.
Algorithmic bytes per unit (scalar-record model): 0.5 B per cop retro-edge
(2-bit outcome read + written) for the cop-turn kernels of the levels they
served, 2 B per robber retro-edge (8-bit escape counter read + written) for
the robber step, 1.25 B per newly won cop-turn state (depth byte + frontier
bit) for the finaliser. solve_roofline is the whole solve's algorithmic bytes
(B_alg) over the step time. roofline.traffic is the dominant kernel's measured
HBM bytes per dispatch (rocprofv3 FETCH_SIZE/WRITE_SIZE in profiles/, scaled by
a device copy of known size), and roofline.traffic_rate those bytes over the
profiled dispatch time: what the kernel moves, beside the model's bytes.

cpu_baseline: the reference solver (oracle/_ref/k_cops_visibility, built from
the reference's own sources) on one host core, on a bounded sample of the
workload -- the same graph with fewer cops, the most that keep the sample to
at most 50,000 configurations (k = 2 for n = 200, ~2 s) -- as GTEPS of that
run. cpu_port: the oracle's multithreaded level-synchronous restatement on
the host's cores. Up to 2e9 states it solves the whole workload and its
table must equal the device's (the run fails otherwise); above that it runs
under a time cap (--cpu-port-cap seconds) and reports the depth levels it
finished and its rate over the relaxations it performed. reference_mt: the
reference's own multithreaded solver (legacy/src/k_cops_4.cpp, stay-move
rules, no depths) on the same graph at k = 3, wall time, for context only.

batch (default line only): batch4096 -- one persistent launch over 4096 small
graphs, the plain per-graph loop on a sample, and a Poisson request stream
through the continuous-batching server (p50/p99 latency).
"""
import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK = 8.0e12
WORKLOADS = {
    "er200_k3": ("er200_p0.03_s1", 3),
    "er200_k2": ("er200_p0.03_s1", 2),
    "planar200_k4": ("planar200_e400_s1", 4),
    "sy_all_k2": ("scotlandyard-all", 2),
}
# limited-visibility workloads: (graph, k, p)
MACRO_WORKLOADS = {
    "grid8x8_k2_p3": ("grid8x8_chords16_s1", 2, 3),
    "grid8x8_k2_p2": ("grid8x8_chords16_s1", 2, 2),
}
# solver timer name -> device kernel name in rocprofv3 summaries
KERNEL_NAMES = {"cop_push": "k_cop_push", "sep_stage": "k_pull_tile + k_pull_tile1 + k_pull_pair", "robber_step": "k_robber_step",
                "cop_finalize": "k_cop_finalize", "clear_frontier": "k_clear_frontier",
                "mx_expand": "k_mx_expand", "mx_emit": "k_mx_expand", "mx_insert": "k_mx_insert",
                "mx_compact": "k_mx_compact", "mx_leaf": "k_mx_leaf", "mx_up": "k_mx_up",
                "mx_planes": "k_mx_planes", "mx_mark": "k_mx_mark"}


def traffic_profile(workload, timer_name):
    """(HBM bytes, duration in us) per dispatch of a kernel from the committed rocprofv3
    summaries (profiles/<workload>_traffic.json: FETCH_SIZE/WRITE_SIZE scaled by the copy
    calibration, durations from the kernel trace of the same command), or (None, None)."""
    path = os.path.join(REPO, "profiles", f"{workload}_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        data = json.load(f)
    kernels = data.get("kernels", {})
    # a timer can cover several device kernels (the pull stages): dispatch-weighted means
    found = [kernels[n] for n in KERNEL_NAMES.get(timer_name, timer_name).split(" + ") if n in kernels]
    if not found:
        return None, None
    calls = sum(k.get("dispatches", 1) for k in found)
    hbm = sum(k["hbm_bytes"] * k.get("dispatches", 1) for k in found) / calls
    timed = [k for k in found if k.get("avg_us")]
    us = (sum(k["avg_us"] * k.get("dispatches", 1) for k in timed) / sum(k.get("dispatches", 1) for k in timed)
          if timed else None)
    return hbm, us


def measured_traffic(workload, timer_name):
    """HBM bytes per dispatch of a kernel (traffic_profile)."""
    return traffic_profile(workload, timer_name)[0]


def traffic_rate(workload, timer_name):
    """Measured HBM bytes per dispatch over the profiled dispatch time, GB/s and as a
    fraction of the 8 TB/s peak: what the kernel moves, beside the model's bytes."""
    hbm, us = traffic_profile(workload, timer_name)
    if not hbm or not us:
        return None
    rate = hbm / (us * 1e-6)
    return {"GBps": rate / 1e9, "frac": rate / HBM_PEAK}


def matrix_path(stem):
    return os.path.join(REPO, "tests", "golden", "matrices", stem + ".txt")


def cpu_reference_sample(stem, k, e_retro):
    """Reference binary (single thread) on a bounded sample; GTEPS from its Main Loop time."""
    exe = os.path.join(REPO, "oracle", "_ref", "k_cops_visibility")
    if not os.path.exists(exe):
        return None
    work = tempfile.mkdtemp(prefix="crbench_")
    try:
        proc = subprocess.run([exe, matrix_path(stem), str(k), "1"], cwd=work, capture_output=True,
                              text=True, timeout=600)
    finally:
        shutil.rmtree(work, ignore_errors=True)
    m = re.search(r"Main Loop\s*-*=>\s*([\d.]+)\s*s", proc.stdout)
    if not m:
        return None
    secs = float(m.group(1))
    return {"value": e_retro / secs / 1e9, "unit": "GTEPS", "cores": 1, "kind": "reference",
            "sample": f"reference k_cops_visibility, same graph at k={k} ({stem}), main loop "
                      f"{secs:.2f} s for {e_retro} retro-edges"}


def progress(*args):
    """One line per finished phase on stderr (stdout carries only the JSON line)."""
    print("bench:", *args, file=sys.stderr, flush=True)


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def cpu_port(stem, k, threads, cap_s=None):
    """The oracle's level-synchronous restatement with on-the-fly moves: the whole
    workload (cap_s None), or until cap_s seconds have passed."""
    from oracle import binding as oracle
    n, adj = oracle.load_matrix(matrix_path(stem))
    states, info = oracle.v1_levelsync_capped(adj, k, threads, 1e9 if cap_s is None else cap_s)
    done = info["level_fraction"] == 0.0 and cap_s is None
    rate = (info["e_cop"] + info["e_rob"]) / info["elapsed_s"] / 1e9
    if cap_s is None:
        sample = f"multithreaded level-synchronous oracle, full workload, {info['elapsed_s']:.2f} s"
    else:
        sample = (f"multithreaded level-synchronous oracle, full workload under a {cap_s:.0f} s cap: captures and "
                  f"{info['levels_done']} depth levels done, {100 * info['level_fraction']:.0f} % of the next; "
                  f"rate over the {info['e_cop'] + info['e_rob']} relaxations performed in {info['elapsed_s']:.1f} s "
                  f"(capture initialisation included)")
    return {"value": rate, "unit": "GTEPS", "cores": threads, "kind": "port", "sample": sample,
            "levels_done": info["levels_done"], "complete": done}, states


def reference_mt(stem, k, threads, timeout=100):
    """The reference's multithreaded solver (legacy/src/k_cops_4.cpp, built by
    oracle/build_ref.sh): stay-move rules and no depths, so a different
    workload; wall time only, for context."""
    exe = os.path.join(REPO, "oracle", "_ref", "k_cops_4")
    if not os.path.exists(exe):
        return None
    work = tempfile.mkdtemp(prefix="crbench_")
    try:
        t0 = time.perf_counter()
        proc = subprocess.run([exe, matrix_path(stem), str(k)], cwd=work, capture_output=True, text=True,
                              timeout=timeout)
        dt = time.perf_counter() - t0
    except subprocess.TimeoutExpired:
        return {"skipped": f"over {timeout} s"}
    finally:
        shutil.rmtree(work, ignore_errors=True)
    m = re.search(r"RESULT: (WIN|LOSS)", proc.stdout)
    return {"wall_s": dt, "threads": threads, "graph": stem, "k": k, "result": m.group(1) if m else None,
            "note": "reference legacy/src/k_cops_4.cpp (hardware_concurrency threads, stay moves for cops and "
                    "robber, win/loss only): a different rule set, timed for context, not a baseline"}


def copy_calibration(nbytes):
    """A plain device-to-device copy of known size: FETCH_SIZE/WRITE_SIZE of this
    kernel give the factor between counter units and bytes."""
    import torch
    x = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "copies": 3}))
    return 0


def cpu_reference_macro(stem, k, p, units):
    """Reference v2 binary (single thread); rate in this solver's units of the same solve."""
    exe = os.path.join(REPO, "oracle", "_ref", "k_cops_visibility_v2")
    if not os.path.exists(exe):
        return None
    work = tempfile.mkdtemp(prefix="crbench_")
    try:
        proc = subprocess.run([exe, matrix_path(stem), str(k), str(p)], cwd=work, capture_output=True,
                              text=True, timeout=600)
    finally:
        shutil.rmtree(work, ignore_errors=True)
    m = re.search(r"Main Loop\s*-*=>\s*([\d.]+)\s*s", proc.stdout)
    if not m:
        return None
    secs = float(m.group(1))
    return {"value": units / secs / 1e9, "unit": "G macro-edges/s", "cores": 1, "kind": "reference",
            "sample": f"reference k_cops_visibility_v2, same graph at p={p} ({stem}, k={k}), main loop "
                      f"{secs:.2f} s; units = this solver's macro-edges for that solve ({units})"}


def bench_macro(args):
    """Limited-visibility solve (k_cops_visibility_v2.cpp). A step = initializeCaptures + mainLoop
    (trajectory DAG build + all snapshot passes). Unit: a macro-edge, one (DAG node, team move)
    evaluation -- enumerated while building the DAG, at the last cop step of every pass, or a DAG
    edge read on the way up in every pass."""
    import torch

    import copsandrobbers_amd as cra
    stem, k, p = MACRO_WORKLOADS[args.workload]
    s = cra.Solver(0)
    s.load_graph_file(matrix_path(stem), k, p, cra.VARIANT_MACRO)
    s.build_aux_graph()
    info = s.info()
    for _ in range(args.warmup):
        s.initialize_captures()
        s.main_loop()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.initialize_captures()
        passes = s.main_loop()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    m = s.macro_stats()
    units = m["build_moves"] + m["leaf_moves"] + m["up_edges"]
    # kernel split, from a separate timed run (events add host overhead)
    s.set_kernel_timing(True)
    s.initialize_captures()
    s.main_loop()
    kt = s.kernel_times()
    s.set_kernel_timing(False)
    verdict = s.find_final_result()
    W = (info.n + 63) // 64
    key = 4 + 8 * W  # configuration rank + cloud words of one DAG node
    nodes = m["nodes_per_step"]
    leaf_nodes, inner_nodes = nodes[-1], sum(nodes[:-1])
    E = m["dag_edges"]
    npass = len(passes)
    alg = {  # scalar-record bytes per solve and kernel
        "mx_expand": inner_nodes * (key + 5),          # node key read, value + count written
        "mx_emit": inner_nodes * key + E * key,        # node key read, candidate key written
        "mx_insert": E * (key + 8),                    # candidate key read, slot CAS, owner written
        "mx_compact": E * 16 + sum(nodes[1:]) * key,   # owner + position read, edge written, node key copied
        "mx_leaf": npass * leaf_nodes * (key + 1) + m["leaf_moves"] * info.n / 8,  # node + N-bit row per move
        "mx_up": npass * (E * 5 + inner_nodes * 9),    # edge + child value read, offsets read, value written
        "mx_planes": npass * info.num_states * 2,      # snapshot row byte read, plane bits written
        "mx_mark": npass * info.num_states * 2,        # value read, table byte read (and written once)
    }
    dominant = max(kt, key=lambda n: kt[n][0])
    rl = None
    if dominant in alg:
        ms, launches = kt[dominant]
        per_launch = alg[dominant] / launches
        achieved = per_launch / (ms / 1e3 / launches)
        rl = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
              "frac": achieved / HBM_PEAK, "traffic": measured_traffic(args.workload, dominant),
              "traffic_rate": traffic_rate(args.workload, dominant),
              "kernel": KERNEL_NAMES[dominant], "alg_bytes_per_launch": per_launch,
              "avg_launch_us": ms * 1e3 / launches}
    out = {
        "metric": "macro-edges/s (limited visibility)", "value": units * args.steps / elapsed / 1e9,
        "unit": "G macro-edges/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8/u64 bitsets (integer)",
        "data": "synthetic (seeded graph generator, committed matrix)",
        "config": {"workload": args.workload, "graph": stem, "n": info.n, "k": k, "p": p,
                   "configs": info.config_count, "states": info.num_states, "passes": len(passes),
                   "verdict": "WIN" if verdict else "LOSS", "dag_nodes_per_step": m["nodes_per_step"],
                   "dag_edges_per_step": m["edges_per_step"], "macro_edges": units,
                   "build_ms": m["build_ms"], "passes_ms": m["passes_ms"]},
        "roofline": rl,
        "kernels_ms_per_step": {n: v[0] for n, v in kt.items()},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        # bounded sample: the reference at p - 1 on the same graph (p = 3 takes ~2 min there)
        sp = max(1, p - 1)
        with cra.Solver(0) as t:
            t.load_graph_file(matrix_path(stem), k, sp, cra.VARIANT_MACRO)
            t.build_aux_graph()
            t.initialize_captures()
            t.main_loop()
            tm = t.macro_stats()
        out["cpu_baseline"] = cpu_reference_macro(stem, k, sp, tm["build_moves"] + tm["leaf_moves"] + tm["up_edges"])
    s.close()
    print(json.dumps(out))
    return 0


def run_batch(steps, warmup, plain_sample, serve_load, autotune_now=False, db=None):
    """4096 seeded connected graphs (n uniform in [8, 48], edge probability in
    [0.08, 0.3], k in {1, 2, 3}), full information. A step = one
    cr_batch_solve over all of them (one persistent launch; tables stay in
    HBM, verdicts come back), timed on the host clock around the call, with
    the graphs packed once into the arrays the call reads and the results
    returned as a record array (the per-graph Python dicts of the default
    result form cost about as much host time as the launch itself). Beside
    it: the plain six-step single-graph loop on a sample of the same graphs,
    and a Poisson request stream through the continuous-batching server.
    Returns (block, graphs, batch results, sample indices)."""
    import threading

    import numpy as np
    import torch

    import copsandrobbers_amd as cra
    from copsandrobbers_amd import batch as crb
    from copsandrobbers_amd import workloads

    from copsandrobbers_amd import autotune
    graphs = workloads.batch_graphs(4096)
    states = sum(2 * __import__("math").comb(a.shape[0] + k - 1, k) * a.shape[0] for a, k in graphs)
    b = crb.Batch(0)
    key = autotune.batch_key(graphs)
    tuning = {"key": key, "source": "defaults", "settings": {}}
    if autotune_now:
        best, t_best, t_def, evals = autotune.tune_batch(b, graphs, log=progress)
        autotune.save_entry(key, autotune.entry(best, t_best, t_def, evals, "MI355X"), db)
        tuning.update(source="autotuned in this run", settings=best, tuned_ms=t_best * 1e3, default_ms=t_def * 1e3)
    else:
        stored = autotune.load_db(db).get(key)
        if stored is not None:
            tuning.update(source="tuning database", settings=stored["settings"], tuned_ms=stored.get("best_ms"),
                          default_ms=stored.get("default_ms"))
    b.set_tuning(**tuning["settings"])
    packed = crb.pack(graphs)
    ref = b.solve(graphs)  # result dicts, for the checks below
    for _ in range(warmup):
        b.solve(packed, results="array")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev_ms = 0.0
    for _ in range(steps):
        arr = b.solve(packed, results="array")
        dev_ms += b.last_timing()[0]
    elapsed = time.perf_counter() - t0
    if not (arr["table_hash"] == np.array([r["table_hash"] for r in ref], np.uint64)).all():
        raise SystemExit("bench: a timed batch launch's tables differ from the first launch's")
    grid = b.last_timing()[1]
    b.close()
    batched = len(graphs) * steps / elapsed
    # the plain loop: the six steps per graph on an evenly spaced sample
    idx = np.linspace(0, len(graphs) - 1, min(plain_sample, len(graphs))).astype(int)
    s = cra.Solver(0)
    s.load_graph_matrix(graphs[0][0], graphs[0][1], 1)
    s.build_aux_graph()
    t1 = time.perf_counter()
    for i in idx:
        adj, k = graphs[i]
        s.load_graph_matrix(adj, k, 1)
        s.build_aux_graph()
        s.initialize_captures()
        s.main_loop()
        v = s.find_final_result()
        assert (v is not None) == ref[i]["win"]
    plain = len(idx) / (time.perf_counter() - t1)
    s.close()
    # serving: Poisson arrivals at a fraction of the batched rate
    rate = serve_load * batched
    rng = np.random.default_rng(12345)
    nreq = len(graphs)
    gaps = rng.exponential(1.0 / rate, nreq)
    order = rng.permutation(len(graphs))[:nreq]
    results = {}
    with crb.Server(0, hbm_budget=64 << 30, max_batch=1024, streams=4) as srv:
        srv.set_tuning(**tuning["settings"])
        # warm-up burst (untimed): every launch slot grows its buffers to working size
        srv.submit_many(graphs)
        warm = 0
        while warm < len(graphs):
            warm += len(srv.poll(wait_us=1000))
        warm_launches = srv.stats()["launches"]
        start = crb.clock_ns() + 2_000_000
        arrivals = (start + (np.cumsum(gaps) * 1e9)).astype(np.uint64)
        tickets = {}

        def produce():
            # hands every request whose arrival time has passed to the server in one call
            j = 0
            while j < nreq:
                now = crb.clock_ns()
                hi = int(np.searchsorted(arrivals, now, side="right"))
                if hi <= j:
                    continue
                ts, _ = srv.submit_many([graphs[gi] for gi in order[j:hi]], arrivals[j:hi])
                for t, gi in zip(ts, order[j:hi]):
                    tickets[t] = gi
                j = hi

        th = threading.Thread(target=produce)
        th.start()
        while len(results) < nreq:
            for r in srv.poll(wait_us=500):
                results[r["ticket"]] = r
        th.join()
        t_end = crb.clock_ns()
        last_arrival = int(arrivals[-1])
        sstats = srv.stats()
        sstats["launches"] -= warm_launches
    lat = np.array([r["latency_ms"] for r in results.values()])
    mismatch = sum(1 for t, r in results.items() if r["table_hash"] != ref[tickets[t]]["table_hash"])
    # algorithmic bytes of the batch: the plain solve's scalar-record model per graph
    # (2 B per robber retro-edge, 0.5 B per cop retro-edge, 1.25 B per won state,
    # 1.75 B per state of initialisation), from each graph's exact relaxation counts
    alg = 0.0
    with cra.Solver(0) as t:
        for adj, k in graphs:
            t.load_graph_matrix(adj, k, 1)
            t.build_aux_graph()
            t.initialize_captures()
            t.main_loop()
            ec, er, won = t.retro_edges()
            alg += 2 * er + 0.5 * ec + 1.25 * won + 1.75 * t.info().num_states
    dev_s = dev_ms / 1e3 / steps
    block = {
        "graphs_per_s": batched, "ms_per_step": elapsed * 1e3 / steps, "device_ms_per_step": dev_ms / steps,
        "device_graphs_per_s": len(graphs) / (dev_ms / 1e3 / steps),
        "graphs": len(graphs), "states": states, "table_bytes": states, "workgroups": grid, "tuning": tuning,
        "roofline": {"bound": "hbm", "achieved": alg / dev_s / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": alg / dev_s / HBM_PEAK,
                     "traffic": measured_traffic("batch4096", "k_batch_solve"),
                     "traffic_rate": traffic_rate("batch4096", "k_batch_solve"), "kernel": "k_batch_solve",
                     "alg_bytes_per_launch": alg, "avg_launch_us": dev_s * 1e6,
                     "note": "achieved counts the scalar-record model's bytes (0.5 B per cop retro-edge, 2 B "
                             "per robber retro-edge), as for the single-graph line. Here it can exceed the HBM peak: "
                             "each graph's work planes stay in the caches for most of its levels and the per-graph "
                             "factorised pull does not visit each cop retro-edge, so those bytes are not moved. "
                             "traffic and traffic_rate are what the launches move (rocprofv3, mean over the "
                             "profiled run's full launches and the server's micro-batches)"},
        "plain_loop": {"graphs_per_s": plain, "sample": int(len(idx)), "speedup_batched": batched / plain},
        "serving": {"arrival_rate_per_s": rate, "requests": nreq,
                    "served_per_s": nreq / ((t_end - start) / 1e9),
                    "drain_ms_after_last_arrival": (t_end - last_arrival) / 1e6,
                    "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                    "launches": sstats["launches"], "peak_admitted_bytes": sstats["peak_bytes"],
                    "results_differing_from_batch": mismatch},
    }
    return block, graphs, ref, idx


def bench_batch(args):
    block, graphs, ref, idx = run_batch(args.steps, args.warmup, args.plain_sample, args.serve_load, args.autotune,
                                        args.tuning_db)
    out = {
        "metric": "batched graphs/s (4096 small graphs, one launch)", "value": block["graphs_per_s"],
        "unit": "graphs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": block["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8/u64 bitsets (integer)", "data": "synthetic (seeded graph generator)",
        "config": {"workload": "batch4096", "graphs": block["graphs"], "states": block["states"],
                   "workgroups": block["workgroups"], "device_ms_per_step": block["device_ms_per_step"],
                   "tuning": block["tuning"]},
        "roofline": block["roofline"], "plain_loop": block["plain_loop"], "serving": block["serving"],
        "host": host_info(), "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_reference_batch(graphs, idx[:: max(1, len(idx) // 32)])
    print(json.dumps(out))
    return 0


def cpu_reference_batch(graphs, sample, budget_s=60.0):
    """The reference binary, one process per graph as its harness runs it, on
    a sample of the batch (stops after budget_s); graphs/s of the graphs done."""
    exe = os.path.join(REPO, "oracle", "_ref", "k_cops_visibility")
    if not os.path.exists(exe):
        return None
    from copsandrobbers_amd import workloads
    work = tempfile.mkdtemp(prefix="crbatch_")
    done, t_total = 0, 0.0
    try:
        for i in sample:
            adj, k = graphs[i]
            path = os.path.join(work, f"g{i}.txt")
            workloads.write_matrix(path, adj)
            t0 = time.perf_counter()
            proc = subprocess.run([exe, path, str(k), "1"], cwd=work, capture_output=True, text=True, timeout=300)
            t_total += time.perf_counter() - t0
            if "RESULT" not in proc.stdout:
                return None
            done += 1
            if t_total > budget_s:
                break
    finally:
        shutil.rmtree(work, ignore_errors=True)
    return {"value": done / t_total, "unit": "graphs/s", "cores": 1, "kind": "reference",
            "sample": f"reference k_cops_visibility, one process per graph, {done} graphs evenly spaced "
                      f"over the batch, {t_total:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="planar200_k4",
                    choices=sorted(WORKLOADS) + sorted(MACRO_WORKLOADS) + ["batch4096"])
    ap.add_argument("--plain-sample", type=int, default=256,
                    help="batch4096: graphs timed through the plain per-graph loop")
    ap.add_argument("--serve-load", type=float, default=0.7,
                    help="batch4096: Poisson arrival rate as a fraction of the batched throughput")
    ap.add_argument("--no-batch", action="store_true", help="plain workloads: leave out the batch block")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-port", type=int, default=-1,
                    help="threads for the oracle-port timing (default: min(16, cores); 0 = skip)")
    ap.add_argument("--cpu-port-cap", type=float, default=20.0,
                    help="seconds the CPU port may run on workloads above 2e9 states")
    ap.add_argument("--autotune", action="store_true",
                    help="run the autotuner for this workload (and the batch block) before timing and store its "
                         "choice in the tuning database")
    ap.add_argument("--tuning-db", default=None, help="tuning database (default: the package's)")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="plain workloads: a solver speed setting (copsandrobbers_amd.TUNING names)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (rocprof counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.gpus != 1:
        print(json.dumps({"error": "this solver runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "design-only)"}))
        return 2

    if args.workload in MACRO_WORKLOADS:
        return bench_macro(args)
    if args.workload == "batch4096":
        return bench_batch(args)
    return bench_plain(args)


def bench_plain(args):
    import torch

    import copsandrobbers_amd as cra

    stem, k = WORKLOADS[args.workload]
    path = matrix_path(stem)
    from copsandrobbers_amd import autotune
    s = cra.Solver(0)
    s.load_graph_file(path, k, 1)
    s.build_aux_graph()
    info = s.info()
    with open(path, "rb") as f:
        edges = int(cra.parse_matrix(f.read())[1].sum())
    key = autotune.plain_entry_key(s, edges)
    tuning = {"key": key, "source": "defaults", "settings": {}}
    if args.autotune:
        best, t_best, t_def, evals = autotune.tune_plain(s, log=progress)
        autotune.save_entry(key, autotune.entry(best, t_best, t_def, evals, "MI355X"), args.tuning_db)
        tuning.update(source="autotuned in this run", settings=best, tuned_ms=t_best * 1e3,
                      default_ms=t_def * 1e3, evaluations=len(evals))
    else:
        stored = autotune.load_db(args.tuning_db).get(key)
        if stored is not None:
            tuning.update(source="tuning database", settings=stored["settings"],
                          tuned_ms=stored.get("best_ms"), default_ms=stored.get("default_ms"))
    overrides = {kv.split("=")[0]: int(kv.split("=")[1]) for kv in args.tune}
    if overrides:
        tuning["overrides"] = overrides
    s.set_tuning(**{**tuning["settings"], **overrides})
    progress(f"{args.workload}: {info.num_states} states, device buffers {info.device_bytes} B")
    for _ in range(args.warmup):
        s.initialize_captures()
        s.main_loop()
    progress("warm-up done")
    # per-kernel breakdown from one untimed solve with every launch timed; the
    # timed steps then carry events around the dominant kernel's launches only
    s.set_kernel_timing(True)
    s.initialize_captures()
    s.main_loop()
    breakdown = {name: ms for name, (ms, _) in s.kernel_times().items()}
    dominant = max(breakdown, key=breakdown.get)
    s.set_kernel_timing(False)
    kt_total = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        if i == args.steps - 1:
            s.set_kernel_timing(True, only=dominant)
        s.initialize_captures()
        passes = s.main_loop()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    progress(f"timed {args.steps} steps: {elapsed * 1e3 / args.steps:.3f} ms per step")
    for name, (ms, launches) in s.kernel_times().items():  # the last timed step
        kt_total[name] = [ms, launches]
    s.set_kernel_timing(False)
    # the work count comes from the table the last timed step solved (untimed)
    e_cop, e_rob, won = s.retro_edges()
    e_retro = e_cop + e_rob
    verdict = s.find_final_result()
    stats = s.stats()
    ms_per_step = elapsed * 1e3 / args.steps
    value = e_retro * args.steps / elapsed / 1e9

    e_cop_d, e_rob_d, pulled = s.work_profile()
    lv = s.level_counts()
    push_units = sum(int(e_cop_d[d]) for d in range(1, 127) if not pulled[d])
    pull_units = sum(int(e_cop_d[d]) for d in range(1, 127) if pulled[d])
    alg_bytes = {
        "cop_push": 0.5 * push_units,
        "sep_stage": 0.5 * pull_units,
        "robber_step": 2.0 * sum(int(x) for x in e_rob_d[1:]),
        "cop_finalize": 1.25 * sum(int(lv[d, 0]) for d in range(2, 127)),
    }
    rl = None
    if dominant in alg_bytes:
        ms, launches = kt_total[dominant]
        bytes_per_launch = alg_bytes[dominant] / launches
        avg_s = ms / 1e3 / launches
        achieved = bytes_per_launch / avg_s
        rl = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
              "frac": achieved / HBM_PEAK, "traffic": measured_traffic(args.workload, dominant),
              "traffic_rate": traffic_rate(args.workload, dominant),
              "kernel": KERNEL_NAMES.get(dominant, dominant),
              "alg_bytes_per_launch": bytes_per_launch, "avg_launch_us": avg_s * 1e6,
              "launches_per_step": launches}
    b_alg = 2 * e_rob + 0.5 * e_cop + won * 1.25 + info.num_states * (0.25 + 1) + info.num_states / 2
    performed = e_retro - int(e_cop_d[0]) - int(e_rob_d[0])
    out = {
        "metric": "retro-edges/s (GTEPS)",
        "value": value,
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/u64 bitsets (integer)",
        "data": "synthetic (seeded graph generator, committed matrix)",
        "config": {"workload": args.workload, "graph": stem, "n": info.n, "k": k, "p": 1,
                   "configs": info.config_count, "states": info.num_states,
                   "e_cop": e_cop, "e_rob": e_rob, "won_states": won, "levels": stats.n_levels,
                   "passes": len(passes), "verdict": "WIN" if verdict else "LOSS",
                   "positions": list(verdict["positions"]) if verdict else None,
                   "plys": verdict["plys"] if verdict else None, "tuning": tuning},
        "performed_gteps": {"value": performed * args.steps / elapsed / 1e9, "relaxations": performed,
                            "note": "relaxations from states of depth >= 1; those of the depth-0 captures are "
                                    "replaced by the closed-form initialisation"},
        "roofline": rl,
        "solve_roofline": {"alg_bytes": b_alg, "achieved_GBps": b_alg / (elapsed / args.steps) / 1e9,
                           "frac": b_alg / (elapsed / args.steps) / HBM_PEAK},
        "kernels_ms_per_step": breakdown,
        "host": host_info(),
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        # bounded sample: the reference on the same graph with fewer cops, the
        # most that keep it to at most 50,000 configurations (~2 s on one core
        # for n = 200)
        from math import comb
        sk = max(1, k - 1)
        while sk > 1 and comb(info.n + sk - 1, sk) > 50_000:
            sk -= 1
        with cra.Solver(0) as t:
            t.load_graph_file(path, sk, 1)
            t.build_aux_graph()
            t.initialize_captures()
            t.main_loop()
            se = sum(t.retro_edges()[:2])
        out["cpu_baseline"] = cpu_reference_sample(stem, sk, se)
        progress("reference sample done")
    if args.cpu_port < 0:
        args.cpu_port = 0 if args.no_cpu_baseline else min(16, len(os.sched_getaffinity(0)))
    mismatch = False
    if args.cpu_port:
        capped = info.num_states > 2_000_000_000
        port, states = cpu_port(stem, k, args.cpu_port, args.cpu_port_cap if capped else None)
        out["cpu_port"] = port
        if not capped:
            import numpy as np
            same = bool(np.array_equal(states, s.states()))
            out["check_vs_port"] = same
            mismatch = not same
        del states
        progress("cpu port done")
        out["reference_mt"] = reference_mt(stem, 3, args.cpu_port)
        progress("reference multithreaded solver done")
    s.close()
    if args.workload == "planar200_k4" and not args.no_batch:
        out["batch"] = run_batch(5, 2, 128, 0.7, args.autotune, args.tuning_db)[0]
        progress("batch block done")
    print(json.dumps(out))
    if mismatch:
        print("bench: the device table differs from the CPU port's", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
