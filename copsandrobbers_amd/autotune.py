"""Autotuner for the solver's speed settings.

The reference has no counterpart: its only sizing logic is the regression
harness's cost model that orders jobs cheapest first
(util/test_k_cops_visibility.py:72-74). Here every speed setting of the
device solver (cr_set_tuning, cr_batch_set_tuning) leaves the results
unchanged, so the tuner may pick any of them by measured time alone.

Search: coordinate descent. Starting from the library's defaults (no
override), each knob in turn is set to each of its candidate values with the
others held at the best found so far; a candidate replaces the incumbent only
if it is faster by more than `min_gain` (relative), which keeps run-to-run
noise from moving the choice. The knobs are visited in the order listed, most
influential first, and the pass is repeated while it still improves (at most
`rounds` passes).

Database: the chosen overrides are stored per problem shape in a JSON file
(default copsandrobbers_amd/tuning/mi355x.json, which ships with the
package; CR_TUNING_DB overrides the path). A plain solve is keyed on
(n, k, p, configurations, row words, edges, stay rules), a batch on (graph
count, largest n, largest k, total states). Lookups apply a stored entry;
a shape without one runs on the defaults.
"""
import json
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_DB = os.path.join(HERE, "tuning", "mi355x.json")

# (knob, candidates) in visiting order; names as in copsandrobbers_amd.TUNING
PLAIN_SPACE = [
    ("pull_scale", [100, 130, 150, 190, 250, 400]),
    ("pull_gate", [1, 4, 16]),
    ("pull_pair", [0, 1, 2]),
    ("pull_tile_rows", [2, 4, 8]),
    ("pull_tile1", [4, 9, 12]),
    ("pull_tile_chunk", [4, 16, 32]),
    ("pull_tile_blocks", [4096, 8192, 16384]),
    ("pair_blocks", [512, 1024, 2048]),
    ("robber_open", [1, 2, 4]),
    ("levels_per_sync", [4, 8, 16]),
]
# names as in copsandrobbers_amd.batch.TUNING
BATCH_SPACE = [
    ("threads", [0, 256, 512, 1024]),
    ("blocks_per_cu", [0, 1, 2, 4, 8]),
    ("push_pct", [0, 30, 100, 300]),
]


def db_path(path=None):
    return path or os.environ.get("CR_TUNING_DB") or DEFAULT_DB


def load_db(path=None):
    p = db_path(path)
    if not os.path.exists(p):
        return {}
    with open(p) as f:
        return json.load(f)


def save_entry(key, entry, path=None):
    """Merges one entry into the database file (written atomically)."""
    p = db_path(path)
    db = load_db(p)
    db[key] = entry
    os.makedirs(os.path.dirname(p) or ".", exist_ok=True)
    tmp = p + ".tmp"
    with open(tmp, "w") as f:
        json.dump(db, f, indent=1, sort_keys=True)
        f.write("\n")
    os.replace(tmp, p)


def plain_key(n, k, p, configs, words, edges, cop_stay=0, robber_stay=0):
    return f"plain/n{n}/k{k}/p{p}/C{configs}/W{words}/e{edges}/stay{int(cop_stay)}{int(robber_stay)}"


def batch_key(graphs):
    """graphs: [(adj, k)]."""
    from math import comb
    states = sum(2 * comb(a.shape[0] + k - 1, k) * a.shape[0] for a, k in graphs)
    nmax = max((a.shape[0] for a, _ in graphs), default=0)
    kmax = max((k for _, k in graphs), default=0)
    return f"batch/g{len(graphs)}/nmax{nmax}/kmax{kmax}/S{states}"


def coordinate_descent(measure, space, rounds=2, min_gain=0.02, log=None):
    """measure(settings) -> seconds (lower is better). Returns (best settings,
    best time, baseline time, evaluations)."""
    best = {}
    base = best_t = measure(best)
    evals = [({}, base)]
    if log:
        log(f"tune: defaults {base * 1e3:.3f} ms")
    for _ in range(rounds):
        improved = False
        for name, values in space:
            for v in values:
                if best.get(name) == v:
                    continue
                trial = dict(best)
                trial[name] = v
                t = measure(trial)
                evals.append((trial, t))
                if t < best_t * (1.0 - min_gain):
                    if log:
                        log(f"tune: {name}={v} {t * 1e3:.3f} ms (was {best_t * 1e3:.3f})")
                    best, best_t, improved = trial, t, True
        if not improved:
            break
    return best, best_t, base, evals


def tune_plain(solver, reps=2, space=PLAIN_SPACE, rounds=2, min_gain=0.02, log=None):
    """Tunes the graph already loaded and built in `solver` (a
    copsandrobbers_amd.Solver). Each measurement is the fastest of `reps`
    solves (initializeCaptures + mainLoop). Leaves the best settings applied;
    returns (settings, best seconds, default seconds, evaluations)."""
    def measure(settings):
        _apply_plain(solver, settings)
        t_best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            solver.initialize_captures()
            solver.main_loop()
            t_best = min(t_best, time.perf_counter() - t0)
        return t_best

    solver.initialize_captures()  # warm-up: buffers, caches, tile tables
    solver.main_loop()
    best, t, base, evals = coordinate_descent(measure, space, rounds, min_gain, log)
    _apply_plain(solver, best)
    return best, t, base, evals


def _apply_plain(solver, settings):
    from . import TUNING
    reset = {name: PLAIN_DEFAULTS[name] for name, _ in PLAIN_SPACE if name not in settings}
    solver.set_tuning(**reset)
    solver.set_tuning(**{k: v for k, v in settings.items() if k in TUNING})


# the library's defaults of the knobs in PLAIN_SPACE (include/crsolve.h);
# pull_scale -1 selects its size-dependent default
PLAIN_DEFAULTS = {"pull_scale": -1, "pull_gate": 4, "pull_pair": 1, "pull_tile_rows": 4, "pull_tile1": 9,
                  "pull_tile_chunk": 16, "pull_tile_blocks": 8192, "pair_blocks": 1024, "robber_open": 2,
                  "levels_per_sync": 8}
BATCH_DEFAULTS = {"threads": 0, "blocks_per_cu": 0, "push_pct": 100}


def tune_batch(batch, graphs, reps=3, space=BATCH_SPACE, rounds=2, min_gain=0.02, log=None):
    """Tunes the launch shape of one cr_batch_solve over `graphs` (device
    time of the launch, fastest of `reps`). Leaves the best applied."""
    def measure(settings):
        batch.set_tuning(**{**BATCH_DEFAULTS, **settings})
        t_best = float("inf")
        for _ in range(reps):
            batch.solve(graphs)
            t_best = min(t_best, batch.last_timing()[0] / 1e3)
        return t_best

    batch.solve(graphs)  # warm-up: buffers grow to size
    best, t, base, evals = coordinate_descent(measure, space, rounds, min_gain, log)
    batch.set_tuning(**{**BATCH_DEFAULTS, **best})
    return best, t, base, evals


def plain_entry_key(solver, edges):
    info = solver.info()
    return plain_key(info.n, info.k, info.p, info.config_count, (info.n + 63) // 64, edges)


def lookup_plain(solver, edges, path=None):
    """The stored settings for the graph loaded in `solver` (None if none)."""
    e = load_db(path).get(plain_entry_key(solver, edges))
    return None if e is None else e["settings"]


def lookup_batch(graphs, path=None):
    e = load_db(path).get(batch_key(graphs))
    return None if e is None else e["settings"]


def entry(settings, best_s, default_s, evals, device=None):
    return {"settings": settings, "best_ms": best_s * 1e3, "default_ms": default_s * 1e3,
            "evaluations": len(evals), "device": device, "when": time.strftime("%Y-%m-%d")}


def main(argv=None):
    """python -m copsandrobbers_amd.autotune --graph FILE --k K  (plain solve),
    or --batch4096 (the seeded 4096-graph batch); stores the choice."""
    import argparse

    import numpy as np

    from . import Solver, parse_matrix
    ap = argparse.ArgumentParser()
    ap.add_argument("--graph")
    ap.add_argument("--k", type=int, default=1)
    ap.add_argument("--batch4096", action="store_true")
    ap.add_argument("--db", default=None)
    ap.add_argument("--reps", type=int, default=2)
    args = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)  # noqa: E731
    if args.batch4096:
        from . import workloads
        from .batch import Batch
        graphs = workloads.batch_graphs(4096)
        with Batch(0) as b:
            best, t, base, evals = tune_batch(b, graphs, reps=max(args.reps, 3), log=log)
        save_entry(batch_key(graphs), entry(best, t, base, evals), args.db)
        return 0
    with open(args.graph, "rb") as f:
        n, adj = parse_matrix(f.read())
    edges = int(np.count_nonzero(adj))
    with Solver(0) as s:
        s.load_graph_file(args.graph, args.k, 1)
        s.build_aux_graph()
        best, t, base, evals = tune_plain(s, reps=args.reps, log=log)
        save_entry(plain_entry_key(s, edges), entry(best, t, base, evals), args.db)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
