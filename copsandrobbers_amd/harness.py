"""Sweep runner that writes the reference's regression CSV.

A sweep solves every graph of a matrix folder for k = 1..max_k cops and
visibility p = 1..max_p with a solver executable (the reference's binaries or
this package's CLI, bin/k_cops_visibility*), one process per case, cheapest
case first, and writes one CSV row per case. The columns are the reference
harness's (util/test_k_cops_visibility.py:231-233), and the values come from
the solver's stdout lines that harness reads (:78-100): verdict, start
positions, capture plies, main-loop seconds, memory footprint, base captures,
cache hit and the last pass number. Cases are costed, and cut at a state
budget, by the reference's estimate n^k * n * 2p bytes (:72-74).

Usage:
  python -m copsandrobbers_amd.harness --matrices tests/golden/matrices \\
      --exe copsandrobbers_amd/bin/k_cops_visibility --csv out.csv
"""
import argparse
import collections
import csv
import os
import re
import subprocess
import sys
import time

# Column order of the reference's CSV.
CSV_COLUMNS = ["graph", "n", "k", "p", "est_state_mb", "status", "result", "capture_plys", "passes",
               "base_caps", "cache_hit", "mem_mb", "mainloop_s", "solver_s", "wall_s", "exit_code",
               "positions"]

# Graphs a sweep leaves out unless asked (minutes per case in the reference).
SLOW_GRAPHS = frozenset({"scotlandyard-all", "scotlandyard-yellow", "scotlandyard-green", "scotlandyard-red",
                         "cycle100", "line100", "tree127"})


def _first(pattern, convert, missing=""):
    """A stdout reader: the first match of `pattern`, converted, or `missing`."""
    rx = re.compile(pattern)

    def read(text):
        m = rx.search(text)
        return convert(m.group(1)) if m else missing
    return read


def _last_pass(text):
    found = re.findall(r"Pass\s+(\d+):", text)
    return int(found[-1]) if found else ""


# CSV column -> how its value is read from the solver's stdout.
STDOUT_READERS = {
    "result": _first(r"RESULT:\s+(WIN|LOSS)", str, missing="?"),
    "positions": _first(r"Optimal Cop Start Positions:\s*\(([^)]*)\)", str.strip),
    "capture_plys": _first(r"Capture Time:\s*(\d+)\s*plys", int),
    "mainloop_s": _first(r"Main Loop\s*-*=>\s*([\d.]+)\s*s", float),
    "mem_mb": _first(r"Total Footprint\s*-*=>\s*(\d+)\s*B", lambda b: round(int(b) / 1e6, 1)),
    "base_caps": _first(r"Initialized\s+(\d+)\s+base captures", int),
    "cache_hit": lambda text: "Loaded AuxGraph from cache" in text,
    "passes": _last_pass,
}


def vertex_count(path):
    """Vertices of an adjacency-matrix file: the length of its first row."""
    with open(path) as f:
        for line in f:
            if line.strip():
                return len(line.strip())
    return 0


def state_bytes_estimate(n, k, p):
    """The reference's cost estimate of a case: n^k configurations x n robber
    positions x 2p turns, one byte each."""
    return (n ** k) * n * (2 * p)


def parse_stdout(text):
    """The CSV values a solver's stdout carries (missing ones empty)."""
    values = {column: read(text) for column, read in STDOUT_READERS.items()}
    values["solver_s"] = ""
    return values


def run_solver(exe, matrix_path, k, p, timeout=90.0, cwd=None):
    """Runs `exe matrix k p` once; returns its CSV values with status, exit
    code and wall time (status TIMEOUT when it does not finish in time)."""
    start = time.perf_counter()
    try:
        proc = subprocess.run([str(exe), str(matrix_path), str(k), str(p)], cwd=cwd, capture_output=True,
                              text=True, timeout=timeout)
    except subprocess.TimeoutExpired:
        return {"status": "TIMEOUT", "wall_s": round(time.perf_counter() - start, 2)}
    values = parse_stdout(proc.stdout)
    values["wall_s"] = round(time.perf_counter() - start, 2)
    values["exit_code"] = proc.returncode
    verdict = values["result"] in ("WIN", "LOSS")
    if proc.returncode == 0:
        values["status"] = "ok" if verdict else "no-result"
    else:
        values["status"] = ("ok-but-" if verdict else "") + f"exit{proc.returncode}"
    values["stderr"] = proc.stderr
    return values


Case = collections.namedtuple("Case", "cost n path graph k p")


def plan_sweep(matrices, max_k, max_p, state_budget, name_filters=(), include_slow=False):
    """The cases of a sweep over the .txt matrices of a folder, cheapest first."""
    cases = []
    for fn in sorted(os.listdir(matrices)):
        graph, ext = os.path.splitext(fn)
        if ext != ".txt" or (name_filters and not any(f in graph for f in name_filters)):
            continue
        if graph in SLOW_GRAPHS and not include_slow:
            continue
        path = os.path.join(matrices, fn)
        n = vertex_count(path)
        if n < 2:
            continue
        for k in range(1, min(max_k, n - 1) + 1):
            for p in range(1, max_p + 1):
                cost = state_bytes_estimate(n, k, p)
                if cost <= state_budget:
                    cases.append(Case(cost, n, path, graph, k, p))
    cases.sort(key=lambda c: c.cost)
    return cases


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--exe", required=True)
    ap.add_argument("--matrices", required=True)
    ap.add_argument("--csv", required=True)
    ap.add_argument("--timeout", type=float, default=90.0)
    ap.add_argument("--max-k", type=int, default=3)
    ap.add_argument("--max-p", type=int, default=3)
    ap.add_argument("--state-budget", type=float, default=1.5e9)
    ap.add_argument("--graphs", default="", help="comma-separated substrings of graph names to keep")
    ap.add_argument("--include-big", action="store_true", help="also sweep the slow graphs")
    ap.add_argument("--cwd", default=None)
    args = ap.parse_args(argv)
    filters = tuple(s.strip() for s in args.graphs.split(",") if s.strip())
    exe = os.path.abspath(args.exe)
    rows = []
    for case in plan_sweep(args.matrices, args.max_k, args.max_p, args.state_budget, filters, args.include_big):
        row = {"graph": case.graph, "n": case.n, "k": case.k, "p": case.p,
               "est_state_mb": round(case.cost / 1e6, 1)}
        row.update(run_solver(exe, os.path.abspath(case.path), case.k, case.p, args.timeout, args.cwd))
        rows.append(row)
        print(f"{case.graph:<20} {case.n:>4} {case.k:>2} {case.p:>2} {row.get('result', ''):<6} "
              f"{row.get('capture_plys', '')!s:>4} {row.get('passes', '')!s:>6} {row.get('status')}", flush=True)
    with open(args.csv, "w", newline="") as f:
        writer = csv.DictWriter(f, fieldnames=CSV_COLUMNS, extrasaction="ignore")
        writer.writeheader()
        writer.writerows(rows)
    return 0


if __name__ == "__main__":
    sys.exit(main())
