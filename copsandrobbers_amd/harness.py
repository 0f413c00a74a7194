"""Regression harness with the reference's CSV contract.

Runs a solver executable over graphs x k x p, parses its stdout and writes
one CSV row per run with the columns of util/test_k_cops_visibility.py:231-233
(graph, n, k, p, est_state_mb, status, result, capture_plys, passes,
base_caps, cache_hit, mem_mb, mainloop_s, solver_s, wall_s, exit_code,
positions). The same stdout lines are parsed (:78-100), so the reference
binaries and this package's CLI (bin/k_cops_visibility*) can be swept alike.

Usage:
  python -m copsandrobbers_amd.harness --matrices tests/golden/matrices \\
      --exe copsandrobbers_amd/bin/k_cops_visibility --csv out.csv
"""
import argparse
import csv
import os
import re
import subprocess
import sys
import time

RE_RESULT = re.compile(r"RESULT:\s+(WIN|LOSS)")
RE_POSITIONS = re.compile(r"Optimal Cop Start Positions:\s*\(([^)]*)\)")
RE_CAPTURE = re.compile(r"Capture Time:\s*(\d+)\s*plys")
RE_MAINLOOP = re.compile(r"Main Loop\s*-*=>\s*([\d.]+)\s*s")
RE_FOOTPRINT = re.compile(r"Total Footprint\s*-*=>\s*(\d+)\s*B")
RE_PASS = re.compile(r"Pass\s+(\d+):")
RE_CACHE_HIT = re.compile(r"Loaded AuxGraph from cache")
RE_BASECAPS = re.compile(r"Initialized\s+(\d+)\s+base captures")

FIELDS = ["graph", "n", "k", "p", "est_state_mb", "status", "result", "capture_plys", "passes",
          "base_caps", "cache_hit", "mem_mb", "mainloop_s", "solver_s", "wall_s", "exit_code",
          "positions"]
BIG_GRAPHS = {"scotlandyard-all", "scotlandyard-yellow", "scotlandyard-green", "scotlandyard-red",
              "cycle100", "line100", "tree127"}


def node_count(path):
    with open(path) as f:
        for line in f:
            row = line.strip()
            if row:
                return len(row)
    return 0


def est_state_bytes(n, k, p):
    """The reference harness's cost model: n^k configs * n * 2p bytes."""
    return (n ** k) * n * (2 * p)


def parse_output(text):
    out = {}
    m = RE_RESULT.search(text)
    out["result"] = m.group(1) if m else "?"
    m = RE_POSITIONS.search(text)
    out["positions"] = m.group(1).strip() if m else ""
    m = RE_CAPTURE.search(text)
    out["capture_plys"] = int(m.group(1)) if m else ""
    m = RE_MAINLOOP.search(text)
    out["mainloop_s"] = float(m.group(1)) if m else ""
    m = RE_FOOTPRINT.search(text)
    out["mem_mb"] = round(int(m.group(1)) / 1e6, 1) if m else ""
    m = RE_BASECAPS.search(text)
    out["base_caps"] = int(m.group(1)) if m else ""
    out["cache_hit"] = bool(RE_CACHE_HIT.search(text))
    passes = RE_PASS.findall(text)
    out["passes"] = int(passes[-1]) if passes else ""
    out["solver_s"] = ""
    return out


def run_one(exe, matrix_path, k, p, timeout=90.0, cwd=None):
    cmd = [str(exe), str(matrix_path), str(k), str(p)]
    t0 = time.perf_counter()
    try:
        proc = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=timeout)
    except subprocess.TimeoutExpired:
        return {"status": "TIMEOUT", "wall_s": round(time.perf_counter() - t0, 2)}
    parsed = parse_output(proc.stdout)
    parsed["wall_s"] = round(time.perf_counter() - t0, 2)
    parsed["exit_code"] = proc.returncode
    got = parsed["result"] in ("WIN", "LOSS")
    if proc.returncode == 0:
        parsed["status"] = "ok" if got else "no-result"
    else:
        parsed["status"] = ("ok-but-" if got else "") + f"exit{proc.returncode}"
    parsed["stderr"] = proc.stderr
    return parsed


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--exe", required=True)
    ap.add_argument("--matrices", required=True)
    ap.add_argument("--csv", required=True)
    ap.add_argument("--timeout", type=float, default=90.0)
    ap.add_argument("--max-k", type=int, default=3)
    ap.add_argument("--max-p", type=int, default=3)
    ap.add_argument("--state-budget", type=float, default=1.5e9)
    ap.add_argument("--graphs", default="")
    ap.add_argument("--include-big", action="store_true")
    ap.add_argument("--cwd", default=None)
    args = ap.parse_args(argv)
    filters = [s.strip() for s in args.graphs.split(",") if s.strip()]
    jobs = []
    for fn in sorted(os.listdir(args.matrices)):
        if not fn.endswith(".txt"):
            continue
        stem = fn[:-4]
        path = os.path.join(args.matrices, fn)
        if filters and not any(f in stem for f in filters):
            continue
        if stem in BIG_GRAPHS and not args.include_big:
            continue
        n = node_count(path)
        if n < 2:
            continue
        for k in range(1, args.max_k + 1):
            if k >= n:
                continue
            for p in range(1, args.max_p + 1):
                est = est_state_bytes(n, k, p)
                if est <= args.state_budget:
                    jobs.append((est, n, path, stem, k, p))
    jobs.sort(key=lambda j: j[0])
    rows = []
    for est, n, path, stem, k, p in jobs:
        res = run_one(os.path.abspath(args.exe), os.path.abspath(path), k, p, args.timeout, args.cwd)
        row = {"graph": stem, "n": n, "k": k, "p": p, "est_state_mb": round(est / 1e6, 1)}
        row.update(res)
        rows.append(row)
        print(f"{stem:<20} {n:>4} {k:>2} {p:>2} {row.get('result', ''):<6} {row.get('capture_plys', '')!s:>4} "
              f"{row.get('passes', '')!s:>6} {row.get('status')}", flush=True)
    with open(args.csv, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS, extrasaction="ignore")
        w.writeheader()
        for r in rows:
            w.writerow(r)
    return 0


if __name__ == "__main__":
    sys.exit(main())
