"""Seeded input graphs for the benchmark and parity workloads.

All randomness comes from SplitMix64 (Steele, Lea, Flood, "Fast splittable
pseudorandom number generators", OOPSLA 2014), so every graph is a pure
function of its seed. Graphs are written in the reference's adjacency-matrix
text format (one row of '0'/'1' per line, a final '-' line), the input the
reference's Graph.cpp:15-79 parses.

Workloads:
  * er_graph        Erdos-Renyi G(200, 0.03), first connected seed >= s0.
  * planar_graph    Delaunay triangulation of 200 uniform points in the unit
                    square, edges deleted in seeded order (never a bridge)
                    until 400 remain (mean degree 4).
  * grid_chords     8x8 grid (row-major ids) plus 16 chords drawn without
                    replacement among non-adjacent vertex pairs.
  * batch_graphs    4096 connected G(n, p) graphs, n ~ U{8..48},
                    p ~ U[0.08, 0.3], k ~ U{1, 2, 3}; graph i uses seed i.
"""
import os

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        """Double in [0, 1) from the top 53 bits."""
        return (self.next() >> 11) * (1.0 / (1 << 53))

    def below(self, m):
        """Integer in [0, m) (multiply-shift on the top 32 bits)."""
        return ((self.next() >> 32) * m) >> 32


def is_connected(adj):
    n = adj.shape[0]
    if n == 0:
        return False
    seen = np.zeros(n, bool)
    seen[0] = True
    stack = [0]
    while stack:
        u = stack.pop()
        for v in np.flatnonzero(adj[u] | adj[:, u]):
            if not seen[v]:
                seen[v] = True
                stack.append(v)
    return bool(seen.all())


def gnp(n, p, rng):
    adj = np.zeros((n, n), np.uint8)
    for i in range(n):
        for j in range(i + 1, n):
            if rng.uniform() < p:
                adj[i, j] = adj[j, i] = 1
    return adj


def er_graph(n=200, p=0.03, seed=1):
    """First connected G(n, p) over seeds seed, seed+1, ... -> (adj, seed)."""
    s = seed
    while True:
        adj = gnp(n, p, SplitMix64(s))
        if is_connected(adj):
            return adj, s
        s += 1


def planar_graph(n=200, target_edges=400, seed=1):
    from scipy.spatial import Delaunay

    rng = SplitMix64(seed)
    pts = np.array([[rng.uniform(), rng.uniform()] for _ in range(n)])
    tri = Delaunay(pts)
    adj = np.zeros((n, n), np.uint8)
    for a, b, c in tri.simplices:
        for u, v in ((a, b), (b, c), (a, c)):
            adj[u, v] = adj[v, u] = 1
    edges = [(i, j) for i in range(n) for j in range(i + 1, n) if adj[i, j]]
    for i in range(len(edges) - 1, 0, -1):  # seeded Fisher-Yates deletion order
        j = rng.below(i + 1)
        edges[i], edges[j] = edges[j], edges[i]
    count = len(edges)
    for u, v in edges:
        if count <= target_edges:
            break
        adj[u, v] = adj[v, u] = 0
        if is_connected(adj):
            count -= 1
        else:  # a bridge: keep it
            adj[u, v] = adj[v, u] = 1
    return adj


def grid_chords(rows=8, cols=8, chords=16, seed=1):
    n = rows * cols
    adj = np.zeros((n, n), np.uint8)
    for r in range(rows):
        for c in range(cols):
            v = r * cols + c
            if c + 1 < cols:
                adj[v, v + 1] = adj[v + 1, v] = 1
            if r + 1 < rows:
                adj[v, v + cols] = adj[v + cols, v] = 1
    pairs = [(i, j) for i in range(n) for j in range(i + 1, n) if not adj[i, j]]
    rng = SplitMix64(seed)
    for t in range(chords):  # partial Fisher-Yates: first `chords` picks
        j = t + rng.below(len(pairs) - t)
        pairs[t], pairs[j] = pairs[j], pairs[t]
        u, v = pairs[t]
        adj[u, v] = adj[v, u] = 1
    return adj


def batch_graphs(count=4096, seed_base=0):
    """[(adj, k)] with graph i drawn from seed seed_base + i."""
    out = []
    for i in range(count):
        rng = SplitMix64(seed_base + i)
        while True:
            n = 8 + rng.below(41)
            p = 0.08 + 0.22 * rng.uniform()
            k = 1 + rng.below(3)
            adj = gnp(n, p, rng)
            if is_connected(adj):
                break
        out.append((adj, k))
    return out


def to_text(adj):
    rows = ["".join("1" if x else "0" for x in row) for row in adj]
    return "\n".join(rows) + "\n-"


def write_matrix(path, adj):
    with open(path, "w", newline="\n") as f:
        f.write(to_text(adj))


ER_SEED = 1
PLANAR_SEED = 1
GRID_SEED = 1


def named():
    """The committed workload graphs: {file stem: builder}."""
    return {
        f"er200_p0.03_s{ER_SEED}": lambda: er_graph(200, 0.03, ER_SEED)[0],
        f"planar200_e400_s{PLANAR_SEED}": lambda: planar_graph(200, 400, PLANAR_SEED),
        f"grid8x8_chords16_s{GRID_SEED}": lambda: grid_chords(8, 8, 16, GRID_SEED),
    }


def write_all(directory):
    for stem, make in named().items():
        write_matrix(os.path.join(directory, stem + ".txt"), make())
