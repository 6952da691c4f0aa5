"""Seeded synthetic instances written in the reference .dat formats.

PRNG: SplitMix64 with unbiased rejection sampling onto [lo, hi] (the mapping
of std::uniform_int_distribution is implementation-defined, so it is not
used).  Draw order is fixed per problem so that a (seed, sizes) pair names
one instance on every machine.

Formats (readers: src/pmdd/instances/knapsack_instance.cpp:17-47,
setpacking_instance.hpp:49-89, tsp_instance.cpp:11-41):
  knapsack     "n 1 d" / d rows of n profits / n weights then the capacity
  set packing  "n m" / "d" / d rows of n coefficients / per constraint
               "k v_1 .. v_k" (1-based)
  TSP          "d n" / d matrices n x n
"""
from __future__ import annotations

import os

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi], unbiased (rejection of the top sliver)."""
        span = hi - lo + 1
        limit = ((1 << 64) // span) * span
        while True:
            x = self.next()
            if x < limit:
                return lo + x % span


def knapsack(seed: int, n: int, d: int, lo: int = 1, hi: int = 1000) -> dict:
    """One-constraint knapsack: weights U[lo,hi], capacity floor(0.5 * sum),
    profits U[lo,hi] independently per objective.  Draws: weights, then the
    profits objective by objective."""
    rng = SplitMix64(seed)
    w = [rng.uniform(lo, hi) for _ in range(n)]
    profits = [[rng.uniform(lo, hi) for _ in range(n)] for _ in range(d)]
    cap = sum(w) // 2
    return {"kind": 1, "n": n, "d": d, "weights": w, "capacity": cap, "profits": profits}


def setpacking(seed: int, n: int, d: int, m: int | None = None, kmin: int = 5, kmax: int = 15,
               lo: int = 1, hi: int = 100) -> dict:
    """Set packing: m = n // 5 constraints unless given; each a uniform random
    subset of U[kmin,kmax] variables; any variable left uncovered joins a
    uniformly chosen constraint; coefficients U[lo,hi].  Draws: subsets
    (size, then a partial Fisher-Yates), repairs, then coefficients."""
    if m is None:
        m = max(1, n // 5)
    rng = SplitMix64(seed)
    cons = []
    for _ in range(m):
        k = min(n, rng.uniform(kmin, kmax))
        pool = list(range(n))
        for i in range(k):
            j = rng.uniform(i, n - 1)
            pool[i], pool[j] = pool[j], pool[i]
        cons.append(sorted(pool[:k]))
    covered = set(v for c in cons for v in c)
    for v in range(n):
        if v not in covered:
            c = rng.uniform(0, m - 1)
            cons[c] = sorted(cons[c] + [v])
    objs = [[rng.uniform(lo, hi) for _ in range(n)] for _ in range(d)]
    return {"kind": 2, "n": n, "d": d, "m": m, "constraints": cons, "objs": objs}


def tsp(seed: int, n: int, d: int, lo: int = 1, hi: int = 1000) -> dict:
    """Asymmetric TSP: d independent n x n matrices U[lo,hi], zero diagonal.
    Draws: matrix by matrix, row-major, skipping the diagonal."""
    rng = SplitMix64(seed)
    mats = []
    for _ in range(d):
        mat = [[0 if i == j else rng.uniform(lo, hi) for j in range(n)] for i in range(n)]
        mats.append(mat)
    return {"kind": 3, "n": n, "d": d, "dist": mats}


def to_text(inst: dict) -> str:
    k = inst["kind"]
    lines = []
    if k == 1:
        lines.append(f"{inst['n']} 1 {inst['d']}")
        for row in inst["profits"]:
            lines.append(" ".join(map(str, row)))
        lines.append(" ".join(map(str, inst["weights"])) + f" {inst['capacity']}")
    elif k == 2:
        lines.append(f"{inst['n']} {inst['m']}")
        lines.append(f"{inst['d']}")
        for row in inst["objs"]:
            lines.append(" ".join(map(str, row)))
        for c in inst["constraints"]:
            lines.append(" ".join(map(str, [len(c)] + [v + 1 for v in c])))
    elif k == 3:
        lines.append(f"{inst['d']} {inst['n']}")
        for mat in inst["dist"]:
            for row in mat:
                lines.append(" ".join(map(str, row)))
    else:
        raise ValueError(f"unknown instance kind {k}")
    return "\n".join(lines) + "\n"


def write(inst: dict, path: str) -> str:
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        f.write(to_text(inst))
    return path


def generate(kind: int, seed: int, n: int, d: int, **kw) -> dict:
    if kind == 1:
        return knapsack(seed, n, d, **kw)
    if kind == 2:
        return setpacking(seed, n, d, **kw)
    if kind == 3:
        return tsp(seed, n, d, **kw)
    raise ValueError(kind)
