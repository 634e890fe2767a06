"""Small seeded test graphs shared by the CPU and GPU tests (sizes the oracle finishes in
well under a second).  Built with the oracle's generators or by hand."""
from __future__ import annotations

import functools

import numpy as np

from oracle import pyoracle as po


def _hub_graph(seed=3, n=2600, hub_deg=1500, extra=6000):
    """A vertex of degree > 1023 (HBM-table path) inside a sparse random graph."""
    rng = np.random.default_rng(seed)
    others = rng.choice(np.arange(1, n), size=hub_deg, replace=False)
    pairs = [(0, int(v)) for v in others]
    a = rng.integers(1, n, size=extra)
    b = rng.integers(1, n, size=extra)
    keep = a != b
    pairs += list(zip(a[keep].tolist(), b[keep].tolist()))
    pairs = np.array(pairs, dtype=np.int64)
    lo, hi = np.minimum(pairs[:, 0], pairs[:, 1]), np.maximum(pairs[:, 0], pairs[:, 1])
    uniq = np.unique(np.stack([lo, hi], 1), axis=0)
    w = rng.uniform(0.05, 1.0, size=len(uniq))
    return po.from_edge_list(n, uniq, w)


def _multi_graph(seed=9, n=700):
    """Rows with repeated tails (parallel edges), self-loops and isolated vertices."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        if i % 37 == 0:
            continue  # isolated
        k = int(rng.integers(1, 9))
        tails = rng.integers(max(0, i - 30), min(n, i + 30), size=k)
        tails = np.concatenate([tails, tails[: k // 3]])  # repeats
        if i % 5 == 0:
            tails = np.concatenate([tails, [i, i]])  # two self-loop entries
        for t in tails:
            if int(t) % 37 == 0:
                continue
            rows.append((i, int(t), float(rng.uniform(0.1, 2.0))))
    rows.sort(key=lambda r: r[0])
    offs = np.zeros(n + 1, np.int64)
    for r in rows:
        offs[r[0] + 1] += 1
    offs = np.cumsum(offs)
    return po.make_graph(offs, [r[1] for r in rows], [r[2] for r in rows])


def _unit_grid(side=40):
    """Unit-weight 2D grid: many exact gain ties (min-id tie-break)."""
    n = side * side
    pairs = []
    for y in range(side):
        for x in range(side):
            v = y * side + x
            if x + 1 < side:
                pairs.append((v, v + 1))
            if y + 1 < side:
                pairs.append((v, v + side))
    return po.from_edge_list(n, pairs)


def _dense_geometric(seed=21, n=2400, mean_deg=70.0, extra=3000):
    """Geometric graph with degrees spread over roughly 35-110, so most rows fall in the
    32-63 and 64-127 degree classes (the two largest group sizes of the local move), plus
    random long edges."""
    rng = np.random.default_rng(seed)
    x, y = rng.random(n), rng.random(n)
    r = np.sqrt(mean_deg / (np.pi * n))
    d = np.sqrt((x[:, None] - x[None, :]) ** 2 + (y[:, None] - y[None, :]) ** 2)
    i, j = np.nonzero(np.triu(d <= r, 1))
    w = d[i, j]
    a = rng.integers(0, n, size=extra)
    b = rng.integers(0, n, size=extra)
    keep = a != b
    pairs = np.concatenate([np.stack([i, j], 1), np.stack([np.minimum(a, b), np.maximum(a, b)], 1)[keep]])
    wts = np.concatenate([w, rng.uniform(0.0, 1.0, size=int(keep.sum()))])
    pairs, first = np.unique(pairs, axis=0, return_index=True)
    return po.from_edge_list(n, pairs, wts[first])


# Zachary's karate club (34 members, 78 friendships; W. W. Zachary, J. Anthropol. Res. 33,
# 1977): for each member (1-based), the higher-numbered members it is tied to.
_KARATE = {1: [2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13, 14, 18, 20, 22, 32], 2: [3, 4, 8, 14, 18, 20, 22, 31],
           3: [4, 8, 9, 10, 14, 28, 29, 33], 4: [8, 13, 14], 5: [7, 11], 6: [7, 11, 17], 7: [17],
           9: [31, 33, 34], 10: [34], 14: [34], 15: [33, 34], 16: [33, 34], 19: [33, 34], 20: [34],
           21: [33, 34], 23: [33, 34], 24: [26, 28, 30, 33, 34], 25: [26, 28, 32], 26: [32], 27: [30, 34],
           28: [34], 29: [32, 34], 30: [33, 34], 31: [33, 34], 32: [33, 34], 33: [34]}


# The club's split after the conflict (Zachary's faction membership, 1-based): members who
# went with the instructor; the rest went with the officer.
KARATE_INSTRUCTOR = [1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13, 14, 17, 18, 20, 22]


def karate_factions():
    """Ground truth for the karate club: 0 = instructor's faction, 1 = officer's."""
    return np.array([0 if v + 1 in KARATE_INSTRUCTOR else 1 for v in range(34)], dtype=np.int64)


def _karate():
    """Zachary's karate club with unit weights, 0-based ids."""
    pairs = [(a - 1, b - 1) for a, bs in _KARATE.items() for b in bs]
    return po.from_edge_list(34, pairs)


@functools.lru_cache(maxsize=None)
def graph(name: str) -> po.Graph:
    if name == "rgg":
        return po.gen_rgg(11, 3000, 1.0)
    if name == "rgg_plain":
        return po.gen_rgg(12, 2500, 0.0)
    if name == "ba":
        return po.gen_ba(5, 4000, 4)
    if name == "vite_rgg":
        return po.gen_vite_rgg(2048)
    if name == "hub":
        return _hub_graph()
    if name == "multi":
        return _multi_graph()
    if name == "grid":
        return _unit_grid()
    if name == "dense":
        return _dense_geometric()
    if name == "karate":
        return _karate()
    raise KeyError(name)


ALL = ["rgg", "rgg_plain", "ba", "vite_rgg", "hub", "multi", "grid", "dense", "karate"]


def random_state(g: po.Graph, seed: int = 0):
    """A consistent mid-run state: communities are vertex ids, each vertex keeps its own id or
    takes a neighbour's; localCinfo is the member count and the in-order degree sum."""
    rng = np.random.default_rng(seed)
    nv = g.nv
    vdeg, constant = po.init_degrees(g)
    curr = np.arange(nv, dtype=np.int64)
    for i in range(nv):
        e0, e1 = g.offs[i], g.offs[i + 1]
        if e1 > e0 and rng.random() < 0.6:
            curr[i] = g.edges["tail"][rng.integers(e0, e1)]
    csize = np.zeros(nv, np.int64)
    cdeg = np.zeros(nv)
    for i in range(nv):
        csize[curr[i]] += 1
        cdeg[curr[i]] += vdeg[i]
    return curr, csize, cdeg, vdeg, constant
