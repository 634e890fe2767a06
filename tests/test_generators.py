"""The workload generators of the oracle, pinned by what their definitions fix.

The seeded RGG and Barabasi-Albert generators have no counterpart in the reference (Vite's
own RGG uses another radius and an O(n^2) pair test; its BA generator comes from NetworKit,
which is not vendored), so they are pinned here without it:
  * RGG: a brute-force all-pairs distance test over the same points finds exactly the
    edges the cell-binned generator emits (same weights, same degree sequence);
  * RGG extra edges: exactly floor(pct * m) / 100 new undirected edges, none duplicating a
    geometric edge or a self-loop, weights in [0, 1);
  * BA: the structural laws of the Batagelj-Brandes process (symmetric, simple, unit
    weights, at most m edges from a vertex to earlier ones, a heavy-tailed degree
    distribution).  Equality with NetworKit's BA generator is unpinned.
The device generators are compared with these oracle generators bit for bit in
tests/test_gpu_parity.py.
"""
import numpy as np
import pytest

from oracle import pyoracle as po


def _points(seed, n):
    lib = po._lib()
    z = np.array([lib.or_rng_u64(seed, 1, k) for k in range(2 * n)], dtype=np.uint64)
    u = (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return u[0::2], u[1::2]


def _directed_pairs(g):
    heads = np.repeat(np.arange(g.nv), np.diff(g.offs))
    return heads, g.edges["tail"].astype(np.int64), g.edges["weight"]


@pytest.mark.parametrize("seed,n", [(1, 2), (3, 500), (7, 2500), (11, 4000)])
def test_rgg_equals_brute_force(seed, n):
    g = po.gen_rgg(seed, n, 0.0)
    x, y = _points(seed, n)
    r = po.rgg_radius(n)
    dx = x[:, None] - x[None, :]
    dy = y[:, None] - y[None, :]
    d = np.sqrt(dx * dx + dy * dy)
    np.fill_diagonal(d, np.inf)
    within = d <= r
    bf_weights = np.sort(d[within])
    bf_degrees = np.sort(within.sum(axis=1))
    assert g.ne == int(within.sum())
    assert np.array_equal(np.sort(g.edges["weight"]).view(np.uint64), bf_weights.view(np.uint64))
    assert np.array_equal(np.sort(np.diff(g.offs)), bf_degrees)


def _edge_set(g):
    h, t, w = _directed_pairs(g)
    return {(int(a), int(b)): float(c) for a, b, c in zip(h, t, w)}


@pytest.mark.parametrize("seed,n,pct", [(2, 3000, 1.0), (5, 2000, 10.0), (9, 1500, 37.5)])
def test_rgg_extra_edges(seed, n, pct):
    base = po.gen_rgg(seed, n, 0.0)
    full = po.gen_rgg(seed, n, pct)
    m = base.ne // 2
    nx = int(pct * m) // 100
    assert full.ne == base.ne + 2 * nx
    eb, ef = _edge_set(base), _edge_set(full)
    for k, w in eb.items():
        assert ef[k] == w
    extra = {k: w for k, w in ef.items() if k not in eb}
    assert len(extra) == 2 * nx
    for (a, b), w in extra.items():
        assert a != b
        assert ef[(b, a)] == w
        assert 0.0 <= w < 1.0


@pytest.mark.parametrize("seed,n,m", [(1, 2, 1), (2, 3000, 4), (3, 20000, 8), (4, 20000, 2)])
def test_ba_structure(seed, n, m):
    g = po.gen_ba(seed, n, m)
    h, t, w = _directed_pairs(g)
    assert np.all(w == 1.0)
    assert np.all(h != t)
    pairs = set(zip(h.tolist(), t.tolist()))
    assert len(pairs) == g.ne                      # simple graph
    assert all((b, a) in pairs for a, b in pairs)  # symmetric
    assert g.ne <= 2 * n * m
    earlier = np.bincount(h[t < h], minlength=n)
    assert earlier.max() <= m
    for i in range(g.nv):                           # rows sorted by tail
        row = g.edges["tail"][g.offs[i]:g.offs[i + 1]]
        assert np.all(np.diff(row) > 0)
    if n >= 3000:
        deg = np.diff(g.offs)
        assert np.median(deg) <= 2 * m + 1
        # preferential attachment: the Barabasi-Albert degree law P(k) = 2m(m+1)/(k(k+1)(k+2)),
        # whose tail count is n m(m+1)/(K(K+1)) vertices of degree >= K, within Poisson noise
        # plus 5% (uniform attachment has an exponential tail: no vertex of degree 10m at
        # n = 20000, m = 8, against 222 here)
        for K in (2 * m, 5 * m, 10 * m):
            law = n * m * (m + 1) / (K * (K + 1))
            assert abs(int((deg >= K).sum()) - law) <= 4 * np.sqrt(law) + 0.05 * law, K
        # the largest hub grows as m sqrt(n) (uniform attachment: about m ln n)
        assert deg.max() >= 0.5 * m * np.sqrt(n)
