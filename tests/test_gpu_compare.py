"""GPU: ground-truth comparison (vt_compare_communities, compare.cpp:58-337).

Against the statistics the compiled reference printed for -g runs (golden vectors), against
the oracle on random partitions, and at 2^20 vertices against the contingency-table closed
form: the pair counts are exact, and precision, recall, F-score and the Gini coefficients
are bit-identical because they are formed with the reference's expressions and order.
"""
import json
import os

import numpy as np
import pytest

import vite_amd
from oracle import pyoracle as po
from tests import graphs

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_runs.json")))
INT_KEYS = ("same_same", "same_diff", "diff_same", "nc_truth", "nc_found")
FLOAT_KEYS = ("precision", "recall", "f_score", "gini_truth", "gini_found")


def _same(a, b):
    """Equal counts and bit-identical floats (NaN included: one vertex has no pairs)."""
    for k in INT_KEYS:
        assert a[k] == b[k], k
    for k in FLOAT_KEYS:
        assert float(a[k]).hex() == float(b[k]).hex(), k


def _truth_vector(kind, nv):
    if kind == "factions":
        return graphs.karate_factions()
    v = np.arange(nv, dtype=np.int64)
    return v // 37 if kind == "blocks" else 3 * (v // 50) + 1


@pytest.mark.parametrize("idx", range(len(GOLDEN["ground_truth"])))
def test_compare_matches_reference(ctx, idx):
    case = GOLDEN["ground_truth"][idx]
    nv = GOLDEN["graphs"][case["graph"]]["nv"]
    st = ctx.compare_communities(_truth_vector(case["truth"], nv), np.array(case["communities"].split(), dtype=np.int64))
    for k in INT_KEYS:
        assert st[k] == case["stats"][k], k
    for k in FLOAT_KEYS:
        assert float(st[k]).hex() == case["stats"][k], k


@pytest.mark.parametrize("seed,n,k1,k2", [(1, 1, 1, 1), (2, 500, 7, 40), (3, 5000, 60, 25), (4, 3000, 3000, 2)])
def test_compare_matches_oracle(ctx, seed, n, k1, k2):
    rng = np.random.default_rng(seed)
    t = rng.integers(0, k1, n)
    f = rng.integers(0, k2, n) * 3  # ids with gaps: empty communities count in the Gini
    _same(ctx.compare_communities(t, f), po.compare(t, f))


def _closed_form(t, f):
    """Pair counts from the contingency table; the reference's float expressions in order."""
    def pairs(x):
        return int(np.sum(x * (x - 1) // 2))

    nc1, nc2 = int(t.max()) + 1, int(f.max()) + 1
    _, n_ab = np.unique(t * nc2 + f, return_counts=True)
    s1 = np.bincount(t, minlength=nc1)
    s2 = np.bincount(f, minlength=nc2)
    ss = pairs(n_ab)
    sd, ds = pairs(s1) - ss, pairs(s2) - ss

    def gini(sizes):
        sizes = np.sort(sizes)
        num = den = 0.0
        for i, s in enumerate(sizes.tolist()):
            num = num + float((i + 1) * s)
            den = den + float(s)
        k = len(sizes)
        return ((2 * num) / (k * den)) - (float(k + 1) / float(k))

    p = ss / (ss + ds)
    r = ss / (ss + sd)
    return dict(same_same=ss, same_diff=sd, diff_same=ds, nc_truth=nc1, nc_found=nc2, precision=p, recall=r,
                f_score=2.0 * ((p * r) / (p + r)), gini_truth=gini(s1), gini_found=gini(s2))


def test_compare_large(ctx):
    rng = np.random.default_rng(9)
    n = 1 << 20
    t = rng.integers(0, 1000, n)
    f = (t + (rng.random(n) < 0.2) * rng.integers(0, 5000, n)) % 4096
    _same(ctx.compare_communities(t, f), _closed_form(t, f))


def test_compare_rejects_bad_input(ctx):
    with pytest.raises(vite_amd.VtError) as e:
        ctx.compare_communities(np.array([0, -1, 2]), np.array([0, 1, 2]))
    assert e.value.code == vite_amd.VT_ERR_ARG
    with pytest.raises(vite_amd.VtError):
        ctx.compare_communities(np.zeros(0, np.int64), np.zeros(0, np.int64))
    with pytest.raises(vite_amd.VtError) as e:
        ctx.compare_communities(np.array([0, 1 << 40]), np.array([0, 1]))
    assert e.value.code == vite_amd.VT_ERR_LIMIT
