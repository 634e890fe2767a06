"""GPU: the batched form (vt_batch_*) against the one-graph path and the oracle.

Every graph of a batch must get exactly what vt_run gives it alone: the same return value
and iteration count for every Louvain call, the same level sizes and the same final
modularity -- with fixed-order tree sums and with sequential sums, with the iteration
replayed from captured HIP graphs and launched directly.  With sequential sums the results
also equal the oracle's (the reference at one thread).
"""
import numpy as np
import pytest

import vite_amd
from oracle import pyoracle as po
from tests import graphs

pytestmark = pytest.mark.gpu


def _mixed_graphs():
    gs = [graphs.graph(n) for n in graphs.ALL]
    gs += [po.gen_rgg(100 + k, n, extra) for k, (n, extra) in
           enumerate([(9, 5.0), (17, 0.0), (300, 1.0), (1200, 0.0), (4000, 1.0), (700, 10.0)])]
    gs += [po.gen_ba(200 + k, n, m) for k, (n, m) in enumerate([(50, 1), (900, 3), (2500, 8)])]
    return gs


@pytest.fixture(scope="module")
def mixed(ctx):
    host = _mixed_graphs()
    dev = [ctx.upload(vite_amd.CSR(g.offs, g.edges)) for g in host]
    return host, dev


RUNS = [dict(), dict(threshold_cycling=True), dict(one_phase=True), dict(threshold=1e-4)]


def _same(a, b):
    assert a.call_q == b.call_q
    assert a.call_iters == b.call_iters
    assert a.level_nv == b.level_nv
    assert a.level_ne == b.level_ne
    assert a.level_q == b.level_q
    assert a.total_iters == b.total_iters
    assert a.final_q == b.final_q or (np.isnan(a.final_q) and np.isnan(b.final_q))


@pytest.mark.parametrize("kw", RUNS, ids=[str(i) for i in range(len(RUNS))])
@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("sequential", [False, True])
def test_batch_equals_one_by_one(ctx, mixed, kw, use_graph, sequential):
    host, dev = mixed
    b = vite_amd.Batch(ctx, dev)
    res, st = b.run(use_graph=use_graph, sequential=sequential, **kw)
    b.free()
    assert st["phases"] >= 1 and st["rounds"] >= st["phases"]
    # phases run in lockstep, call k of every graph in batch phase k, and a phase lasts as
    # many rounds as its longest-running graph iterates (no round is counted twice or lost)
    calls = max(len(r.call_iters) for r in res)
    assert st["rounds"] == sum(max(r.call_iters[k] for r in res if len(r.call_iters) > k) for k in range(calls))
    for g, r in zip(dev, res):
        _same(r, g.run(sequential=sequential, **kw))


@pytest.mark.parametrize("kw", RUNS[:2], ids=["0", "1"])
def test_batch_sequential_equals_oracle(ctx, mixed, kw):
    host, dev = mixed
    b = vite_amd.Batch(ctx, dev)
    res, _ = b.run(sequential=True, **kw)
    b.free()
    for g, r in zip(host, res):
        o = po.run(g, **kw)
        assert r.call_q == o.call_q
        assert r.call_iters == o.call_iters
        assert r.level_nv == o.level_nv
        assert r.final_q == o.final_q


def test_batch_order_and_repeats(ctx, mixed):
    """A graph's result does not depend on its neighbours in the batch or its position,
    and a batch can be run again."""
    host, dev = mixed
    b1 = vite_amd.Batch(ctx, dev)
    r1, st1 = b1.run(threshold_cycling=True)
    # graphs finish at different rounds: the sweep lists drop finished graphs' vertices
    assert st1["relists"] >= 1
    r1b, _ = b1.run(threshold_cycling=True)
    b1.free()
    rev = list(reversed(dev))
    b2 = vite_amd.Batch(ctx, rev + rev[:3])
    r2, _ = b2.run(threshold_cycling=True)
    b2.free()
    for a, c in zip(r1, r1b):
        _same(a, c)
    for a, c in zip(r1, list(reversed(r2[:len(dev)]))):
        _same(a, c)
    for a, c in zip(r2[:3], r2[len(dev):]):
        _same(a, c)


def test_batch_autotune_keeps_results(ctx, mixed):
    """The batch's own autotuning changes launch shapes only: every graph still gets what it
    gets alone under the default setting."""
    host, dev = mixed
    default = ctx.tuning
    alone = [g.run(threshold_cycling=True) for g in dev]
    b = vite_amd.Batch(ctx, dev)
    try:
        t = b.autotune(prep=1, reps=1)
        assert t == ctx.tuning
        res, _ = b.run(threshold_cycling=True)
    finally:
        b.free()
        ctx.set_tuning(**default)
    for a, r in zip(alone, res):
        _same(a, r)


def test_batch_with_community_records(ctx, mixed):
    """The batched sweeps reading the per-vertex community records (vertex_records = 1, the
    records of the whole union rebuilt before every captured round) give every graph what
    it gets alone with separate arrays."""
    host, dev = mixed
    default = ctx.tuning
    alone = [g.run(threshold_cycling=True) for g in dev]
    b = vite_amd.Batch(ctx, dev)
    try:
        ctx.set_tuning(vertex_records=1)
        res, _ = b.run(threshold_cycling=True)
    finally:
        b.free()
        ctx.set_tuning(**default)
    for a, r in zip(alone, res):
        _same(a, r)


def test_batch_rejects(ctx):
    g_empty = ctx.upload(vite_amd.CSR(np.zeros(2, np.int64), np.zeros(0, po.EDGE_DTYPE)))
    with pytest.raises(vite_amd.VtError):
        vite_amd.Batch(ctx, [g_empty])
    g = ctx.upload(vite_amd.CSR(graphs.graph("rgg").offs, graphs.graph("rgg").edges))
    b = vite_amd.Batch(ctx, [g])
    for kw in [dict(coloring=True), dict(et_type=2, et_delta=0.5), dict(vertex_order=True)]:
        with pytest.raises(vite_amd.VtError):
            b.run(**kw)
    b.free()
