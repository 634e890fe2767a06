"""GPU: the serving form (vt_server_*).

Each request's result is the same alone and in the mix (exact: every job runs on its own
stream, and a batched job of small requests gives each graph what vt_run gives it alone),
equals vt_run on the same generated graph, and the admission rule holds: arrival order, and
running reservations never above the budget, with every job's measured HBM high-water mark
inside its reservation.
"""
import itertools

import numpy as np
import pytest

import vite_amd

pytestmark = pytest.mark.gpu


def _requests(seed=5, count=24):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(round(2.0 ** rng.uniform(8, 12.5)))
        if k % 2 == 0:
            out.append(("rgg", 1000 + k, n, float(rng.choice([0.0, 1.0]))))
        else:
            out.append(("ba", 2000 + k, n, int(rng.integers(1, 9))))
    return out


RUN_KW = [dict(threshold_cycling=True), dict()]


def _same(a, b):
    assert a.call_q == b.call_q
    assert a.call_iters == b.call_iters
    assert a.level_nv == b.level_nv
    assert a.final_q == b.final_q


@pytest.mark.parametrize("kw", RUN_KW, ids=["cycling", "plain"])
def test_mix_equals_alone(ctx, kw):
    reqs = _requests()
    with vite_amd.Server(slots=4) as srv:
        tickets = [srv.submit(*r, **kw) for r in reqs]
        mixed = [srv.wait(t) for t in tickets]
    with vite_amd.Server(slots=1) as one:
        alone = [one.wait(one.submit(*r, **kw)) for r in reqs]
    for (rm, im), (ra, ia) in zip(mixed, alone):
        assert im["status"] == 0 and ia["status"] == 0
        assert (im["nv"], im["ne"]) == (ia["nv"], ia["ne"])
        _same(rm, ra)
    # and equal to the one-graph entry point on the same generated graph
    for (kind, seed, n, p), (rm, _) in list(zip(reqs, mixed))[:6]:
        g = ctx.gen_rgg(seed, n, p) if kind == "rgg" else ctx.gen_ba(seed, n, p)
        _same(rm, g.run(**kw))
        g.free()


def test_admission_order_and_budget(ctx):
    reqs = _requests(seed=9, count=16)
    res = [vite_amd.request_reservation(*r, threshold_cycling=True) for r in reqs]
    budget = max(res) + min(res)  # the largest job fits only beside the smallest
    with vite_amd.Server(hbm_budget=budget, slots=4) as srv:
        tickets = [srv.submit(*r, threshold_cycling=True) for r in reqs]
        infos = [srv.wait(t)[1] for t in tickets]
    # each admission takes the oldest queued request (a batched job's head, its
    # lowest-numbered member; its other members may have arrived later)
    groups = {}
    for k, i in enumerate(infos):
        groups.setdefault((i["slot"], i["start_s"]), []).append(k)
    heads = [min(m) for _, m in sorted(groups.items(), key=lambda kv: kv[0][1])]
    assert heads == sorted(heads)
    for i in infos:
        assert i["status"] == 0
        assert i["peak_bytes"] <= i["reserved_bytes"]
    # at every job start, the jobs running then fit in the budget together (the members of
    # a batched job share one slot, start and reservation)
    jobs = {(i["slot"], i["start_s"]): i for i in infos}
    for t in [j["start_s"] for j in jobs.values()]:
        running = [j for j in jobs.values() if j["start_s"] <= t < j["end_s"]]
        assert sum(j["reserved_bytes"] for j in running) <= budget


@pytest.mark.parametrize("kw", RUN_KW, ids=["cycling", "plain"])
def test_queued_small_requests_batched(ctx, kw):
    """Small requests queued behind a running one are taken together as one batched job,
    and every result still equals vt_run on that graph alone; larger requests and requests
    with other run parameters are not batched."""
    reqs = _requests(seed=13, count=20)
    with vite_amd.Server(slots=1, max_batch=6, batch_max_n=3000) as srv:
        tickets = [srv.submit(*r, **kw) for r in reqs]
        out = [srv.wait(t) for t in tickets]
        # a request with other run parameters, queued between batchable ones
        t1 = srv.submit(*reqs[0], **kw)
        t2 = srv.submit(*reqs[1], threshold=1e-4)
        t3 = srv.submit(*reqs[2], **kw)
        r1, r2, r3 = srv.wait(t1), srv.wait(t2), srv.wait(t3)
    assert max(i["batch"] for _, i in out) > 1
    assert all(i["batch"] <= 6 for _, i in out)
    for (kind, seed, n, p), (r, i) in zip(reqs, out):
        assert i["status"] == 0
        if n > 3000:
            assert i["batch"] == 1
        g = ctx.gen_rgg(seed, n, p) if kind == "rgg" else ctx.gen_ba(seed, n, p)
        _same(r, g.run(**kw))
        g.free()
    _same(r1[0], out[0][0])
    assert r2[1]["start_s"] != r1[1]["start_s"] and r2[1]["start_s"] != r3[1]["start_s"]


def test_small_requests_batched_past_a_large_one(ctx):
    """A batched job takes the matching small requests wherever they wait in the queue: with a
    large request queued between small ones, the small ones behind it join the job of the
    small one ahead of it, and the large one runs after that job, as the oldest request left."""
    kw = dict(threshold_cycling=True)
    big = [("rgg", 31, 40000, 1.0), ("rgg", 33, 30000, 0.0)]
    small = [("rgg", 32, 900, 1.0), ("ba", 34, 700, 4), ("rgg", 35, 1500, 0.0)]
    with vite_amd.Server(slots=1, max_batch=8, batch_max_n=3000) as srv:
        t_big0 = srv.submit(*big[0], **kw)  # holds the one slot while the others queue
        t_s0 = srv.submit(*small[0], **kw)
        t_big1 = srv.submit(*big[1], **kw)
        t_s1 = srv.submit(*small[1], **kw)
        t_s2 = srv.submit(*small[2], **kw)
        out = {t: srv.wait(t) for t in (t_big0, t_s0, t_big1, t_s1, t_s2)}
    start = {t: out[t][1]["start_s"] for t in out}
    assert out[t_s0][1]["batch"] == 3
    assert start[t_s0] == start[t_s1] == start[t_s2]
    assert start[t_big0] < start[t_s0] < start[t_big1]
    for t, req in zip((t_big0, t_s0, t_big1, t_s1, t_s2), (big[0], small[0], big[1], small[1], small[2])):
        kind, seed, n, p = req
        g = ctx.gen_rgg(seed, n, p) if kind == "rgg" else ctx.gen_ba(seed, n, p)
        _same(out[t][0], g.run(**kw))
        g.free()


def test_batching_settings_rejected(ctx):
    with vite_amd.Server(slots=1) as srv:
        for mb, mn in [(0, 100), (5000, 100), (4, 0)]:
            with pytest.raises(vite_amd.VtError):
                srv.set_batching(mb, mn)


def test_oversized_request_rejected(ctx):
    with vite_amd.Server(hbm_budget=256 << 20, slots=2) as srv:
        with pytest.raises(vite_amd.VtError):
            srv.submit("rgg", 1, 1 << 20, 1.0)
        r, info = srv.wait(srv.submit("rgg", 2, 300, 0.0))
        assert info["status"] == 0 and r.final_q > 0


def test_threads_submit_and_wait(ctx):
    """Submit and wait from several host threads at once, with the same results as one
    thread gets; a ticket is collected once: a wait on a ticket another thread is waiting on
    is refused with VT_ERR_STATE, and one on a collected ticket with VT_ERR_ARG."""
    import threading
    import time

    reqs = _requests(seed=8, count=12)
    with vite_amd.Server(slots=3) as srv:
        tickets = [None] * len(reqs)
        results = [None] * len(reqs)

        def submit(k):
            tickets[k] = srv.submit(*reqs[k])

        def wait(k):
            results[k] = srv.wait(tickets[k])

        for fn in (submit, wait):
            threads = [threading.Thread(target=fn, args=(k,)) for k in range(len(reqs))]
            for t in threads:
                t.start()
            for t in threads:
                t.join()
        for r, info in results:
            assert info["status"] == 0
        with pytest.raises(vite_amd.VtError) as e:
            srv.wait(tickets[0])
        assert e.value.code == vite_amd.VT_ERR_ARG

        big = srv.submit("rgg", 77, 1 << 18, 1.0)
        first = []
        waiter = threading.Thread(target=lambda: first.append(srv.wait(big)))
        waiter.start()
        time.sleep(0.05)
        with pytest.raises(vite_amd.VtError) as e:
            srv.wait(big)
        # refused while the other thread waits, or already collected by it
        assert e.value.code in (vite_amd.VT_ERR_STATE, vite_amd.VT_ERR_ARG)
        waiter.join()
        assert first and first[0][1]["status"] == 0
    with vite_amd.Server(slots=1) as one:
        for r, (rm, _) in zip(reqs, results):
            _same(rm, one.wait(one.submit(*r))[0])


# Requests at the serving configuration's sizes: RGG (half with 1% extra edges) and BA
# (m from 1 to 8), 50/50, n from 2^10 to 2^20.  Big ones first, so the small ones queue
# behind them and a one-slot server takes them as batched jobs.
CONFIG_REQS = [("rgg", 5101, 1 << 20, 1.0), ("ba", 5102, 1 << 19, 8), ("rgg", 5103, 1 << 15, 0.0),
               ("ba", 5104, 1 << 13, 4), ("rgg", 5105, 1 << 10, 1.0), ("ba", 5106, 150000, 2),
               ("rgg", 5107, 3000, 0.0), ("ba", 5108, 1500, 8)]


def _oracle_run(req, kw):
    from oracle import pyoracle as po
    kind, seed, n, p = req
    g = po.gen_rgg(seed, n, p) if kind == "rgg" else po.gen_ba(seed, n, p)
    return po.run(g, **kw)


@pytest.mark.timeout(600)
def test_serving_matches_oracle_at_config_sizes(ctx):
    """Every served request's result equals the oracle's run of the same seeded graph: every
    Louvain call's Q and iteration count, the level sizes and the final Q, with sums in the
    reference's order.  The same requests are served in a 4-slot mix and on one slot, where
    the queued small requests run as batched jobs (one vt_batch_run over their union)."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import pyoracle as po
    po._lib()  # load once before the threads
    kw = dict(sequential=True)
    with ThreadPoolExecutor(max_workers=8) as pool:  # the oracle's C calls release the GIL
        expected = list(pool.map(lambda r: _oracle_run(r, {}), CONFIG_REQS))
    with vite_amd.Server(slots=4) as srv:
        mixed = [srv.wait(t) for t in [srv.submit(*r, **kw) for r in CONFIG_REQS]]
    with vite_amd.Server(slots=1) as one:
        queued = [one.wait(t) for t in [one.submit(*r, **kw) for r in CONFIG_REQS]]
    for req, want, (rm, im), (rq, iq) in zip(CONFIG_REQS, expected, mixed, queued):
        for r, info in ((rm, im), (rq, iq)):
            assert info["status"] == 0, req
            assert r.call_q == want.call_q, req
            assert r.call_iters == want.call_iters, req
            assert r.level_nv == want.level_nv, req
            assert r.final_q == want.final_q, req
    assert max(i["batch"] for _, i in queued) > 1  # some small requests ran batched
