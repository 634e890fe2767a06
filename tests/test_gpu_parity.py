"""GPU parity: every device entry point against the CPU oracle on the same inputs.

Integer and index outputs (targets, communities, CSR structure, colors) must be
bit-identical.  FP64 outputs computed in the same order as the oracle (row folds, the
per-community counters, the delta folds, the coarse-edge folds) must be bit-identical too;
sums the device forms as fixed-order trees (2m, modularity) are compared with
VT_FLAG_SEQUENTIAL_SUMS for exactness and without it within 1e-12 relative.
"""
import numpy as np
import pytest

import vite_amd
from oracle import pyoracle as po
from tests import graphs

pytestmark = pytest.mark.gpu


def _upload(ctx, name):
    g = graphs.graph(name)
    return g, ctx.upload(vite_amd.CSR(g.offs, g.edges))


@pytest.mark.parametrize("name", graphs.ALL)
def test_upload_roundtrip(ctx, name):
    g, dg = _upload(ctx, name)
    back = dg.download()
    assert np.array_equal(back.offs, g.offs)
    assert np.array_equal(back.edges.view(np.uint8), g.edges.view(np.uint8))


@pytest.mark.parametrize("name", graphs.ALL)
def test_init_degrees(ctx, name):
    g, dg = _upload(ctx, name)
    vdeg_o, c_o = po.init_degrees(g)
    vdeg, c_seq = dg.init_degrees(sequential=True)
    assert np.array_equal(vdeg, vdeg_o)
    assert c_seq == c_o
    _, c_tree = dg.init_degrees(sequential=False)
    assert abs(c_tree - c_o) <= 1e-12 * abs(c_o)


def test_long_event_runs(ctx):
    """A community that thousands of vertices join in one sweep (the centre of a star with
    weighted leaves, plus a ring among the leaves): its run of move events is folded by a
    wave, the short runs by a thread each; the per-community degree and size changes stay
    bit-identical to the oracle's in-order folds."""
    rng = np.random.default_rng(23)
    n = 5001
    pairs = [(0, v) for v in range(1, n)] + [(v, v + 1) for v in range(1, n - 1)]
    w = np.concatenate([rng.uniform(0.5, 1.0, n - 1), rng.uniform(0.01, 0.1, n - 2)])
    g = po.from_edge_list(n, np.array(pairs), w)
    dg = ctx.upload(vite_amd.CSR(g.offs, g.edges))
    vdeg, c = po.init_degrees(g)
    st = (np.arange(n), np.ones(n, np.int64), vdeg.copy(), vdeg, c)
    t_o, cw_o, ds_o, dd_o = po.sweep(g, *st)
    assert np.count_nonzero(t_o == 0) > 2 * 1024  # the long run this test is about
    t, cw, ds, dd = dg.sweep(*st)
    assert np.array_equal(t, t_o)
    assert np.array_equal(cw.view(np.uint64), cw_o.view(np.uint64))
    assert np.array_equal(ds, ds_o)
    assert np.array_equal(dd.view(np.uint64), dd_o.view(np.uint64))
    q_o, cv_o, it_o = po.louvain_phase(g)
    q, cv, it = dg.louvain_phase(sequential=True)
    assert (q, it) == (q_o, it_o) and np.array_equal(cv, cv_o)


def test_long_rows_degree_fold(ctx):
    """Rows of degree 2048 and more are folded by a wave each, the rest by a thread each:
    every vDegree is still the row's left fold from 0.0 in CSR order."""
    rng = np.random.default_rng(17)
    n = 9000
    pairs = [(0, int(v)) for v in rng.choice(np.arange(1, n), 7000, replace=False)]
    pairs += [(1, int(v)) for v in rng.choice(np.arange(2, n), 2048, replace=False)]
    pairs += [(2, int(v)) for v in rng.choice(np.arange(3, n), 2047, replace=False)]
    a, b = rng.integers(3, n, 20000), rng.integers(3, n, 20000)
    pairs += [(int(x), int(y)) for x, y in zip(a, b) if x != y]
    pairs = np.unique(np.sort(np.array(pairs), axis=1), axis=0)
    g = po.from_edge_list(n, pairs, rng.uniform(0.01, 1.0, len(pairs)))
    assert np.diff(g.offs).max() >= 7000
    dg = ctx.upload(vite_amd.CSR(g.offs, g.edges))
    vdeg_o, c_o = po.init_degrees(g)
    vdeg, c = dg.init_degrees(sequential=True)
    assert np.array_equal(vdeg.view(np.uint64), vdeg_o.view(np.uint64))
    assert c == c_o
    q_o, cv_o, it_o = po.louvain_phase(g)
    q, cv, it = dg.louvain_phase(sequential=True)
    assert (q, it) == (q_o, it_o) and np.array_equal(cv, cv_o)


@pytest.mark.parametrize("name", graphs.ALL)
@pytest.mark.parametrize("seed", [None, 1, 2])
def test_sweep_replay(ctx, name, seed):
    """One Jacobi sweep from the same state: targets, clusterWeight and the per-community
    delta folds are bit-exact."""
    g, dg = _upload(ctx, name)
    if seed is None:
        vdeg, c = po.init_degrees(g)
        nv = g.nv
        curr, csize, cdeg = np.arange(nv), np.ones(nv, np.int64), vdeg.copy()
    else:
        curr, csize, cdeg, vdeg, c = graphs.random_state(g, seed)
    t_o, cw_o, ds_o, dd_o = po.sweep(g, curr, csize, cdeg, vdeg, c)
    t, cw, ds, dd = dg.sweep(curr, csize, cdeg, vdeg, c)
    assert np.array_equal(t, t_o)
    assert np.array_equal(cw.view(np.uint64), cw_o.view(np.uint64))
    assert np.array_equal(ds, ds_o)
    assert np.array_equal(dd.view(np.uint64), dd_o.view(np.uint64))


@pytest.mark.parametrize("base", ["multi", "rgg", "dense"])
def test_signed_zero_weights(ctx, base):
    """Weights of -0.0 and +0.0: a new per-community counter starts as its first weight and
    the own community's from 0.0 (louvain.cpp:2456-2462), so counters and clusterWeight keep
    the sign of zero exactly as the oracle does."""
    g0 = graphs.graph(base)
    w = g0.edges["weight"].copy()
    k = np.arange(len(w))
    w[k % 3 == 0] = -0.0
    w[k % 7 == 0] = 0.0
    g = po.make_graph(g0.offs, g0.edges["tail"], w)
    dg = ctx.upload(vite_amd.CSR(g.offs, g.edges))
    for seed in (None, 1, 2):
        if seed is None:
            vdeg, c = po.init_degrees(g)
            curr, csize, cdeg = np.arange(g.nv), np.ones(g.nv, np.int64), vdeg.copy()
        else:
            curr, csize, cdeg, vdeg, c = graphs.random_state(g, seed)
        t_o, cw_o, ds_o, dd_o = po.sweep(g, curr, csize, cdeg, vdeg, c)
        t, cw, ds, dd = dg.sweep(curr, csize, cdeg, vdeg, c)
        assert np.array_equal(t, t_o)
        assert np.array_equal(cw.view(np.uint64), cw_o.view(np.uint64))
        assert np.array_equal(ds, ds_o)
        assert np.array_equal(dd.view(np.uint64), dd_o.view(np.uint64))
    q_o, cv_o, it_o = po.louvain_phase(g)
    q, cv, it = dg.louvain_phase(sequential=True)
    assert it == it_o and q == q_o
    assert np.array_equal(cv, cv_o)


@pytest.mark.parametrize("name", graphs.ALL)
def test_modularity(ctx, name):
    g = graphs.graph(name)
    curr, csize, cdeg, vdeg, c = graphs.random_state(g, 4)
    _, cw, _, _ = po.sweep(g, curr, csize, cdeg, vdeg, c)
    q_o = po.modularity(cw, cdeg, c)
    assert ctx.modularity(cw, cdeg, c, sequential=True) == q_o
    assert abs(ctx.modularity(cw, cdeg, c) - q_o) <= 1e-12


PHASES = [("plain", 0), ("plain", 1), ("plain", 2), ("plain", 3), ("plain", 4),
          ("color", 0), ("color", 1), ("color", 2), ("color", 4), ("vorder", 0), ("vorder", 3), ("vorder", 2)]


@pytest.mark.parametrize("name", graphs.ALL)
@pytest.mark.parametrize("mode,et", PHASES)
def test_phase_exact(ctx, name, mode, et):
    """A whole phase with the reference's summation order: return value, iteration count and
    partition identical to the oracle."""
    g, dg = _upload(ctx, name)
    kw = dict(lower=-1.0, thresh=1e-6, et_type=et, et_delta=0.25)
    if mode != "plain":
        nc, colors = po.color(g, 4)
        kw.update(mode=mode, num_color=nc + 1, colors=colors)
    q_o, cv_o, it_o = po.louvain_phase(g, **kw)
    q, cv, it = dg.louvain_phase(sequential=True, **kw)
    assert it == it_o
    assert np.array_equal(cv, cv_o)
    assert q == q_o


@pytest.mark.parametrize("name", graphs.ALL)
def test_phase_tree_sums(ctx, name):
    """Default device summation (fixed-order trees for 2m and modularity): modularity within
    1e-6 of the oracle's."""
    g, dg = _upload(ctx, name)
    q_o, cv_o, it_o = po.louvain_phase(g)
    q, cv, it = dg.louvain_phase()
    assert abs(q - q_o) <= 1e-6
    q2, _, _ = dg.louvain_phase()
    assert q2 == q  # deterministic run to run


@pytest.mark.parametrize("name", graphs.ALL)
def test_rebuild(ctx, name):
    g, dg = _upload(ctx, name)
    _, cv, _ = po.louvain_phase(g)
    r_o = po.rebuild(g, cv)
    r = dg.rebuild(cv).download()
    assert np.array_equal(r.offs, r_o.offs)
    assert np.array_equal(r.edges["tail"], r_o.edges["tail"])
    assert np.array_equal(r.edges["weight"].view(np.uint64), r_o.edges["weight"].view(np.uint64))


@pytest.mark.parametrize("name", graphs.ALL)
@pytest.mark.parametrize("part", ["identity", "random", "one", "blocks64", "random200"])
def test_rebuild_partitions(ctx, name, part):
    """Partitions no phase produces: singletons (every fine row becomes a coarse row; the hub
    has more distinct neighbours than the per-community table holds, so it takes the sort
    path), random communities, one community holding every vertex, blocks of 64 consecutive
    vertices (long concatenated rows, few neighbours), and 200 random communities (on most
    graphs, communities with more distinct neighbours than the fold's table holds)."""
    g, dg = _upload(ctx, name)
    rng = np.random.default_rng(7)
    cv = {"identity": np.arange(g.nv), "random": rng.integers(0, max(1, g.nv // 3), g.nv),
          "one": np.zeros(g.nv), "blocks64": np.arange(g.nv) // 64,
          "random200": rng.integers(0, min(200, g.nv), g.nv)}[part].astype(np.int64)
    r_o = po.rebuild(g, cv)
    r = dg.rebuild(cv).download()
    assert np.array_equal(r.offs, r_o.offs)
    assert np.array_equal(r.edges.view(np.uint8), r_o.edges.view(np.uint8))


@pytest.mark.parametrize("name", graphs.ALL)
def test_rebuild_from_device_partition(ctx, name):
    g, dg = _upload(ctx, name)
    _, cv, _ = dg.louvain_phase(sequential=True)
    r = dg.rebuild(None).download()
    r_o = po.rebuild(g, cv)
    assert np.array_equal(r.offs, r_o.offs)
    assert np.array_equal(r.edges.view(np.uint8), r_o.edges.view(np.uint8))


def test_rebuild_rejects_ids_outside_the_graph(ctx):
    """A caller's partition with a community id of -1 or nv is refused with an error (the
    partition a phase leaves on the device is not checked: its ids are vertex ids); the
    context and the graph stay usable."""
    g, dg = _upload(ctx, "karate")
    for bad in (-1, g.nv):
        cv = np.zeros(g.nv, np.int64)
        cv[5] = bad
        with pytest.raises(vite_amd.VtError):
            dg.rebuild(cv)
    cv = np.arange(g.nv) % 3
    r = dg.rebuild(cv).download()
    r_o = po.rebuild(g, cv)
    assert np.array_equal(r.offs, r_o.offs)
    assert np.array_equal(r.edges.view(np.uint8), r_o.edges.view(np.uint8))


@pytest.mark.parametrize("name", graphs.ALL)
@pytest.mark.parametrize("nhash,single", [(4, False), (2, False), (4, True), (1, False)])
def test_color(ctx, name, nhash, single):
    g, dg = _upload(ctx, name)
    nc_o, col_o = po.color(g, nhash, 70, single)
    nc, col = dg.color(nhash, 70, single)
    assert nc == nc_o
    assert np.array_equal(col, col_o)


@pytest.mark.parametrize("seed,n,extra", [(1, 1, 0.0), (2, 5000, 0.0), (3, 6000, 1.0), (4, 20000, 1.0), (5, 3000, 10.0)])
def test_gen_rgg(ctx, seed, n, extra):
    g_o = po.gen_rgg(seed, n, extra)
    g = ctx.gen_rgg(seed, n, extra).download()
    assert np.array_equal(g.offs, g_o.offs)
    assert np.array_equal(g.edges.view(np.uint8), g_o.edges.view(np.uint8))


@pytest.mark.parametrize("seed,n,m", [(1, 1, 1), (2, 3000, 4), (3, 20000, 8)])
def test_gen_ba(ctx, seed, n, m):
    g_o = po.gen_ba(seed, n, m)
    g = ctx.gen_ba(seed, n, m).download()
    assert np.array_equal(g.offs, g_o.offs)
    assert np.array_equal(g.edges.view(np.uint8), g_o.edges.view(np.uint8))


RUNS = [dict(), dict(threshold_cycling=True), dict(coloring=True, et_type=2, et_delta=0.25),
        dict(et_type=1), dict(vertex_order=True, et_type=3), dict(coloring=True, single_color_iter=True, et_type=4,
                                                                   et_delta=0.5), dict(one_phase=True)]


@pytest.mark.parametrize("name", ["rgg", "ba", "vite_rgg", "hub", "grid", "karate"])
@pytest.mark.parametrize("kw", RUNS, ids=[str(i) for i in range(len(RUNS))])
def test_run_exact(ctx, name, kw):
    """The whole phase loop (main.cpp:242-513): every Louvain call's return value and
    iteration count equal the oracle's when sums follow the reference order."""
    g, dg = _upload(ctx, name)
    r_o = po.run(g, **kw)
    r = dg.run(sequential=True, **kw)
    assert r.call_iters == r_o.call_iters
    assert r.call_q == r_o.call_q
    assert r.level_nv == r_o.level_nv
    assert r.final_q == r_o.final_q


@pytest.mark.parametrize("name", ["rgg", "ba", "vite_rgg"])
def test_run_tree_sums(ctx, name):
    g, dg = _upload(ctx, name)
    r_o = po.run(g, threshold_cycling=True)
    r = dg.run(threshold_cycling=True)
    assert abs(r.final_q - r_o.final_q) <= 1e-6


@pytest.mark.parametrize("name", ["rgg", "ba", "hub", "multi"])
@pytest.mark.parametrize("et", [0, 1, 2])
def test_phase_stepwise_equals_oneshot(ctx, name, et):
    """vt_phase_begin/step/end reproduce vt_louvain_phase exactly."""
    g, dg = _upload(ctx, name)
    q1, cv1, it1 = dg.louvain_phase(et_type=et, et_delta=0.4)
    st = dg.phase_begin(et_type=et, et_delta=0.4)
    done = False
    while not done:
        _, done = st.step()
    q2, cv2, it2 = st.end()
    assert (q1, it1) == (q2, it2)
    assert np.array_equal(cv1, cv2)


@pytest.mark.parametrize("name", ["rgg", "ba", "hub", "vite_rgg"])
@pytest.mark.parametrize("prep", [0, 1, 3])
def test_phase_sweep_state_replay(ctx, name, prep):
    """A sweep from a device mid-run state equals the oracle's sweep from the downloaded
    state, and repeating it changes nothing."""
    g, dg = _upload(ctx, name)
    st = dg.phase_begin(thresh=-1.0)
    for _ in range(prep):
        st.step()
    st.sweep()
    s1 = st.state()
    st.sweep()
    s2 = st.state()
    assert np.array_equal(s1["target"], s2["target"])
    assert np.array_equal(s1["cw"].view(np.uint64), s2["cw"].view(np.uint64))
    t_o, cw_o, _, _ = po.sweep(g, s1["curr"], s1["csize"], s1["cdeg"], s1["vdeg"], s1["constant"])
    assert np.array_equal(s1["target"], t_o)
    assert np.array_equal(s1["cw"].view(np.uint64), cw_o.view(np.uint64))
    st.end(want_cvect=False)


TUNINGS = [dict(bucket_max_degree=[3, 7, 15, 31, 63, 1023], blocks_per_bucket=16384),
           dict(bucket_max_degree=[5, 11, 23, 47, 95, 300], blocks_per_bucket=8, wide_table_max_degree=[383, 767]),
           dict(bucket_max_degree=[-1, -1, -1, -1, 127, 1023], blocks_per_bucket=64, wide_table_max_degree=[-1, -1]),
           dict(bucket_max_degree=[7, 15, 31, 63, 127, 130], blocks_per_bucket=4096),
           dict(bucket_max_degree=[-1, -1, -1, -1, -1, 1023], blocks_per_bucket=256, wide_table_max_degree=[150, 600]),
           # bucket 2 (16-lane groups, a launch of its own) alone, then everything but it
           dict(bucket_max_degree=[-1, -1, 47, -1, 127, 1023], blocks_per_bucket=512),
           dict(bucket_max_degree=[7, 15, -1, 63, 127, 1023], blocks_per_bucket=512),
           # the group kernels reading the per-vertex community records
           dict(bucket_max_degree=[7, 15, 47, 63, 191, 1023], blocks_per_bucket=16384, vertex_records=1),
           dict(bucket_max_degree=[3, 7, 15, 31, 63, 1023], blocks_per_bucket=8, vertex_records=1),
           # rows above degree 130 are hubs: a workgroup per hub, then the flat launches, whatever
           # the hub count
           dict(bucket_max_degree=[7, 15, 31, 63, 127, 130], blocks_per_bucket=4096, hub_block_min=0),
           dict(bucket_max_degree=[7, 15, 31, 63, 127, 130], blocks_per_bucket=4096, hub_block_min=1 << 30)]


@pytest.mark.parametrize("tuning", TUNINGS, ids=[str(i) for i in range(len(TUNINGS))])
@pytest.mark.parametrize("name", ["rgg", "ba", "hub", "dense", "multi"])
def test_tuning_does_not_change_results(ctx, name, tuning):
    """Bucket bounds and block caps only change the launch shape: sweeps and phases stay
    bit-identical to the oracle under every valid setting."""
    g, dg = _upload(ctx, name)
    default = ctx.tuning
    try:
        ctx.set_tuning(**tuning)
        curr, csize, cdeg, vdeg, c = graphs.random_state(g, 7)
        t_o, cw_o, ds_o, dd_o = po.sweep(g, curr, csize, cdeg, vdeg, c)
        t, cw, ds, dd = dg.sweep(curr, csize, cdeg, vdeg, c)
        assert np.array_equal(t, t_o)
        assert np.array_equal(cw.view(np.uint64), cw_o.view(np.uint64))
        q_o, cv_o, it_o = po.louvain_phase(g)
        q, cv, it = dg.louvain_phase(sequential=True)
        assert (q, it) == (q_o, it_o)
        assert np.array_equal(cv, cv_o)
    finally:
        ctx.set_tuning(**default)


def test_tuning_records_in_color_and_et_phases(ctx):
    """The community records are rebuilt before each color's sweep (the degrees change
    between colors) and read under early termination: whole phases stay exact."""
    default = ctx.tuning
    try:
        ctx.set_tuning(vertex_records=1)
        for name in ("rgg", "ba"):
            g, dg = _upload(ctx, name)
            nc, colors = po.color(g, 4)
            for kw in (dict(mode="color", num_color=nc + 1, colors=colors), dict(et_type=2, et_delta=0.25),
                       dict(mode="vorder", num_color=nc + 1, colors=colors, et_type=1)):
                q_o, cv_o, it_o = po.louvain_phase(g, **kw)
                q, cv, it = dg.louvain_phase(sequential=True, **kw)
                assert (q, it) == (q_o, it_o), (name, kw)
                assert np.array_equal(cv, cv_o), (name, kw)
    finally:
        ctx.set_tuning(**default)


def test_tuning_rejects_bounds_beyond_kernel_capacity(ctx):
    default = ctx.tuning
    for bad in [[8, 15, 31, 63, 127, 1023], [7, 15, 31, 200, 127, 1023], [7, 15, 31, 63, 127, 1024]]:
        with pytest.raises(vite_amd.VtError):
            ctx.set_tuning(bucket_max_degree=bad)
    with pytest.raises(vite_amd.VtError):
        ctx.set_tuning(blocks_per_bucket=4)
    for bad in [[384, 511], [255, 768], [511, 300], [-2, 511]]:
        with pytest.raises(vite_amd.VtError):
            ctx.set_tuning(wide_table_max_degree=bad)
    for bad in [-1, (1 << 30) + 1]:
        with pytest.raises(vite_amd.VtError):
            ctx.set_tuning(hub_block_min=bad)
    assert ctx.tuning == default


def test_autotune_keeps_results(ctx):
    g, dg = _upload(ctx, "dense")
    default = ctx.tuning
    try:
        t = ctx.autotune(dg, prep=1, reps=1)
        assert t == ctx.tuning
        q_o, cv_o, it_o = po.louvain_phase(g)
        q, cv, it = dg.louvain_phase(sequential=True)
        assert (q, it) == (q_o, it_o)
        assert np.array_equal(cv, cv_o)
    finally:
        ctx.set_tuning(**default)


@pytest.mark.parametrize("name", ["rgg", "ba", "dense", "multi", "grid"])
def test_lane_per_vertex_kernel(ctx, name, monkeypatch):
    """The lane-per-vertex kernel (selected for rows up to VITE_AMD_LPV_MAXDEG entries) gives
    the same sweep and phase as the oracle."""
    monkeypatch.setenv("VITE_AMD_LPV_MAXDEG", "64")
    g, dg = _upload(ctx, name)
    curr, csize, cdeg, vdeg, c = graphs.random_state(g, 3)
    t_o, cw_o, ds_o, dd_o = po.sweep(g, curr, csize, cdeg, vdeg, c)
    t, cw, ds, dd = dg.sweep(curr, csize, cdeg, vdeg, c)
    assert np.array_equal(t, t_o)
    assert np.array_equal(cw.view(np.uint64), cw_o.view(np.uint64))
    assert np.array_equal(dd.view(np.uint64), dd_o.view(np.uint64))
    q_o, cv_o, it_o = po.louvain_phase(g, et_type=1)
    q, cv, it = dg.louvain_phase(sequential=True, et_type=1)
    assert (q, it) == (q_o, it_o)
    assert np.array_equal(cv, cv_o)
