"""GPU parity at larger sizes: a 2^20-vertex RGG (29 M directed entries) swept from a
mid-run device state, and a hub-heavy Barabasi-Albert graph (2^17 vertices, m = 8, rows up
to a few thousand entries, so the sorted hub path runs) through whole phases and runs.
The graphs are generated on the device (the generators' parity is tested at smaller sizes)
and downloaded for the oracle."""
import numpy as np
import pytest

import vite_amd
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _host(dg):
    csr = dg.download()
    return po.Graph(csr.offs, csr.edges)


def test_rgg_2p20_sweep_from_device_state(ctx):
    dg = ctx.gen_rgg(77, 1 << 20, 1.0)
    g = _host(dg)
    st = dg.phase_begin(thresh=1e-3)
    for _ in range(2):
        st.step()
    st.sweep()
    s = st.state()
    st.end(want_cvect=False)
    t_o, cw_o, _, _ = po.sweep(g, s["curr"], s["csize"], s["cdeg"], s["vdeg"], s["constant"])
    assert np.array_equal(s["target"], t_o)
    assert np.array_equal(s["cw"].view(np.uint64), cw_o.view(np.uint64))
    dg.free()


@pytest.fixture(scope="module")
def ba_hubs(ctx):
    dg = ctx.gen_ba(78, 1 << 17, 8)
    g = _host(dg)
    assert int(np.diff(g.offs).max()) > 1023  # rows for the sorted hub path
    yield dg, g
    dg.free()


@pytest.mark.parametrize("et", [0, 2])
def test_ba_hubs_phase_exact(ctx, ba_hubs, et):
    dg, g = ba_hubs
    kw = dict(et_type=et, et_delta=0.25)
    q_o, cv_o, it_o = po.louvain_phase(g, **kw)
    q, cv, it = dg.louvain_phase(sequential=True, **kw)
    assert (q, it) == (q_o, it_o)
    assert np.array_equal(cv, cv_o)


def test_ba_hubs_run_exact(ctx, ba_hubs):
    dg, g = ba_hubs
    r_o = po.run(g, threshold_cycling=True)
    r = dg.run(sequential=True, threshold_cycling=True)
    assert r.call_q == r_o.call_q
    assert r.call_iters == r_o.call_iters
    assert r.level_nv == r_o.level_nv
